"""Host setup of a shot batch (no device call): every shot's located sources, source arrays, time functions
and t0 are those of a simulation that has only that shot's sources, and the oracle's setup of that simulation
(itself pinned against the reference's golden vectors) agrees with them."""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config
from specfem2d_amd.config import Source
from specfem2d_amd.driver import ShotBatch, Simulation


def shots_for(cfg):
    """Four shots: a single Ricker source, a rotated first-derivative-of-Gaussian source, a shot of two
    sources with different frequencies and time shifts (so its t0 differs from the others), and a Ricker source
    near a corner of the mesh."""
    w = cfg.xmax - cfg.xmin
    h = cfg.interfaces[-1] - cfg.interfaces[0]
    return [
        [Source(x=cfg.xmin + 0.31 * w, z=cfg.interfaces[-1] - 0.4 * h, f0=15.0, factor=1.0e9, anglesource=30.0)],
        [Source(x=cfg.xmin + 0.66 * w, z=cfg.interfaces[-1] - 0.55 * h, f0=12.0, factor=2.0e9, anglesource=120.0,
                time_function_type=2)],
        [Source(x=cfg.xmin + 0.45 * w, z=cfg.interfaces[-1] - 0.25 * h, f0=18.0, factor=1.0e9, tshift=0.01),
         Source(x=cfg.xmin + 0.52 * w, z=cfg.interfaces[-1] - 0.7 * h, f0=9.0, factor=-5.0e8, anglesource=75.0,
                tshift=0.004)],
        [Source(x=cfg.xmin + 0.9 * w, z=cfg.interfaces[-1] - 0.88 * h, f0=14.0, factor=1.5e9, anglesource=200.0)],
    ]


def test_each_shot_is_its_own_source_setup():
    cfg = config.small_test_mesh(nx=12, nz=9, nstep=60, gll=True, seed=5)
    shots = shots_for(cfg)
    batch = ShotBatch(cfg, shots, device=False)
    bi, bx, bg, ba, bs = batch.sources()
    t0 = batch.shot_t0()
    k0 = 0
    for k, shot in enumerate(shots):
        single = Simulation(cfg.replace(sources=shot), device=False)
        si, sx, sg, sa, st = single.sources()
        sl = slice(k0, k0 + len(shot))
        np.testing.assert_array_equal(bi[sl], si)
        np.testing.assert_array_equal(bx[sl], sx)
        np.testing.assert_array_equal(bg[sl], sg)
        np.testing.assert_array_equal(ba[sl], sa)
        np.testing.assert_array_equal(bs[sl], st)
        assert t0[k] == single.t0
        orc = OracleSim(cfg.replace(sources=shot), "sp")
        oi, _, _, osa, ostf = orc.sources()
        np.testing.assert_array_equal(si, oi)
        np.testing.assert_allclose(sa, osa, rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(st, ostf, rtol=1e-6, atol=1e-6 * np.abs(ostf).max())
        assert single.t0 == pytest.approx(orc.t0, rel=1e-15)  # the oracle forms 1.2 / f0 + tshift in another order
        k0 += len(shot)
    # the two-source shot keeps its own t0
    assert t0[2] != t0[0]
    batch.close()


def test_shot_batch_rejects_empty_shots():
    cfg = config.small_test_mesh(nstep=10)
    with pytest.raises(ValueError):
        ShotBatch(cfg, [], device=False)
    with pytest.raises(ValueError):
        ShotBatch(cfg, [[cfg.sources[0]], []], device=False)
