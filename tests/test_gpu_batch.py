"""GPU parity of shot batches (all shots of one mesh and model advanced together, one launch per kernel).

* forward: every shot's seismograms and final fields equal those of a single simulation with only that
  shot's sources (rel L2 2e-6 on the traces: FP32 summation order), and the FP32 oracle's traces of that
  simulation to the parity tolerance 1e-5; shot-minor and shot-major batch layouts agree;
* gradient: the batch misfit is the sum of the single misfits (per shot too), the batch gradient the sum of
  the single gradients (rel L2 1e-5), and the sum of the FP64 oracle's gradients to the tolerance of the
  single-shot gradient test; checkpoint length does not change it; the torch.autograd misfit runs on a batch.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config
from tests.test_batch_setup import shots_for

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def case():
    cfg = config.small_test_mesh(nx=12, nz=9, nstep=300, gll=True, seed=5).replace(
        absorbtop=True, absorbbottom=True, absorbleft=True, absorbright=True)
    return cfg, shots_for(cfg)


@pytest.fixture(scope="module")
def singles(case):
    from specfem2d_amd.driver import Simulation
    cfg, shots = case
    out = []
    for shot in shots:
        sim = Simulation(cfg.replace(sources=shot))
        seis, _ = sim.forward()
        out.append((seis, sim.fields()))
        sim.close()
    return out


def test_batch_forward_matches_single_runs(case, singles):
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    b = ShotBatch(cfg, shots)
    seis, ms = b.forward(timing=True)
    assert seis.shape == (len(shots), cfg.nrec, 2, cfg.nstep)
    for k, (ref, fields) in enumerate(singles):
        err = _rel(seis[k], ref)
        print(f"shot {k}: batch vs single seismograms rel L2 {err:.2e}")
        assert err < 2e-6, (k, err)
        for name, x, y in zip(("displ", "veloc", "accel"), b.fields(k), fields):
            e = _rel(x, y)
            print(f"  {name} rel L2 {e:.2e}")
            assert e < {"displ": 5e-6, "veloc": 2e-5, "accel": 5e-5}[name], (k, name, e)
    print(f"batch of {len(shots)} shots, {cfg.nstep} steps: {ms:.1f} ms")
    b.close()


def test_shot_minor_and_shot_major_layouts_agree(case, observed, monkeypatch):
    """The lane-per-shot kernels on shot-minor fields (SF2D_SHOT_MINOR) and the patch kernels on shot-major
    fields (the default) give the same seismograms and the same summed gradient to FP32 rounding."""
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    out = {}
    monkeypatch.delenv("SF2D_SHOT_MAJOR", raising=False)
    for major in (False, True):
        if major:
            monkeypatch.delenv("SF2D_SHOT_MINOR", raising=False)
        else:
            monkeypatch.setenv("SF2D_SHOT_MINOR", "1")
        b = ShotBatch(cfg, shots)
        assert b.shot_minor() == (not major)
        seis = b.forward()
        b.gradient_forward(observed, chunk=64)
        out[major] = (seis, b.gradient_backward(), b.fields(2))
        b.close()
    err = _rel(out[False][0], out[True][0])
    print(f"shot-minor vs shot-major: seismograms rel L2 {err:.2e}")
    assert err < 2e-6
    for name in ("rho", "vp", "vs"):
        e = _rel(out[False][1][name], out[True][1][name])
        print(f"  gradient {name}: {e:.2e}")
        assert e < 1e-5
    for x, y, tol in zip(out[False][2][:2], out[True][2][:2], (5e-6, 2e-5)):
        assert _rel(x, y) < tol


def test_batch_forward_matches_the_oracle(case):
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    b = ShotBatch(cfg, shots)
    seis = b.forward()
    b.close()
    for k, shot in enumerate(shots):
        ref = OracleSim(cfg.replace(sources=shot), "sp").forward()
        err = _rel(seis[k], ref)
        print(f"shot {k}: batch vs FP32 oracle rel L2 {err:.2e}")
        assert err < 1e-5, (k, err)


@pytest.fixture(scope="module")
def observed(case):
    cfg, shots = case
    rng = np.random.default_rng(1)
    vs0 = np.asarray(cfg.vs, float)
    true = cfg.replace(vs=vs0 * (1 + 0.02 * rng.standard_normal(vs0.size)))
    return np.stack([OracleSim(true.replace(sources=shot), "dp").forward() for shot in shots])


@pytest.fixture(scope="module")
def batch_grad(case, observed):
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    b = ShotBatch(cfg, shots)
    chi, seis = b.gradient_forward(observed, chunk=64)
    per_shot = b.shot_misfits()
    g = b.gradient_backward()
    b.close()
    return chi, per_shot, seis, g


def test_batch_gradient_is_the_sum_of_single_gradients(case, observed, batch_grad):
    from specfem2d_amd.driver import Simulation
    cfg, shots = case
    chi, per_shot, seis, g = batch_grad
    total = {k: 0.0 for k in ("rho", "vp", "vs")}
    chis = []
    for k, shot in enumerate(shots):
        sim = Simulation(cfg.replace(sources=shot))
        c, s = sim.gradient_forward(observed[k], chunk=64)
        gk = sim.gradient_backward()
        sim.close()
        chis.append(c)
        assert _rel(seis[k], s) < 2e-6
        for name in total:
            total[name] = total[name] + gk[name]
    print(f"chi batch {chi:.6e}, sum of singles {sum(chis):.6e}, per shot {per_shot} vs {chis}")
    assert abs(chi - sum(chis)) < 1e-6 * sum(chis)
    # per shot against the total: a shot whose synthetics nearly equal its observations has a misfit made of
    # rounding-sensitive residuals
    np.testing.assert_allclose(per_shot, chis, rtol=0, atol=1e-6 * sum(chis))
    for name in total:
        err = _rel(g[name], total[name])
        print(f"batch gradient {name}: rel L2 vs sum of single gradients {err:.2e}")
        assert err < 1e-5, (name, err)


def test_batch_gradient_against_oracle(case, observed, batch_grad):
    cfg, shots = case
    chi, _, _, g = batch_grad
    gd = {k: 0.0 for k in ("rho", "vp", "vs")}
    gs = {k: 0.0 for k in ("rho", "vp", "vs")}
    cd = 0.0
    for k, shot in enumerate(shots):
        c, d, _ = OracleSim(cfg.replace(sources=shot), "dp").gradient(observed[k])
        _, s, _ = OracleSim(cfg.replace(sources=shot), "sp").gradient(observed[k])
        cd += c
        for name in gd:
            gd[name] = gd[name] + d[name]
            gs[name] = gs[name] + s[name]
    assert abs(chi - cd) < 1e-4 * cd
    for name in gd:
        tol = max(1e-5, 3 * _rel(gs[name], gd[name]))
        err = _rel(g[name], gd[name])
        print(f"batch gradient {name}: rel L2 vs FP64 oracle {err:.2e} (tol {tol:.1e})")
        assert err < tol, (name, err, tol)


def test_batch_checkpoint_length_does_not_matter(case, observed, batch_grad):
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    _, _, _, g64 = batch_grad
    b = ShotBatch(cfg, shots)
    for chunk in (1, 37, cfg.nstep):
        b.gradient_forward(observed, chunk=chunk)
        g = b.gradient_backward()
        for name in g64:
            err = _rel(g[name], g64[name])
            assert err < 1e-5, (chunk, name, err)
    b.close()


def test_autograd_misfit_on_a_batch(case, observed, batch_grad):
    import torch
    from specfem2d_amd.autograd import waveform_misfit
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    chi_ref, _, _, g = batch_grad
    b = ShotBatch(cfg, shots)
    dev = torch.device("cuda:0")
    params = [torch.tensor(np.asarray(getattr(cfg, k), np.float64).reshape(cfg.nspec, 25), device=dev,
                           dtype=torch.float32, requires_grad=True) for k in ("rho", "vp", "vs")]
    chi = waveform_misfit(*params, b, observed, chunk=50)
    chi.backward()
    assert abs(chi.item() - chi_ref) < 1e-4 * chi_ref
    for t, k in zip(params, ("rho", "vp", "vs")):
        err = _rel(t.grad.double().cpu().numpy(), g[k])
        assert err < 1e-4, (k, err)
    b.close()


def test_autotuner_picks_a_candidate_and_keeps_the_results(case, monkeypatch):
    """The autotuner times every (elements per chain patch, shots per wave) candidate on the workload and
    returns the fastest; any candidate gives the same seismograms (FP32 rounding)."""
    from specfem2d_amd import autotune
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    for k in ("SF2D_CHAIN_ELEMS", "SF2D_SHOTS_PER_WAVE"):
        monkeypatch.delenv(k, raising=False)
    choice = autotune.tune(cfg, shots, steps=10, chain_elems=(4, 12), shots_per_wave=(1, 4))
    print("autotuner:", choice)
    assert {(r["chain_elems"], r["shots_per_wave"]) for r in choice["table"]} == {(4, 1), (4, 4), (12, 1), (12, 4)}
    assert choice["ms_per_step"] == min(r["ms_per_step"] for r in choice["table"])
    out = []
    for n in (4, 12):
        b = ShotBatch(cfg, shots, **autotune.settings({"chain_elems": n, "shots_per_wave": 1 if n == 4 else 4}))
        assert b.chain_elems() == n
        out.append(b.forward())
        b.close()
    assert _rel(out[0], out[1]) < 2e-6


def test_graph_replayed_steps_match_launched_steps(case, observed, monkeypatch):
    """With SF2D_BATCH_GRAPH, every batch step after the first of a run is one replay of a captured HIP graph
    (the step index read from device memory); it gives the same seismograms as launching every kernel, and so
    do repeated runs that replay the same graph with eager resets and checkpoint copies in between."""
    from specfem2d_amd.driver import ShotBatch
    cfg, shots = case
    monkeypatch.delenv("SF2D_BATCH_GRAPH", raising=False)
    b = ShotBatch(cfg, shots)
    ref = b.forward()
    b.gradient_forward(observed, chunk=37)
    g_ref = b.gradient_backward()
    monkeypatch.setenv("SF2D_BATCH_GRAPH", "1")
    runs = [b.forward() for _ in range(3)]
    b.gradient_forward(observed, chunk=37)  # replays interleaved with checkpoint copies
    g = b.gradient_backward()
    b.close()
    for k, r in enumerate(runs):
        err = _rel(r, ref)
        print(f"graph run {k}: seismograms rel L2 vs launched steps {err:.2e}")
        assert err < 1e-6
    for name in g_ref:
        assert _rel(g[name], g_ref[name]) < 1e-6, name
