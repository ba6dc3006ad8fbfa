"""Meshes read from files (read_external_mesh = .true.): the reference's external-mesh example (5587 unstructured
4-node elements with topography, a fluid region over a solid one, 328 fluid-solid edges, a free surface on the fluid)
and structured meshes handed over through the external path.

* the files' reader: element, node and side counts of the example, side codes from node pairs;
* equivalence: a structured mesh given as control nodes and side lists gives the internal mesher's numbering, fluid-
  solid edges and free surface, and the oracle's traces agree to FP32 rounding (the same arithmetic, possibly another
  edge order);
* the host driver's set-up of the example against the oracle's: numbering, receiver and source location bit-exact
  (the located xi, gamma equal those of the reference's own log), fluid-solid edges and mass;
* the oracle against the reference's committed traces of the example over its first 4000 steps (1 s).

CPU only.
"""
import os

import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config, io
from specfem2d_amd.config import ExternalMesh, Source
from specfem2d_amd.driver import Simulation

EXAMPLE = "fluid_solid_external_mesh"


def _example(golden_dir, nstep):
    cfg, _ = io.config_from_par_file(os.path.join(golden_dir, "examples", EXAMPLE), nstep=nstep)
    return cfg


def test_reader_counts(golden_dir):
    m = _example(golden_dir, 10).mesh
    assert m.nspec == 5587 and m.coorg.shape == (5861, 2)
    assert len(m.free_elem) == 328 and len(m.abs_elem) == 217
    assert np.bincount(m.material).tolist() == [0, 2737, 2850]
    # every listed side's nodes are two consecutive control nodes of its element
    for e, c in zip(m.free_elem[:20], m.free_code[:20]):
        a, b = config.SIDE_NODES[int(c)]
        assert m.side_code(int(e), int(m.knods[e, a]), int(m.knods[e, b])) == c


def _as_external(cfg):
    """The internal mesher's structured mesh of ``cfg`` as control nodes, elements and side lists."""
    nx, nz = cfg.nx, cfg.nz
    xs = cfg.xmin + (cfg.xmax - cfg.xmin) * np.arange(nx + 1) / nx
    zs = []
    for il, nzl in enumerate(cfg.nz_layers):
        bot, top = cfg.interfaces[il], cfg.interfaces[il + 1]
        zs.append((1 - np.arange(nzl) / nzl) * bot + np.arange(nzl) / nzl * top)
    zs = np.concatenate(zs + [np.array([cfg.interfaces[-1]])])
    coorg = np.stack(np.meshgrid(xs, zs), axis=-1).reshape(-1, 2)
    ex, ez = np.tile(np.arange(nx), nz), np.repeat(np.arange(nz), nx)
    n0 = ez * (nx + 1) + ex
    knods = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1)
    free, absorb = [], []
    sides = {1: (cfg.absorbbottom, ez == 0), 2: (cfg.absorbright, ex == nx - 1), 3: (cfg.absorbtop, ez == nz - 1),
             4: (cfg.absorbleft, ex == 0)}
    order = {3: 0, 1: 1, 4: 2, 2: 3}  # the internal mesher's free-surface order: top, bottom, left, right
    for code in sorted(sides, key=order.get):
        ab, on = sides[code]
        if not ab:
            free += [(e, code) for e in np.flatnonzero(on)]
    for e in range(nx * nz):  # absorbing sides in element order, codes 1, 2, 3, 4
        for code in (1, 2, 3, 4):
            ab, on = sides[code]
            if ab and on[e]:
                absorb.append((e, code))
    a = lambda lst, k: np.array([t[k] for t in lst], np.int64)  # noqa: E731
    mesh = ExternalMesh(knods=knods, coorg=coorg, material=np.ones(nx * nz, np.int64), free_elem=a(free, 0),
                        free_code=a(free, 1), abs_elem=a(absorb, 0), abs_code=a(absorb, 1))
    return cfg.replace(mesh=mesh)


@pytest.fixture(scope="module")
def layered():
    rng = np.random.default_rng(11)
    pert = np.repeat(0.05 * rng.standard_normal(14 * 10), 25)
    src = [Source(x=520.0, z=-610.0, f0=14.0, factor=1e9, anglesource=15.0), Source(x=900.0, z=-150.0, f0=14.0,
                                                                                  factor=1e4)]
    rec = np.array([[300.0, -80.0], [1000.0, -420.0], [640.0, -880.0]])
    return config.layered_fluid_solid(14, 10, 3, 100.0, 600, 1.5e-3, pert=pert, sources=src, receivers=rec)


def test_structured_mesh_through_the_external_path(layered):
    cfg = layered
    ext = _as_external(cfg)
    a, b = OracleSim(cfg, "sp"), OracleSim(ext, "sp")
    np.testing.assert_array_equal(a.ibool(), b.ibool())
    assert a.ncoupling == b.ncoupling > 0 and a.nspec_acoustic == b.nspec_acoustic > 0
    assert sorted(map(tuple, a.coupling())) == sorted(map(tuple, b.coupling()))
    sa, sb = a.forward(), b.forward()
    assert np.linalg.norm(sa - sb) / np.linalg.norm(sa) < 1e-6
    # the host driver as well: the same numbering and fluid-solid edges as for the internal mesh
    d1, d2 = Simulation(cfg, device=False), Simulation(ext, device=False)
    np.testing.assert_array_equal(d1.ibool(), d2.ibool())
    assert sorted(map(tuple, d1.coupling())) == sorted(map(tuple, d2.coupling()))
    np.testing.assert_array_equal(d1.rmass_inverse(), d2.rmass_inverse())
    np.testing.assert_array_equal(d1.rmass_acoustic(), d2.rmass_acoustic())


def test_stacey_sides_through_the_external_path():
    cfg = config.small_test_mesh(nx=9, nz=7, nstep=300, gll=True, seed=4).replace(absorbtop=True)
    ext = _as_external(cfg)
    a, b = OracleSim(cfg, "sp"), OracleSim(ext, "sp")
    for x, y in zip(a.absorbing(), b.absorbing()):
        np.testing.assert_array_equal(x, y)
    sa, sb = a.forward(), b.forward()
    np.testing.assert_array_equal(sa, sb)
    d = Simulation(ext, device=False)
    isp, ij, nrm, jw, edge, ib = d.stacey()
    oisp, ocode, _, oib = b.absorbing()
    np.testing.assert_array_equal(isp, oisp)
    np.testing.assert_array_equal(edge, ocode)


@pytest.fixture(scope="module")
def example(golden_dir):
    cfg = _example(golden_dir, 50)
    return cfg, Simulation(cfg, device=False), OracleSim(cfg, "sp")


def test_example_setup_matches_oracle(example):
    cfg, sim, orc = example
    np.testing.assert_array_equal(sim.ibool(), orc.ibool())
    assert sim.nglob == orc.nglob == 90485
    assert sim.nspec_acoustic == orc.nspec_acoustic == 2737
    assert sim.ncoupling == orc.ncoupling == 328
    np.testing.assert_array_equal(sim.coupling(), orc.coupling())
    np.testing.assert_array_equal(sim.rmass_acoustic().astype(np.float64), orc.rmass_acoustic())
    np.testing.assert_allclose(sim.rmass_inverse(), orc.rmass_inverse(), rtol=2e-6)
    for name in ("xix", "xiz", "gammax", "gammaz", "jacobian"):
        b = orc.elem_array(name).astype(np.float32)
        np.testing.assert_allclose(sim.elem_array(name), b, rtol=4e-7, atol=1e-6 * np.abs(b).max(), err_msg=name)
    si, sx, sg, _, _ = sim.sources()
    oi, ox, og, _, _ = orc.sources()
    np.testing.assert_array_equal(si, oi)
    np.testing.assert_allclose(sx, ox, atol=1e-12)
    np.testing.assert_allclose(sg, og, atol=1e-12)
    ri, rx, rg, _, _ = sim.receivers()
    qi, qx, qg, _, _ = orc.receivers()
    np.testing.assert_array_equal(ri, qi)
    np.testing.assert_allclose(rx, qx, atol=1e-12)
    np.testing.assert_allclose(rg, qg, atol=1e-12)
    # the reference's own log of this example: the source and receivers 6 and 10 at these xi, gamma
    np.testing.assert_allclose([sx[0], sg[0]], [-0.21282643504158869, -0.50965125943626033], atol=1e-12)
    np.testing.assert_allclose([rx[5], rg[5]], [4.1873623501770506e-2, -0.65885025816479270], atol=1e-12)
    np.testing.assert_allclose([rx[9], rg[9]], [0.99122759245559688, 0.79606123353950664], atol=1e-12)


@pytest.mark.slow
def test_oracle_against_reference_traces(golden_dir):
    """The first 4000 steps: receivers 5 and 6 (the closest to the source) carry the direct wave.  The bar, 5e-5 of
    each trace's peak, covers this example's own FP32 spread: the FP32 oracle is 2.1e-5 of the peak away from the
    FP64 one at receiver 6, and the reference's trace was computed in FP32 on 4 MPI partitions."""
    cfg = _example(golden_dir, 4000)
    seis = OracleSim(cfg, "sp").forward()
    g = np.load(os.path.join(golden_dir, f"{EXAMPLE}_ref_seis.npz"))
    worst = 0.0
    for r in range(12):
        for c, ch in enumerate(("BXX", "BXZ")):
            full = g[f"AA.S{r + 1:04d}.{ch}"].astype(np.float64)
            worst = max(worst, float(np.abs(seis[r, c] - full[:4000]).max() / np.abs(full).max()))
    ref6 = g["AA.S0006.BXX"][:4000]
    assert np.abs(ref6).max() > 0.5 * np.abs(g["AA.S0006.BXX"]).max()  # the window holds the direct wave's peak
    assert worst < 5e-5
