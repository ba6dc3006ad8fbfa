"""GPU checks at the full sizes of the large configurations (the other GPU tests use small meshes):

* the 2000 x 1000 fluid-solid mesh with its CPML frame (water layer over a random-field solid): traces in the causal
  window against the FP32 oracle run on a cropped piece of the same mesh and model around the source and receivers
  -- before time T no wave that left the source can have reached the crop's edges and come back, so the cropped
  and the full mesh give the same traces up to FP32 rounding;
* the single-shot 800 x 400-element, 2000-step adjoint-gradient workload: central finite differences of the GPU
  misfit (FP64 misfit sums over FP32 traces) along 8 seeded smooth directions at eps 1e-3, 1e-4, 1e-5 against the
  exact discrete gradient, and the reference's SIMULATION_TYPE=3 kernels (backward reconstruction, reference time
  pairing) against the same gradient;
* the batched 600 x 300-element, 1500-step FWI workload: the gradient of a batch of 4 shots equals the sum of the
  4 single-shot gradients, and central finite differences of the batched misfit along 2 smooth directions agree
  with it.

Every measured error is printed (run pytest with -s to see them).
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config
from specfem2d_amd.config import Source

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _smooth_direction(X, Z, seed, nmodes=4, lmin=1500.0, lmax=5000.0):
    """A seeded smooth field of unit amplitude at the GLL points: a sum of plane waves of wavelengths lmin..lmax."""
    rng = np.random.default_rng(seed)
    f = np.zeros(X.size)
    for _ in range(nmodes):
        lam = rng.uniform(lmin, lmax)
        th = rng.uniform(0.0, np.pi)
        k = 2 * np.pi / lam
        f += rng.standard_normal() * np.cos(k * (np.cos(th) * X + np.sin(th) * Z) + rng.uniform(0, 2 * np.pi))
    return f / np.abs(f).max()


# ---------------------------------------------------------------------------------------------------------------
# 2000 x 1000 fluid-solid mesh with CPML: causal window against the oracle on a cropped piece
def test_c2_fluid_solid_cpml_causal_window():
    from specfem2d_amd.driver import Simulation
    nx, nz, h = 2000, 1000, 10.0
    nstep, dt = 1200, 2.0e-4          # T = 0.24 s
    vmax = 3000.0 * 1.10              # the random field stays within +-10 % of vp 3000
    src = [Source(x=10000.0, z=-1150.0, f0=15.0, factor=1.0e10, anglesource=20.0)]
    rec = np.array([[10000.0, -900.0], [10230.0, -960.0], [9790.0, -1080.0], [10150.0, -1320.0],
                    [9850.0, -1500.0], [10400.0, -1200.0]])
    print(f"case: 2000x1000 fluid-solid mesh with CPML 20, {nstep} steps, causal window {nstep * dt:.2f} s",
          flush=True)
    full = config.c2_mesh_coupled(nstep=nstep, nelem_pml=20).replace(sources=src, receivers=rec)
    sim = Simulation(full)
    seis, t = sim.forward()
    stats = sim.pml_stats()
    sim.close()
    # the crop: 240 x 250 elements around the source (x 8800..11200 m, z -2500..0 m), Stacey on its three
    # buried sides; its model is the full mesh's, element by element
    cx0, cx1, cz0 = 880, 1120, nz - 250
    margin = min(src[0].x - cx0 * h, cx1 * h - src[0].x, src[0].z + nz * h - cz0 * h)
    assert margin > vmax * nstep * dt + 100.0  # no wave reaches a crop edge (let alone comes back) before T
    ncx, ncz = cx1 - cx0, nz - cz0
    sel = (np.arange(cz0, nz)[:, None] * nx + np.arange(cx0, cx1)[None, :]).reshape(-1)
    pick = lambda a: np.asarray(a, np.float64).reshape(nx * nz, 25)[sel].reshape(-1)  # noqa: E731
    crop = config.SimulationConfig(
        name="c2_crop", nx=ncx, nz=ncz, xmin=cx0 * h, xmax=cx1 * h, interfaces=[-ncz * h, 0.0], nz_layers=[ncz],
        nstep=nstep, dt=dt, material_mode=1, stacey=True, absorbbottom=True, absorbright=True, absorbleft=True,
        absorbtop=False, sources=src, receivers=rec, rho=pick(full.rho), vp=pick(full.vp), vs=pick(full.vs))
    ref = OracleSim(crop, "sp").forward()
    exact = OracleSim(crop, "dp").forward()
    gap = _rel(ref, exact)  # the FP32 oracle's own distance from the FP64 one (the water layer rings)
    r, r64 = _rel(seis, ref), _rel(seis, exact)
    per = [_rel(seis[k], ref[k]) for k in range(len(rec))]
    per64 = [_rel(ref[k], exact[k]) for k in range(len(rec))]
    print(f"C2 causal window: rel-L2 GPU (full mesh, {stats[0]} + {stats[1]} CPML elements) vs FP32 oracle (crop "
          f"{ncx}x{ncz}) {r:.3e}, vs FP64 oracle {r64:.3e} (FP32 oracle vs FP64: {gap:.3e}); per receiver GPU vs "
          f"FP32 {', '.join(f'{p:.1e}' for p in per)}; FP32 vs FP64 oracle {', '.join(f'{p:.1e}' for p in per64)}; "
          f"GPU loop {t['loop_ms']:.0f} ms")
    # the GPU is an FP32 path of its own: within twice the FP32 oracle's own gap of either oracle
    bar = max(1e-5, 2.0 * gap)
    assert np.abs(ref).max() > 0 and r < bar and r64 < bar


# ---------------------------------------------------------------------------------------------------------------
# 800 x 400 adjoint-gradient workload
@pytest.fixture(scope="module")
def c3():
    from specfem2d_amd.driver import Simulation
    cfg = config.adjoint_gradient_800x400(anomaly=False)
    true_cfg = config.adjoint_gradient_800x400(anomaly=True)
    print("case: 800x400 adjoint gradient, 2000 steps", flush=True)
    sim = Simulation(cfg)
    sim.set_model(true_cfg.rho, true_cfg.vp, true_cfg.vs)
    sim.reset()
    obs, _ = sim.forward()
    sim.set_model(cfg.rho, cfg.vp, cfg.vs)
    chi, seis = sim.gradient_forward(obs, chunk=100)
    g, parts = sim.gradient_backward(parts=True)
    yield sim, cfg, obs, chi, seis, g, parts
    sim.close()


def test_c3_gradient_finite_differences(c3):
    sim, cfg, obs, chi0, _, g, _ = c3
    X, Z = (a.reshape(-1) for a in cfg.gll_point_coords())
    base = {k: np.asarray(getattr(cfg, k), np.float64).reshape(-1) for k in ("rho", "vp", "vs")}
    keys = ("rho", "vp", "vs")
    worst = 0.0
    print("C3 finite differences (rel error |FD - <g, dm>| / |<g, dm>|):")
    print("  dir    <g, dm>        eps=1e-3   eps=1e-4   eps=1e-5")
    for d in range(8):
        dm = {k: _smooth_direction(X, Z, seed=1000 + 3 * d + i) for i, k in enumerate(keys)}
        an = sum(float(np.dot(g[k].reshape(-1), base[k] * dm[k])) for k in keys)
        errs = []
        for eps in (1e-3, 1e-4, 1e-5):
            chis = []
            for sgn in (1.0, -1.0):
                m = {k: base[k] * (1.0 + sgn * eps * dm[k]) for k in keys}
                sim.set_model(m["rho"], m["vp"], m["vs"])
                sim.reset()
                s, _ = sim.forward()
                chis.append(0.5 * cfg.dt * float(np.sum((s - obs) ** 2)))
            fd = (chis[0] - chis[1]) / (2 * eps)
            errs.append(abs(fd - an) / abs(an))
        print(f"  {d}   {an: .6e}   {errs[0]:.2e}   {errs[1]:.2e}   {errs[2]:.2e}")
        worst = max(worst, min(errs))
    sim.set_model(base["rho"], base["vp"], base["vs"])
    print(f"C3 finite differences: worst over the 8 directions of the best eps: {worst:.2e}")
    assert worst < 1e-3


def test_c3_reference_adjoint_kernels(c3):
    from specfem2d_amd.driver import Simulation
    sim, cfg, obs, _, _, _, parts = c3
    fwd = Simulation(cfg, simulation_type=1, save_forward=True)
    seis, _, store = fwd.forward(save_stacey=True)
    last = fwd.fields()
    fwd.close()
    adj = Simulation(cfg, simulation_type=3)
    _, ker = adj.adjoint(seis - obs, store, last)
    w1 = 2.0 / (20.0 * np.polynomial.legendre.Legendre.basis(4)(config.XIGLL) ** 2)
    W = np.outer(w1, w1).reshape(-1)[None, :] * adj.elem_array("jacobian")
    kap, mu, rho = (adj.elem_array(k) for k in ("kappa", "mu", "rho"))
    adj.close()
    k_kappa = parts["kappa"] * kap / W
    k_rho = parts["rho_mass"] * rho / W
    k_mu = parts["mu"] * mu / W + mu / (3 * kap) * k_kappa
    errs = {}
    for name, ours in (("kappa", k_kappa), ("rho", k_rho), ("mu", k_mu)):
        errs[name] = _rel(ker[name], ours)
    print("C3 SIMULATION_TYPE=3 kernels (backward reconstruction) vs the exact discrete gradient: "
          + ", ".join(f"K_{k} rel-L2 {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-4


# ---------------------------------------------------------------------------------------------------------------
# 600 x 300 batched FWI workload
@pytest.fixture(scope="module")
def c4():
    from specfem2d_amd.driver import ShotBatch
    cfg = config.batched_fwi_600x300()
    true_cfg = config.batched_fwi_600x300(true_model=True)
    shots = config.surface_shots(cfg, 64)[8::16]  # 4 of the 64 shots, across the surface
    print("case: 600x300 batched FWI, 4 shots, 1500 steps", flush=True)
    b = ShotBatch(true_cfg, shots)
    obs = b.forward()
    b.close()
    b = ShotBatch(cfg, shots)
    chi, _ = b.gradient_forward(obs, chunk=39)
    g = b.gradient_backward()
    yield b, cfg, shots, obs, chi, g
    b.close()


def test_c4_batch_gradient_is_sum_of_single_shots(c4):
    from specfem2d_amd.driver import Simulation
    b, cfg, shots, obs, chi, g = c4
    total = {k: np.zeros_like(g[k]) for k in g}
    chis = 0.0
    for k, shot in enumerate(shots):
        sim = Simulation(cfg.replace(sources=shot))
        c, _ = sim.gradient_forward(obs[k], chunk=100)
        gk = sim.gradient_backward()
        sim.close()
        chis += c
        for p in total:
            total[p] += gk[p]
    errs = {p: _rel(g[p], total[p]) for p in g}
    print(f"C4 batch of 4 vs sum of single shots: misfit {chi:.9e} vs {chis:.9e}; gradient rel-L2 "
          + ", ".join(f"{p} {e:.2e}" for p, e in errs.items()))
    assert abs(chi - chis) < 1e-6 * chis
    assert max(errs.values()) < 1e-5


def test_c4_batched_misfit_finite_differences(c4):
    b, cfg, shots, obs, chi0, g = c4
    X, Z = (a.reshape(-1) for a in cfg.gll_point_coords())
    keys = ("rho", "vp", "vs")
    base = {k: np.asarray(getattr(cfg, k), np.float64).reshape(-1) for k in keys}
    worst = 0.0
    for d in range(2):
        dm = {k: _smooth_direction(X, Z, seed=2000 + 3 * d + i, lmin=1000.0, lmax=4000.0) for i, k in enumerate(keys)}
        an = sum(float(np.dot(g[k].reshape(-1), base[k] * dm[k])) for k in keys)
        errs = []
        for eps in (1e-3, 1e-4):
            chis = []
            for sgn in (1.0, -1.0):
                m = {k: base[k] * (1.0 + sgn * eps * dm[k]) for k in keys}
                b.set_model(m["rho"], m["vp"], m["vs"])
                s = b.forward()
                chis.append(0.5 * cfg.dt * float(np.sum((s - obs) ** 2)))
            fd = (chis[0] - chis[1]) / (2 * eps)
            errs.append(abs(fd - an) / abs(an))
        print(f"C4 batched misfit FD, direction {d}: <g, dm> {an:.6e}; rel error eps=1e-3 {errs[0]:.2e}, "
              f"eps=1e-4 {errs[1]:.2e}")
        worst = max(worst, min(errs))
    b.set_model(base["rho"], base["vp"], base["vs"])
    assert worst < 1e-3
