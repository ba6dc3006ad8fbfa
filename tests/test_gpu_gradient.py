"""GPU parity of the exact discrete gradient (checkpointed forward recomputation + reverse-mode kernels) and of
the differentiable misfit behind torch.autograd.

* against the oracle's gradient (orc_gradient, itself pinned by FP64 central finite differences): relative
  L2 within max(1e-5, 3 x the FP32 oracle's own distance from the FP64 one), for rho, vp and vs, on a
  heterogeneous mesh with Stacey on all four sides, several receivers and checkpoint segments of 64 steps
  (the last one partial);
* the checkpoint length does not change the result (segments of 1, 37 and NSTEP steps);
* central finite differences of the GPU misfit itself (FP32 forward, eps = 1e-3 relative) along random
  directions agree with the GPU gradient to 3e-3;
* the torch.autograd.Function returns the same gradient through .backward(), on tensors on cuda:0;
* the GPU's SIMULATION_TYPE=3 kernels (reference entry points, backward reconstruction) are this gradient:
  K_kappa = kappa g_kappa / (w J) to the FP32 level of the reconstruction.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _case(nstep=300):
    cfg = config.small_test_mesh(nx=12, nz=9, nstep=nstep, gll=True, seed=5).replace(
        absorbtop=True, absorbbottom=True, absorbleft=True, absorbright=True)
    rng = np.random.default_rng(1)
    vs0 = np.asarray(cfg.vs, float)
    obs = OracleSim(cfg.replace(vs=vs0 * (1 + 0.02 * rng.standard_normal(vs0.size))), "dp").forward()
    return cfg, obs


@pytest.fixture(scope="module")
def case():
    return _case()


@pytest.fixture(scope="module")
def gpu_grad(case):
    from specfem2d_amd.driver import Simulation
    cfg, obs = case
    sim = Simulation(cfg)
    chi, seis = sim.gradient_forward(obs, chunk=64)
    g, parts = sim.gradient_backward(parts=True)
    sim.close()
    return chi, g, parts


def test_gradient_against_oracle(case, gpu_grad):
    cfg, obs = case
    chi, g, _ = gpu_grad
    cd, gd, _ = OracleSim(cfg, "dp").gradient(obs)
    cs, gs, _ = OracleSim(cfg, "sp").gradient(obs)
    print(f"chi GPU {chi:.6e}  FP64 {cd:.6e}  FP32 oracle {cs:.6e}")
    assert abs(chi - cd) < 1e-4 * cd
    for k in ("rho", "vp", "vs"):
        tol = max(1e-5, 3 * _rel(gs[k], gd[k]))
        err = _rel(g[k], gd[k])
        print(f"grad {k}: GPU vs FP64 oracle rel L2 {err:.3e} (FP32 oracle {_rel(gs[k], gd[k]):.3e}, tol {tol:.1e})")
        assert err < tol, (k, err, tol)


def test_checkpoint_length_does_not_matter(case, gpu_grad):
    from specfem2d_amd.driver import Simulation
    cfg, obs = case
    _, g64, _ = gpu_grad
    sim = Simulation(cfg)
    for chunk in (1, 37, cfg.nstep):
        sim.gradient_forward(obs, chunk=chunk)
        g = sim.gradient_backward()
        for k in ("rho", "vp", "vs"):
            err = _rel(g[k], g64[k])
            print(f"chunk {chunk}: {k} rel L2 vs chunk 64: {err:.2e}")
            assert err < 1e-5, (chunk, k, err)
    sim.close()


def test_gpu_finite_differences(case, gpu_grad):
    from specfem2d_amd.driver import Simulation
    cfg, obs = case
    _, g, _ = gpu_grad
    sim = Simulation(cfg)
    base = {k: np.asarray(getattr(cfg, k), float).reshape(-1) for k in ("rho", "vp", "vs")}
    rng = np.random.default_rng(7)
    eps = 1e-3
    for k in ("rho", "vp", "vs"):
        dm = rng.standard_normal(base[k].size)
        chis = []
        for sgn in (1, -1):
            m = dict(base)
            m[k] = base[k] * (1 + sgn * eps * dm)
            sim.set_model(m["rho"], m["vp"], m["vs"])
            chis.append(sim.gradient_forward(obs, chunk=cfg.nstep)[0])
        fd = (chis[0] - chis[1]) / (2 * eps)
        an = float(np.sum(g[k].reshape(-1) * base[k] * dm))
        print(f"{k}: FD {fd:.6e} gradient {an:.6e} rel {abs(fd - an) / abs(an):.2e}")
        assert abs(fd - an) < 3e-3 * abs(an), (k, fd, an)
    sim.close()


def test_autograd_function(case, gpu_grad):
    import torch
    from specfem2d_amd.autograd import waveform_misfit
    from specfem2d_amd.driver import Simulation
    cfg, obs = case
    chi_ref, g, _ = gpu_grad
    sim = Simulation(cfg)
    dev = torch.device("cuda:0")
    params = [torch.tensor(np.asarray(getattr(cfg, k), np.float64).reshape(cfg.nspec, 25), device=dev,
                           dtype=torch.float32, requires_grad=True) for k in ("rho", "vp", "vs")]
    chi = waveform_misfit(*params, sim, obs, chunk=50)
    chi.backward()
    assert abs(chi.item() - chi_ref) < 1e-4 * chi_ref
    for t, k in zip(params, ("rho", "vp", "vs")):
        assert t.grad is not None and t.grad.device == dev
        err = _rel(t.grad.double().cpu().numpy(), g[k])
        print(f"autograd {k}: rel L2 vs direct call {err:.2e}")
        assert err < 1e-4, (k, err)
    sim.close()


def test_reference_adjoint_kernels_match_the_gradient(case, gpu_grad):
    from specfem2d_amd.driver import Simulation
    cfg, obs = case
    _, _, parts = gpu_grad
    fwd = Simulation(cfg, simulation_type=1, save_forward=True)
    seis, _, store = fwd.forward(save_stacey=True)
    last = fwd.fields()
    fwd.close()
    adj_sim = Simulation(cfg, simulation_type=3)
    _, ker = adj_sim.adjoint(seis - obs, store, last)
    w1 = 2.0 / (20.0 * np.polynomial.legendre.Legendre.basis(4)(config.XIGLL) ** 2)
    W = np.outer(w1, w1).reshape(-1)[None, :] * adj_sim.elem_array("jacobian")
    kap, mu, rho = (adj_sim.elem_array(k) for k in ("kappa", "mu", "rho"))
    adj_sim.close()
    k_kappa = parts["kappa"] * kap / W
    k_rho = parts["rho_mass"] * rho / W
    k_mu = parts["mu"] * mu / W + mu / (3 * kap) * k_kappa
    for name, ours in (("kappa", k_kappa), ("rho", k_rho), ("mu", k_mu)):
        err = _rel(ker[name], ours)
        print(f"SIMULATION_TYPE=3 K_{name} vs exact gradient: rel L2 {err:.2e}")
        assert err < 1e-3, (name, err)
