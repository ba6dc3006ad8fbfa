"""Pins the oracle's exact discrete gradient (orc_gradient) of the L2 waveform misfit
chi = 1/2 dt sum_n sum_r |s_n - o_n|^2 with respect to the MODEL=gll parameters (rho, vp, vs).

* central finite differences of the FP64 forward misfit along random directions: the relative error falls
  as eps^2 (1e-4 -> 1e-5), and is below 1e-5 at eps = 1e-5 for rho, vp and vs;
* every term of the gradient is visible to that check: dropping the mass-matrix term, the Stacey terms or
  the bulk-modulus term moves the directional derivative by far more than the finite-difference error;
* the reference's SIMULATION_TYPE=3 kernels (oracle restatement of compute_kernels.f90 +
  save_adjoint_kernels.f90, FP64, backward reconstruction, waveform adjoint source s - o) are the same
  gradient: K_kappa = kappa g_kappa / (w_i w_j J) and K_rho = rho g_rho(mass) / (w_i w_j J) to 1e-12, and
  K_mu = mu g_mu / (w J) + mu / (3 kappa) K_kappa -- the reference forms the shear kernel with the 3D
  deviator (D:D - div div / 3) while the 2D solver's bulk modulus is kappa = lambda + mu.  This pins the
  kernel scaling constants (-rho dt, -2 mu dt, -kappa dt) of save_adjoint_kernels.f90:100-140.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config


def _case(nx=8, nz=6, nstep=300):
    cfg = config.small_test_mesh(nx=nx, nz=nz, nstep=nstep, gll=True, seed=5).replace(
        absorbtop=True, absorbbottom=True, absorbleft=True, absorbright=True)
    rng = np.random.default_rng(1)
    vs0 = np.asarray(cfg.vs, float)
    obs = OracleSim(cfg.replace(vs=vs0 * (1 + 0.02 * rng.standard_normal(vs0.size))), "dp").forward()
    return cfg, obs


def _misfit(cfg, obs):
    s = OracleSim(cfg, "dp").forward()
    return 0.5 * cfg.dt * float(np.sum((s - obs) ** 2))


@pytest.fixture(scope="module")
def case():
    cfg, obs = _case()
    chi, grad, parts = OracleSim(cfg, "dp").gradient(obs)
    return cfg, obs, chi, grad, parts


def _fd(cfg, obs, name, dm, eps):
    base = np.asarray(getattr(cfg, name), float)
    return (_misfit(cfg.replace(**{name: base * (1 + eps * dm)}), obs)
            - _misfit(cfg.replace(**{name: base * (1 - eps * dm)}), obs)) / (2 * eps)


def test_misfit_value(case):
    cfg, obs, chi, _, _ = case
    assert chi == pytest.approx(_misfit(cfg, obs), rel=1e-12)


@pytest.mark.parametrize("name", ["rho", "vp", "vs"])
def test_gradient_matches_central_differences(case, name):
    cfg, obs, _, grad, _ = case
    rng = np.random.default_rng({"rho": 11, "vp": 12, "vs": 13}[name])
    dm = rng.standard_normal(cfg.nspec * 25)
    base = np.asarray(getattr(cfg, name), float)
    an = float(np.sum(grad[name].reshape(-1) * base * dm))
    e4 = abs(_fd(cfg, obs, name, dm, 1e-4) - an) / abs(an)
    e5 = abs(_fd(cfg, obs, name, dm, 1e-5) - an) / abs(an)
    assert e5 < 1e-5, (name, e4, e5)
    assert e5 < 0.05 * e4 or e5 < 1e-8, (name, e4, e5)  # second-order convergence: the gradient is exact


def test_every_term_is_visible_to_the_check(case):
    cfg, obs, _, grad, parts = case
    rng = np.random.default_rng(21)
    n = cfg.nspec * 25
    rho, vp, vs = (np.asarray(getattr(cfg, k), float).reshape(-1) for k in ("rho", "vp", "vs"))
    dm = rng.standard_normal(n)
    fd_rho = _fd(cfg, obs, "rho", dm, 1e-5)
    fd_vp = _fd(cfg, obs, "vp", dm, 1e-5)
    g = {k: v.reshape(-1) for k, v in parts.items()}
    full_rho = np.sum(grad["rho"].reshape(-1) * rho * dm)
    tol = 1e-5 * abs(fd_rho)
    assert abs(full_rho - fd_rho) < tol
    # a dropped mass term, Stacey term or bulk term would each fail the check by orders of magnitude
    assert abs(np.sum((grad["rho"].reshape(-1) - g["rho_mass"]) * rho * dm) - fd_rho) > 100 * tol
    assert abs(np.sum((grad["rho"].reshape(-1) - g["rho_vp"] * vp - g["rho_vs"] * vs) * rho * dm) - fd_rho) > 100 * tol
    full_vp = np.sum(grad["vp"].reshape(-1) * vp * dm)
    assert abs(full_vp - fd_vp) < 1e-5 * abs(fd_vp)
    assert abs(np.sum((grad["vp"].reshape(-1) - g["rho_vp"] * rho) * vp * dm) - fd_vp) > 1e-3 * abs(fd_vp)


def test_reference_adjoint_kernels_are_the_exact_gradient():
    cfg = config.small_test_mesh(nx=14, nz=10, nstep=500, gll=True, seed=5)
    rng = np.random.default_rng(1)
    vs0 = np.asarray(cfg.vs, float)
    obs = OracleSim(cfg.replace(vs=vs0 * (1 + 0.02 * rng.standard_normal(vs0.size))), "dp").forward()
    o = OracleSim(cfg, "dp")
    _, _, parts = o.gradient(obs)
    fwd = OracleSim(cfg, "dp")
    seis, store = fwd.forward(save_stacey=True)
    _, ker = OracleSim(cfg, "dp").adjoint(seis - obs, store, fwd.fields())
    xi = config.XIGLL
    leg = np.polynomial.legendre.Legendre.basis(4)
    w = 2.0 / (20.0 * leg(xi) ** 2)
    W = np.outer(w, w).reshape(-1)[None, :] * o.elem_array("jacobian")
    kap, mu, rho = (o.elem_array(k) for k in ("kappa", "mu", "rho"))
    k_kappa = parts["kappa"] * kap / W
    k_rho = parts["rho_mass"] * rho / W
    k_mu = parts["mu"] * mu / W + mu / (3 * kap) * k_kappa

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)

    assert rel(k_kappa, ker["kappa"]) < 1e-12
    assert rel(k_rho, ker["rho"]) < 1e-12
    assert rel(k_mu, ker["mu"]) < 1e-12
    # without the deviator conversion the shear kernels differ at the percent level
    assert rel(parts["mu"] * mu / W, ker["mu"]) > 1e-3
