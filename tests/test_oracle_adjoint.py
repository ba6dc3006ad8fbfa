"""Pins the oracle's adjoint path (SIMULATION_TYPE=3, backward reconstruction, kernels).

* adjoint wavefield: bit-exact against the reference's committed adjoint-run traces
  ``EXAMPLES/Tromp2005_kernel/REF_KERNEL/AA.S0001.BX{X,Z}.semd``, with the adjoint source built by
  the restatement of ``src/auxiliaries/adj_seismogram.f90`` from the forward traces (the example runs
  ``xadj_seismogram 27.0 32.0 AA.S0001 1``);
* backward reconstruction (b_deltat = -dt, stored Stacey traction, source at NSTEP-it+1, last frame
  read after step 1): the Newmark scheme is exactly reversible in exact arithmetic, so the
  reconstructed field after adjoint step it equals the forward field after step NSTEP-it+1 to FP32
  rounding;
* kernel accumulation (compute_kernels.f90): one increment against an independent numpy evaluation of
  rho: a.b, kappa: div u div b, mu: D:D_b - div div_b / 3 with symmetric strains;
* final scaling (save_adjoint_kernels.f90): K_rho = -rho dt sum, K_mu = -2 mu dt sum, K_kappa = -kappa dt sum.

The elastic kernel .dat files of the reference examples are missing blobs, so the kernel values
themselves are pinned only by the checks above (the finite-difference test of the differentiable
path covers the gradient)."""
import numpy as np
import pytest

from oracle.adjoint_source import traveltime_adjoint_source
from oracle.oracle import OracleSim
from specfem2d_amd import config


def test_tromp2005_ref_kernel_traces_bit_exact(golden_dir):
    g = np.load(f"{golden_dir}/tromp2005_kernel_ref_seis.npz")
    gk = np.load(f"{golden_dir}/tromp2005_kernel_ref_kernel_traces.npz")
    cfg = config.tromp2005_kernel()
    fwd = OracleSim(cfg, "sp")
    seis, store = fwd.forward(save_stacey=True)
    np.testing.assert_array_equal(seis[0, 0], g["AA.S0001.BXX"])
    last = fwd.fields()
    adj = np.zeros((1, 2, cfg.nstep))
    adj[0, 0] = traveltime_adjoint_source(g["time"], g["AA.S0001.BXX"], 27.0, 32.0)
    aseis, ker = OracleSim(cfg, "sp").adjoint(adj, store, last)
    for c, comp in enumerate(("BXX", "BXZ")):
        np.testing.assert_array_equal(aseis[0, c], gk[f"AA.S0001.{comp}"])
    for k in ("rho", "mu", "kappa"):
        assert np.all(np.isfinite(ker[k])) and np.abs(ker[k]).max() > 0


def test_backward_reconstruction_retraces_the_forward_field():
    cfg = config.small_test_mesh(nx=10, nz=8, nstep=400, gll=True, seed=2)
    fwd = OracleSim(cfg, "sp")
    steps = [1, 50, 200, 399]
    snaps = {}
    seis, store = None, []
    done = 0
    for n in sorted({cfg.nstep - s + 1 for s in steps}):
        s_, st = fwd.forward(nstep=n - done, save_stacey=True)
        store.append(st)
        done = n
        snaps[n] = fwd.fields()[0].copy()
    s_, st = fwd.forward(nstep=cfg.nstep - done, save_stacey=True)
    store.append(st)
    store = np.concatenate(store, axis=0)
    last = fwd.fields()
    adj = np.zeros((cfg.nrec, 2, cfg.nstep))
    _, _, dumps = OracleSim(cfg, "sp").adjoint(adj, store, last, dump_steps=steps)
    scale = max(np.abs(v).max() for v in snaps.values())
    for q, it in enumerate(steps):
        ref = snaps[cfg.nstep - it + 1]
        err = np.abs(dumps[q] - ref).max() / scale
        assert err < 1e-4, (it, err)


def _numpy_kernel_increment(cfg, ibool, xix, xiz, gax, gaz, d, a, b):
    xi = config.XIGLL
    D = np.zeros((5, 5))
    for k in range(5):
        yk = np.zeros(5)
        yk[k] = 1.0
        c = np.polynomial.polynomial.polyfit(xi, yk, 4)
        D[:, k] = np.polynomial.polynomial.polyval(xi, np.polynomial.polynomial.polyder(c))

    def grads(u):
        ue = u[ibool - 1].reshape(-1, 5, 5, 2)  # [e, j, i, c]
        dxi = np.einsum("ik,ejkc->ejic", D, ue)
        dga = np.einsum("jk,ekic->ejic", D, ue)
        X, Z, G, H = (q.reshape(-1, 5, 5, 1) for q in (xix, xiz, gax, gaz))
        return dxi * X + dga * G, dxi * Z + dga * H  # d/dx, d/dz of (ux, uz)

    ux, uz = grads(d)
    bx, bz = grads(b)
    exx, ezz, exz = ux[..., 0], uz[..., 1], 0.5 * (ux[..., 1] + uz[..., 0])
    bxx, bzz, bxz = bx[..., 0], bz[..., 1], 0.5 * (bx[..., 1] + bz[..., 0])
    div, bdiv = exx + ezz, bxx + bzz
    kappa = div * bdiv
    mu = exx * bxx + ezz * bzz + 2 * exz * bxz - kappa / 3
    ag, bg = a[ibool - 1], b[ibool - 1]
    rho = (ag * bg).sum(axis=-1)
    return {"rho": rho.reshape(-1, 25), "mu": mu.reshape(-1, 25), "kappa": kappa.reshape(-1, 25)}


def test_kernel_increment_vs_numpy():
    cfg = config.small_test_mesh(nx=5, nz=4, nstep=2).replace(
        interfaces=[-400.0, -150.0, 0.0], nz_layers=[2, 2], xmax=630.0)
    orc = OracleSim(cfg, "dp")
    rng = np.random.default_rng(11)
    d, a, b = (rng.standard_normal((orc.nglob, 2)) for _ in range(3))
    got = orc.kernel_increment(d, a, b)
    ib = orc.ibool()
    ref = _numpy_kernel_increment(cfg, ib, *(orc.elem_array(n) for n in ("xix", "xiz", "gammax", "gammaz")), d, a, b)
    for k in ("rho", "mu", "kappa"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-10, atol=1e-10 * np.abs(ref[k]).max())


def test_fluid_kernel_increment_vs_numpy():
    """compute_kernels.f90:426-433: rho_ac_kl += 1/rho grad(pot).grad(b_pot) dt, kappa_ac_kl += 1/kappa pdd b_pot dt,
    against an independent numpy evaluation (derivative matrix from polynomial algebra) on a fluid mesh with
    uneven elements and per-point densities and velocities."""
    cfg = config.small_test_mesh(nx=5, nz=4, nstep=2).replace(
        interfaces=[-400.0, -150.0, 0.0], nz_layers=[2, 2], xmax=630.0, material_mode=1)
    rng = np.random.default_rng(5)
    n = cfg.nspec * 25
    cfg = cfg.replace(rho=1000.0 + 200.0 * rng.random(n), vp=1500.0 + 300.0 * rng.random(n), vs=np.zeros(n))
    orc = OracleSim(cfg, "dp")
    assert orc.nspec_acoustic == orc.nspec
    pot, pdd, bpot = (rng.standard_normal(orc.nglob) for _ in range(3))
    dt = 1.7e-3
    got = orc.kernel_ac_increment(pot, pdd, bpot, dt)
    ib = orc.ibool()
    xi = config.XIGLL
    D = np.zeros((5, 5))
    for k in range(5):
        yk = np.zeros(5)
        yk[k] = 1.0
        c = np.polynomial.polynomial.polyfit(xi, yk, 4)
        D[:, k] = np.polynomial.polynomial.polyval(xi, np.polynomial.polynomial.polyder(c))
    X, Z, G, H = (orc.elem_array(q).reshape(-1, 5, 5) for q in ("xix", "xiz", "gammax", "gammaz"))

    def grad(u):
        ue = u[ib - 1].reshape(-1, 5, 5)  # [e, j, i]
        dxi = np.einsum("ik,ejk->eji", D, ue)
        dga = np.einsum("jk,eki->eji", D, ue)
        return dxi * X + dga * G, dxi * Z + dga * H

    (ax, az), (bx, bz) = grad(pot), grad(bpot)
    rho, kappa = orc.elem_array("rho"), orc.elem_array("kappa")
    ref_rho = ((ax * bx + az * bz).reshape(-1, 25) / rho) * dt
    ref_kappa = (pdd[ib - 1] * bpot[ib - 1] / kappa) * dt
    for k, ref in (("rho", ref_rho), ("kappa", ref_kappa)):
        np.testing.assert_allclose(got[k], ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
