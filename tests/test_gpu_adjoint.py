"""GPU parity of the adjoint path (SIMULATION_TYPE=3): a SAVE_FORWARD run on the GPU, then an
adjoint run with backward reconstruction through the reference's entry points
(prepare_sim2_or_3_const_device, update_displacement with b_deltat = -dt, the stored Stacey traction,
add_sources_el_sim_type_2_or_3, compute_add_sources_el_s3, transfer_b_fields_to_device after step 1,
compute_kernels_elastic, transfer_kernels_el_to_host).

* Tromp2005_kernel: the adjoint traces against the reference's REF_KERNEL/AA.S0001.BX{X,Z}.semd
  (relative L2 over both components < 2e-5; BXZ alone is ~2.6e-3 of BXX in L2, so each component's error is
  measured against the norm of the pair), the kernels against the FP32 oracle run of the same sequence
  (relative L2 < 2e-5; the oracle's own FP32 and FP64 kernels differ by ~2e-6);
* a heterogeneous mesh with Stacey on all four sides, several receivers and band-limited random adjoint
  traces: traces and kernels against the FP64 oracle, within max(1e-5, 3 x the FP32 oracle's own
  distance from the FP64 oracle) -- K_kappa there sums products of divergences, whose FP32 rounding
  alone puts the FP32 oracle ~7e-5 away from FP64.
"""
import numpy as np
import pytest

from oracle.adjoint_source import traveltime_adjoint_source
from oracle.oracle import OracleSim
from specfem2d_amd import config

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _gpu_forward_then_adjoint(cfg, adj):
    from specfem2d_amd.driver import Simulation
    fwd = Simulation(cfg, simulation_type=1, save_forward=True)
    seis, _, store = fwd.forward(save_stacey=True)
    last = fwd.fields()
    fwd.close()
    sim = Simulation(cfg, simulation_type=3)
    aseis, ker = sim.adjoint(adj, store, last)
    sim.close()
    return seis, store, last, aseis, ker


def test_tromp2005_adjoint_traces_and_kernels(golden_dir):
    g = np.load(f"{golden_dir}/tromp2005_kernel_ref_seis.npz")
    gk = np.load(f"{golden_dir}/tromp2005_kernel_ref_kernel_traces.npz")
    cfg = config.tromp2005_kernel()
    adj = np.zeros((1, 2, cfg.nstep))
    adj[0, 0] = traveltime_adjoint_source(g["time"], g["AA.S0001.BXX"], 27.0, 32.0)
    seis, store, last, aseis, ker = _gpu_forward_then_adjoint(cfg, adj)
    print("forward vs REF_SEIS", _rel(seis[0, 0], g["AA.S0001.BXX"]))
    ref = np.stack([gk["AA.S0001.BXX"], gk["AA.S0001.BXZ"]])
    scale = np.linalg.norm(ref)
    for c, comp in enumerate(("BXX", "BXZ")):
        err = float(np.linalg.norm(aseis[0, c] - ref[c]) / scale)
        print(f"adjoint {comp} vs REF_KERNEL, L2 error / L2 of both components {err:.3e}")
        assert err < 2e-5, (comp, err)
    orc = OracleSim(cfg, "sp")
    oseis, ostore = orc.forward(save_stacey=True)
    olast = orc.fields()
    _, oker = OracleSim(cfg, "sp").adjoint(adj, ostore, olast)
    for k in ("rho", "mu", "kappa"):
        assert np.all(np.isfinite(ker[k]))
        err = _rel(ker[k], oker[k])
        print(f"K_{k} vs oracle rel L2 {err:.3e}")
        assert err < 2e-5, (k, err)


def test_heterogeneous_mesh_adjoint_vs_oracle():
    cfg = config.small_test_mesh(nx=14, nz=10, nstep=500, gll=True, seed=5).replace(
        absorbtop=True, absorbbottom=True, absorbleft=True, absorbright=True)
    rng = np.random.default_rng(3)
    t = np.arange(cfg.nstep)
    env = np.exp(-((t - 0.6 * cfg.nstep) / (0.08 * cfg.nstep)) ** 2)
    # band-limit the noise to the source band (15 Hz Ricker, dt 2 ms): 9-point moving average
    noise = rng.standard_normal((cfg.nrec, 2, cfg.nstep + 8))
    smooth = sum(noise[..., k:k + cfg.nstep] for k in range(9)) / 9.0
    adj = smooth * env * 1e-3
    seis, store, last, aseis, ker = _gpu_forward_then_adjoint(cfg, adj)
    orc = OracleSim(cfg, "dp")
    oseis, ostore = orc.forward(save_stacey=True)
    assert _rel(seis, oseis) < 1e-4
    assert _rel(store, ostore) < 1e-4
    oa, oker = OracleSim(cfg, "dp").adjoint(adj, ostore, orc.fields())
    sp = OracleSim(cfg, "sp")
    _, sstore = sp.forward(save_stacey=True)
    sa, sker = OracleSim(cfg, "sp").adjoint(adj, sstore, sp.fields())
    tol = max(1e-5, 3 * _rel(sa, oa))
    err = _rel(aseis, oa)
    print(f"adjoint traces vs FP64 oracle rel L2 {err:.3e} (tol {tol:.1e})")
    assert err < tol
    for k in ("rho", "mu", "kappa"):
        tol = max(1e-5, 3 * _rel(sker[k], oker[k]))
        err = _rel(ker[k], oker[k])
        print(f"K_{k} vs FP64 oracle rel L2 {err:.3e} (tol {tol:.1e})")
        assert err < tol, (k, err)
