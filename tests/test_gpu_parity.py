"""Parity of the HIP path (called through the Fortran-style C-ABI by the host driver) with the CPU
oracle and with the reference's committed traces.

Tolerance: FP32 seismograms within rel-L2 1e-5 of the FP32 oracle (the reference's own spread across
precision, MPI partitioning and OpenMP is <= 2.6e-6); GPU atomics and FMA contraction change the
rounding, not the scheme.  The error against the FP64 oracle is reported alongside.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config
from specfem2d_amd.driver import Simulation

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_tromp2005_against_reference_traces(golden_dir):
    cfg = config.tromp2005_kernel()
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    g = np.load(f"{golden_dir}/tromp2005_kernel_ref_seis.npz")
    for c, comp in enumerate(("BXX", "BXZ")):
        r = rel(seis[0, c], g[f"AA.S0001.{comp}"])
        print(f"Tromp2005 {comp}: rel-L2 vs REF_SEIS = {r:.3e}")
        assert r < TOL
    dp = OracleSim(cfg, "dp").forward()
    print(f"Tromp2005: rel-L2 vs FP64 oracle = {rel(seis, dp):.3e}")
    sim.close()


def test_plain_elastic_80x60_against_oracle():
    cfg = config.plain_elastic_c1()
    sim = Simulation(cfg)
    seis, t = sim.forward()
    ref = OracleSim(cfg, "sp").forward()
    r = rel(seis, ref)
    print(f"80x60 plain elastic, 1600 steps, 20 receivers: rel-L2 vs FP32 oracle = {r:.3e}, loop {t['loop_ms']:.1f} ms")
    assert r < TOL
    for k in range(cfg.nrec):
        assert rel(seis[k], ref[k]) < 5 * TOL
    sim.close()


@pytest.mark.parametrize("case", ["gll_random", "layers_all_absorbing", "free_boundaries"])
def test_small_meshes_against_oracle(case):
    if case == "gll_random":
        cfg = config.small_test_mesh(nx=14, nz=11, nstep=600, gll=True, seed=5)
    elif case == "layers_all_absorbing":
        cfg = config.small_test_mesh(nx=13, nz=10, nstep=600).replace(
            interfaces=[-1000.0, -620.0, -330.0, 0.0], nz_layers=[3, 4, 3], xmax=1390.0, absorbtop=True)
    else:
        cfg = config.small_test_mesh(nx=10, nz=8, nstep=500, stacey=False)
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    orc = OracleSim(cfg, "sp")
    ref = orc.forward()
    r = rel(seis, ref)
    print(f"{case}: rel-L2 vs FP32 oracle = {r:.3e}")
    assert r < TOL
    # final wavefields: late-time fields are small (energy has left through Stacey), so their
    # tolerance is set by the FP32 oracle's own distance to the FP64 oracle
    d, v, a = sim.fields()
    od, ov, oa = orc.fields()
    dp = OracleSim(cfg, "dp")
    dp.forward()
    dd, dv, _ = dp.fields()
    tol_d = max(TOL, 3 * rel(od, dd))
    tol_v = max(TOL, 3 * rel(ov, dv))
    print(f"{case}: fields rel-L2 displ {rel(d, od):.2e} (tol {tol_d:.1e}), veloc {rel(v, ov):.2e} (tol {tol_v:.1e})")
    assert rel(d, od) < tol_d and rel(v, ov) < tol_v
    sim.close()


def test_save_forward_stacey_storage():
    cfg = config.small_test_mesh(nx=9, nz=7, nstep=300)
    sim = Simulation(cfg, save_forward=True)
    seis, _, store = sim.forward(save_stacey=True)
    _, ref_store = OracleSim(cfg, "sp").forward(save_stacey=True)
    assert store.shape == ref_store.shape
    assert rel(store, ref_store) < TOL
    sim.close()


def test_partial_run_then_continue_matches_full_run():
    # FP32 atomics make the assembly order (and the last bits) vary from run to run, so two runs
    # agree to rounding, not bitwise
    cfg = config.small_test_mesh(nx=8, nz=6, nstep=200)
    a = Simulation(cfg)
    full, _ = a.forward()
    b = Simulation(cfg)
    p1, _ = b.forward(nstep=120)
    p2, _ = b.forward(nstep=80)
    assert b.steps_done == 200
    both = np.concatenate([p1, p2], axis=2)
    assert rel(both, full) < 1e-6
    b.reset()
    again, _ = b.forward()
    assert rel(again, full) < 1e-6
    a.close()
    b.close()


def test_max_norm_entry_point():
    import ctypes
    cfg = config.small_test_mesh(nx=8, nz=6, nstep=150)
    sim = Simulation(cfg)
    sim.forward()
    d, _, _ = sim.fields()
    norm = ctypes.c_float(0)
    mp = ctypes.c_long(sim.mesh_pointer)
    fwd = ctypes.c_int(1)
    sim.gpu.get_norm_elastic_from_device_(ctypes.byref(norm), ctypes.byref(mp), ctypes.byref(fwd))
    assert norm.value == pytest.approx(float(np.sqrt((d.astype(np.float64) ** 2).sum(axis=1)).max()), rel=1e-6)
    sim.close()


@pytest.mark.parametrize("mode", ["patch", "colour", "atomic"])
def test_assembly_modes_agree_with_oracle(mode, monkeypatch):
    # patch: coloured patches of <= 10 elements assembled in LDS (default);
    # colour: coloured elements, plain read-modify-write per colour;
    # atomic: FP32 atomics, as the reference GPU code assembles
    monkeypatch.setenv("SF2D_ASSEMBLY", mode)
    cfg = config.small_test_mesh(nx=23, nz=9, nstep=500, gll=True, seed=9)
    sim = Simulation(cfg)
    if mode in ("patch", "colour"):
        assert sim.colours(2) == 4  # structured quads: 2 x 2 checkerboard of elements or row patches
    seis, _ = sim.forward()
    ref = OracleSim(cfg, "sp").forward()
    r = rel(seis, ref)
    print(f"assembly {mode}: rel-L2 vs FP32 oracle = {r:.3e}")
    assert r < TOL
    sim.close()


def test_affine_geometry_records_are_exact(monkeypatch):
    """Patches of elements whose geometry is constant over the 25 points (to 2^-22 of the largest entry) read
    one record per element; the per-point path (SF2D_NO_AFFINE) must give the same wavefield to FP32 rounding,
    and the oracle's traces to the usual tolerance."""
    from specfem2d_amd.driver import Simulation
    cfg = config.small_test_mesh(nx=14, nz=9, nstep=300, gll=True, seed=3, stacey=False)
    monkeypatch.delenv("SF2D_NO_AFFINE", raising=False)
    monkeypatch.setenv("SF2D_ASSEMBLY", "patch")
    a = Simulation(cfg)
    assert a.affine_elements() == cfg.nspec  # rectangular elements
    sa, _ = a.forward()
    fa = a.fields()
    a.close()
    monkeypatch.setenv("SF2D_NO_AFFINE", "1")
    b = Simulation(cfg)
    assert b.affine_elements() == 0
    sb, _ = b.forward()
    fb = b.fields()
    b.close()
    err = np.linalg.norm(sa - sb) / np.linalg.norm(sb)
    ferr = [rel(x, y) for x, y in zip(fa, fb)]
    print(f"affine records vs per-point geometry: traces rel-L2 {err:.2e}, displ/veloc/accel {ferr[0]:.2e} "
          f"{ferr[1]:.2e} {ferr[2]:.2e}")
    # the record path of a mesh of rectangles is the chain kernel's rectangle form: other rounding than the
    # per-point geometry path
    assert err < 1e-6
    assert ferr[0] < 5e-6 and ferr[1] < 2e-5 and ferr[2] < 5e-5
    ref = OracleSim(cfg, "sp").forward()
    assert np.linalg.norm(sa - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("no_affine", [False, True])
def test_chain_kernel_matches_the_general_patch_kernel(monkeypatch, no_affine):
    """Meshes whose patches are element chains run the row-lane force kernel; the general patch kernel
    (SF2D_NO_CHAIN) must give the same wavefield to FP32 rounding (the two kernels sum the contractions and
    the shared points in different orders: 5e-6 after 300 steps), with per-element and per-point geometry,
    and the oracle's traces to the usual tolerance."""
    from specfem2d_amd.driver import Simulation
    cfg = config.small_test_mesh(nx=23, nz=9, nstep=300, gll=True, seed=4).replace(
        absorbtop=True, absorbbottom=True, absorbleft=True, absorbright=True)
    monkeypatch.setenv("SF2D_ASSEMBLY", "patch")
    if no_affine:
        monkeypatch.setenv("SF2D_NO_AFFINE", "1")
    else:
        monkeypatch.delenv("SF2D_NO_AFFINE", raising=False)
    monkeypatch.delenv("SF2D_NO_CHAIN", raising=False)
    a = Simulation(cfg)
    assert a.chain_kernel()
    sa, _ = a.forward()
    fa = a.fields()
    a.close()
    monkeypatch.setenv("SF2D_NO_CHAIN", "1")
    b = Simulation(cfg)
    assert not b.chain_kernel()
    sb, _ = b.forward()
    fb = b.fields()
    b.close()
    err = rel(sa, sb)
    ferr = [rel(x, y) for x, y in zip(fa, fb)]
    oerr = rel(sa, OracleSim(cfg, "sp").forward())
    print(f"chain vs general patch kernel (no_affine={no_affine}): traces rel-L2 {err:.2e}, displ/veloc/accel "
          f"{ferr[0]:.2e} {ferr[1]:.2e} {ferr[2]:.2e}; chain vs FP32 oracle {oerr:.2e}")
    # the acceleration is the last step's force itself, not smoothed by the time integration
    assert err < 5e-6 and ferr[0] < 5e-6 and ferr[1] < 5e-6 and ferr[2] < 5e-5
    assert oerr < TOL


def test_first_touch_assembly_matches_read_modify_write(monkeypatch):
    """First-touch assembly (the predictor leaves accel alone except at the Stacey and source points; the first
    force update of a point in a step writes instead of adding) against plain read-modify-write with the same
    chain kernel (SF2D_NO_FIRST_TOUCH): the same wavefield to FP32 rounding, the oracle's traces to the usual
    tolerance."""
    from specfem2d_amd.driver import Simulation
    cfg = config.small_test_mesh(nx=23, nz=9, nstep=300, gll=True, seed=4).replace(
        absorbtop=True, absorbbottom=True, absorbleft=True, absorbright=True)
    monkeypatch.setenv("SF2D_ASSEMBLY", "patch")
    monkeypatch.delenv("SF2D_NO_CHAIN", raising=False)
    monkeypatch.delenv("SF2D_NO_FIRST_TOUCH", raising=False)
    a = Simulation(cfg)
    sa, _ = a.forward()
    fa = a.fields()
    a.close()
    monkeypatch.setenv("SF2D_NO_FIRST_TOUCH", "1")
    b = Simulation(cfg)
    sb, _ = b.forward()
    fb = b.fields()
    b.close()
    err = rel(sa, sb)
    ferr = [rel(x, y) for x, y in zip(fa, fb)]
    oerr = rel(sa, OracleSim(cfg, "sp").forward())
    print(f"first-touch vs read-modify-write: traces {err:.2e}, displ/veloc/accel {ferr[0]:.2e} {ferr[1]:.2e} "
          f"{ferr[2]:.2e}; vs FP32 oracle {oerr:.2e}")
    assert err < 5e-6 and ferr[0] < 5e-6 and ferr[1] < 5e-6 and ferr[2] < 5e-5
    assert oerr < TOL


def test_rectangle_form_matches_the_general_chain_kernel(monkeypatch):
    """Axis-aligned affine elements (xi_z = gamma_x = 0) take the chain kernel's rectangle form; the general
    form (SF2D_NO_RECT) gives the same wavefield to FP32 rounding. A mesh with layers of different thickness
    keeps rectangles; the oracle's traces within the usual tolerance."""
    from specfem2d_amd.driver import Simulation
    cfg = config.small_test_mesh(nx=23, nz=9, nstep=300, gll=True, seed=6).replace(
        interfaces=[-900.0, -530.0, -210.0, 0.0], nz_layers=[3, 4, 2], absorbtop=True)
    monkeypatch.setenv("SF2D_ASSEMBLY", "patch")
    for k in ("SF2D_NO_CHAIN", "SF2D_NO_AFFINE", "SF2D_NO_RECT"):
        monkeypatch.delenv(k, raising=False)
    a = Simulation(cfg)
    assert a.chain_kernel() and a.rect_kernel()
    sa, _ = a.forward()
    fa = a.fields()
    a.close()
    monkeypatch.setenv("SF2D_NO_RECT", "1")
    b = Simulation(cfg)
    assert b.chain_kernel() and not b.rect_kernel()
    sb, _ = b.forward()
    fb = b.fields()
    b.close()
    err = rel(sa, sb)
    ferr = [rel(x, y) for x, y in zip(fa, fb)]
    oerr = rel(sa, OracleSim(cfg, "sp").forward())
    print(f"rectangle vs general chain kernel: traces {err:.2e}, displ/veloc/accel {ferr[0]:.2e} {ferr[1]:.2e} "
          f"{ferr[2]:.2e}; vs FP32 oracle {oerr:.2e}")
    assert err < 5e-6 and ferr[0] < 5e-6 and ferr[1] < 2e-5 and ferr[2] < 5e-5
    assert oerr < TOL


def test_moment_tensor_example_from_its_input_files(golden_dir):
    """A moment-tensor source (source_type 2), set up from the reference example's own input files (CPML
    replaced by Stacey edges): traces against the FP32 oracle and the reference's traces before the waves reach
    the absorbing layer."""
    import os

    from specfem2d_amd import io
    cfg, _ = io.config_from_par_file(os.path.join(golden_dir, "examples", "moment_tensor"), cpml="stacey")
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    ref = OracleSim(cfg, "sp").forward()
    g = np.load(f"{golden_dir}/moment_tensor_ref_seis.npz")
    w = g["time"] < 0.22
    err = max(np.abs(seis[0, c][w] - g[f"AA.S0001.{ch}"][w]).max() / np.abs(g[f"AA.S0001.{ch}"]).max()
              for c, ch in enumerate(("BXX", "BXZ")))
    r = rel(seis, ref)
    print(f"moment-tensor example: rel-L2 vs FP32 oracle {r:.3e}; max error vs REF_SEIS before 0.22 s {err:.2e} "
          f"of the peak")
    assert r < TOL and err < 1e-5
    sim.close()
