"""Fluid elements on the GPU (acoustic potential, free surface, acoustic Stacey edges, pressure sources and
seismograms, fluid-solid coupling), called through the Fortran-style C-ABI by the host driver, against the CPU
oracle (itself pinned by the reference's pressure-source trace and by plane-wave reflection coefficients) and
against the reference's committed trace.

Tolerance: FP32 seismograms within rel-L2 1e-5 of the FP32 oracle, as for the solid path (atomics and FMA
contraction change the rounding, not the scheme).
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config
from specfem2d_amd.config import Source
from specfem2d_amd.driver import Simulation

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_pressure_source_against_oracle_and_reference(golden_dir):
    cfg = config.pressure_source_acoustic_44x44()
    sim = Simulation(cfg)
    assert sim.nspec_acoustic == cfg.nspec
    seis, t = sim.forward()
    o = OracleSim(cfg, "sp")
    ref = o.forward()
    r = rel(seis, ref)
    g = np.load(f"{golden_dir}/pressure_source_acoustic_ref_seis.npz")
    tr, pr = g["time"], g["AA.S0001.PRE"]
    early = tr < 0.35
    e_ref = np.abs(seis[0, 0][early] - pr[early]).max() / np.abs(pr).max()
    print(f"pressure source 44x44 fluid, 1400 steps: rel-L2 vs FP32 oracle {r:.3e}; max error vs REF_SEIS before "
          f"0.35 s {e_ref:.2e} of the peak; loop {t['loop_ms']:.1f} ms")
    assert r < TOL
    assert e_ref < 1e-5
    assert np.all(seis[0, 1] == 0.0)
    # the final fields are the residue left after the Stacey edges absorbed the pulse, where FP32 rounding
    # dominates (the FP32 oracle is 1.6e-5 / 1.2e-4 from the FP64 one): no further from FP64 than the FP32 oracle
    dp = OracleSim(cfg, "dp")
    dp.forward()
    for gpu_f, sp_f, dp_f in zip(sim.acoustic_fields()[:2], o.acoustic_fields()[:2], dp.acoustic_fields()[:2]):
        assert rel(gpu_f, dp_f) < 2 * rel(sp_f, dp_f) + TOL
    sim.close()


def test_pressure_without_trick_against_oracle():
    # -potential_dot_dot carries the FP32 rounding noise that USE_TRICK_FOR_BETTER_PRESSURE exists to avoid (the
    # FP32 oracle is 6e-4 from the FP64 one here): the GPU must be no further from FP64 than the FP32 oracle
    cfg = config.pressure_source_acoustic_44x44(nstep=900).replace(use_trick=False)
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    sp, dp = OracleSim(cfg, "sp").forward(), OracleSim(cfg, "dp").forward()
    e_gpu, e_sp = rel(seis, dp), rel(sp, dp)
    print(f"pressure (-potential_dot_dot, Gaussian source), 900 steps: rel-L2 vs FP64 oracle {e_gpu:.3e} "
          f"(FP32 oracle {e_sp:.3e}), vs FP32 oracle {rel(seis, sp):.3e}")
    assert e_gpu < 2 * e_sp + TOL
    sim.close()


def _coupled_case(seismotype):
    """24 x 16 elements of 50 m, the top 6 rows fluid under a free surface, a random solid below; a pressure
    source in the fluid and a tilted force in the solid; receivers in the fluid (pressure, with
    USE_TRICK_FOR_BETTER_PRESSURE) or the solid (displacement)."""
    rng = np.random.default_rng(11)
    pert = 0.05 * rng.standard_normal(24 * 16 * 25)
    if seismotype == 4:
        rec = np.array([[310.0, -90.0], [640.0, -210.0], [1010.0, -260.0], [575.0, -150.0]])
    else:
        rec = np.array([[320.0, -320.0], [650.0, -510.0], [990.0, -700.0], [600.0, -301.0]])
    src = [Source(x=430.0, z=-140.0, f0=9.0, factor=1.0e4), Source(x=800.0, z=-520.0, f0=7.0, factor=1.0e9,
                                                                  anglesource=35.0, tshift=0.02)]
    cfg = config.layered_fluid_solid(24, 16, 6, 50.0, 900, 1.0e-3, pert=pert, sources=src, receivers=rec,
                                     seismotype=seismotype, name=f"coupled_24x16_s{seismotype}")
    return cfg.replace(use_trick=seismotype == 4)


@pytest.mark.parametrize("seismotype", [4, 1])
def test_coupled_fluid_solid_against_oracle(seismotype):
    cfg = _coupled_case(seismotype)
    sim = Simulation(cfg)
    assert sim.nspec_acoustic == 24 * 6 and sim.ncoupling == 24
    seis, t = sim.forward()
    o = OracleSim(cfg, "sp")
    ref = o.forward()
    r = rel(seis, ref)
    print(f"coupled 24x16 (6 fluid rows), 900 steps, seismotype {seismotype}: rel-L2 vs FP32 oracle {r:.3e}, "
          f"loop {t['loop_ms']:.1f} ms")
    assert r < TOL
    for k in range(cfg.nrec):
        assert rel(seis[k], ref[k]) < 5 * TOL
    d, v, _ = sim.fields()
    od, ov, _ = o.fields()
    pot, pd, _ = sim.acoustic_fields()
    opot, opd, _ = o.acoustic_fields()
    assert rel(d, od) < 5 * TOL and rel(pot, opot) < 5 * TOL
    # the free surface holds the potential at zero
    isp, code = o.free_surface()
    top = o.ibool()[isp][:, 20:25].reshape(-1) - 1
    assert np.all(code == 3) and np.all(pot[top] == 0) and np.all(pd[top] == 0)
    sim.close()


@pytest.mark.parametrize("orientation", ["above", "below", "beside"])
def test_plane_wave_reflection_coefficient(orientation):
    cfg = config.fluid_solid_plane_wave(orientation=orientation)
    sim = Simulation(cfg)
    seis, t = sim.forward()
    ref = OracleSim(cfg, "sp").forward()
    tt = np.arange(cfg.nstep) * cfg.dt - sim.t0
    z1, z2 = 1020.0 * 1500.0, 2500.0 * 3000.0
    r_exact = (z2 - z1) / (z2 + z1)
    p = seis[1, 0]
    d = p[tt < 0.7]
    direct = d[np.argmax(np.abs(d))]
    m = (tt > 0.8) & (tt < 1.3)
    refl = p[m][np.argmax(np.abs(p[m]))]
    print(f"plane wave, fluid {orientation} the solid: R = {refl / direct:.5f} (exact {r_exact:.5f}); rel-L2 vs FP32 "
          f"oracle {rel(seis, ref):.3e}; loop {t['loop_ms']:.1f} ms")
    assert refl / direct == pytest.approx(r_exact, rel=2e-3)
    assert rel(seis, ref) < TOL
    sim.close()


def test_solid_only_paths_refuse_fluid_meshes():
    cfg = _coupled_case(1).replace(nstep=10)
    sim = Simulation(cfg)
    with pytest.raises(NotImplementedError):
        sim.gradient_forward(np.zeros((cfg.nrec, 2, cfg.nstep)))
    sim.close()


def test_pressure_example_from_its_input_files(golden_dir, tmp_path):
    """The reference example's own Par_file / SOURCE / interfaces (CPML replaced by Stacey edges) run on the GPU
    and written as the reference writes its traces."""
    import os

    from specfem2d_amd import io
    cfg, stations = io.config_from_par_file(os.path.join(golden_dir, "examples", "pressure_source_acoustic"),
                                            cpml="stacey")
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    (path,) = io.write_ascii_seismograms(str(tmp_path), seis, stations, cfg.dt, sim.t0, seismotype=4)
    assert os.path.basename(path) == "AA.S0001.PRE.semp"
    t, p = io.read_ascii_trace(path)
    g = np.load(f"{golden_dir}/pressure_source_acoustic_ref_seis.npz")
    np.testing.assert_array_equal(t, g["time"])
    early = t < 0.35
    err = np.abs(p[early] - g["AA.S0001.PRE"][early]).max() / np.abs(g["AA.S0001.PRE"]).max()
    print(f"pressure example from its input files: max error vs REF_SEIS before 0.35 s {err:.2e} of the peak")
    assert err < 1e-5
    sim.close()


def test_fluid_rectangle_form_matches_per_point_geometry(monkeypatch):
    """Axis-aligned rectangular fluid elements read one {xi_x, gamma_z} record per element; the per-point form
    (SF2D_NO_RECT) gives the same traces to FP32 rounding."""
    cfg = _coupled_case(4).replace(nstep=500)
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    sim.close()
    monkeypatch.setenv("SF2D_NO_RECT", "1")
    sim2 = Simulation(cfg)
    seis2, _ = sim2.forward()
    sim2.close()
    r = rel(seis, seis2)
    print(f"fluid rectangle form vs per-point geometry: rel-L2 {r:.2e}")
    assert r < TOL


@pytest.mark.parametrize("rect", [True, False])
def test_fluid_chain_kernel_matches_per_point_kernel(rect, monkeypatch):
    """The fluid chain kernel (chains of fluid elements, lanes along GLL columns) gives the per-point fluid kernel's
    traces (SF2D_NO_FLUID_CHAIN) to FP32 rounding -- only the order of the sums at a chain's shared columns differs --
    in the rectangle form and with per-point geometry (SF2D_NO_RECT)."""
    if not rect:
        monkeypatch.setenv("SF2D_NO_RECT", "1")
    cfg = _coupled_case(4).replace(nstep=500)
    print(f"case: coupled mesh, fluid chain kernel vs per-point kernel, rect {rect}", flush=True)
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    chain = sim.fluid_chain()
    sim.close()
    monkeypatch.setenv("SF2D_NO_FLUID_CHAIN", "1")
    sim2 = Simulation(cfg)
    seis2, _ = sim2.forward()
    chain2 = sim2.fluid_chain()
    sim2.close()
    r = rel(seis, seis2)
    print(f"fluid chain kernel vs per-point kernel (rect {rect}): rel-L2 {r:.2e}")
    assert chain and not chain2
    assert r < TOL


def test_luoyang_fluid_solid_against_reference_trace(golden_dir, tmp_path):
    """The LuoYang fluid-solid example (a solid strip between two fluids, fluid-solid edges on vertical sides,
    fluid Stacey edges next to them), run from its own input files, against the reference's forward potential
    trace (bit-exact for the oracle) and written as the reference writes a potential trace."""
    import os

    from specfem2d_amd import io
    cfg, stations = io.config_from_par_file(os.path.join(golden_dir, "examples", "LuoYang_fluid_solid_kernel"))
    sim = Simulation(cfg)
    assert sim.nspec_acoustic == 4800 and sim.ncoupling == 120
    seis, t = sim.forward()
    g = np.load(f"{golden_dir}/luoyang_fluid_solid_forward_potential.npz")
    w = g["time"] <= float(g["window_end"])
    r = rel(seis[0, 0][w], g["AA.S0001.POT"][w])
    print(f"LuoYang fluid-solid example, 1600 steps: potential rel-L2 vs the reference's trace {r:.3e}, "
          f"loop {t['loop_ms']:.1f} ms")
    assert r < TOL
    (path,) = io.write_ascii_seismograms(str(tmp_path), seis, stations, cfg.dt, sim.t0, seismotype=6)
    assert os.path.basename(path) == "AA.S0001.POT.semx"
    tt, _ = io.read_ascii_trace(path)
    np.testing.assert_array_equal(tt, g["time"])
    sim.close()


@pytest.mark.parametrize("seismotype", [1, 2, 3])
def test_vector_seismograms_in_fluid_and_solid(seismotype):
    """Displacement, velocity, acceleration (seismotypes 1-3) of receivers in fluid elements (gradient of the
    potential or its time derivatives over the density) next to receivers in solid elements, rotated by anglerec."""
    cfg = _coupled_case(1).replace(nstep=600, anglerec=25.0, seismotype=seismotype, receivers=np.array(
        [[310.0, -90.0], [640.0, -210.0], [1010.0, -299.0], [320.0, -320.0], [650.0, -510.0], [990.0, -301.0]]))
    sim = Simulation(cfg)
    seis, t = sim.forward()
    sim.close()
    sp, dp = OracleSim(cfg, "sp").forward(), OracleSim(cfg, "dp").forward()
    r, e_gpu, e_sp = rel(seis, sp), rel(seis, dp), rel(sp, dp)
    print(f"coupled 24x16, seismotype {seismotype}, 3 fluid + 3 solid receivers: rel-L2 vs FP32 oracle {r:.3e}; vs "
          f"FP64 oracle {e_gpu:.3e} (FP32 oracle {e_sp:.3e}); loop {t['loop_ms']:.1f} ms")
    assert e_gpu < 2 * e_sp + TOL
    for k in range(cfg.nrec):
        assert rel(seis[k], dp[k]) < 2 * rel(sp[k], dp[k]) + 5 * TOL


def test_displacement_in_a_pure_fluid():
    cfg = config.pressure_source_acoustic_44x44(nstep=600).replace(
        seismotype=1, use_trick=False, receivers=np.array([[1500.0, 1500.0], [1210.0, 830.0], [900.0, 1600.0]]))
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    sim.close()
    sp, dp = OracleSim(cfg, "sp").forward(), OracleSim(cfg, "dp").forward()
    e_gpu, e_sp = rel(seis, dp), rel(sp, dp)
    print(f"pure fluid 44x44, displacement of 3 receivers, 600 steps: rel-L2 vs FP32 oracle {rel(seis, sp):.3e}, vs "
          f"FP64 oracle {e_gpu:.3e} (FP32 oracle {e_sp:.3e})")
    assert e_gpu < 2 * e_sp + TOL
