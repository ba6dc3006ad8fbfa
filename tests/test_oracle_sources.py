"""Moment-tensor sources (``source_type = 2``, compute_arrays_source.f90:34-160) in the oracle and the host driver,
and more of the reference's examples set up from their own input files.

Pinned by the reference's ``EXAMPLES/infinite_homogeneous_moment_tensor_vertical_dip_slip/REF_SEIS`` traces: the
oracle, set up from the example's own input files, follows them to 1e-6 of their peak until the waves reach the
reference's CPML layer (Stacey edges here), a window that holds the peak and more than half of the energy.  The
host driver's moment-tensor source arrays equal the oracle's bit for bit."""
import os

import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config, io
from specfem2d_amd.driver import Simulation

WINDOW_END = 0.22  # s: the waves reach the absorbing layer afterwards


def _cfg(golden_dir):
    return io.config_from_par_file(os.path.join(golden_dir, "examples", "moment_tensor"), cpml="stacey")


def test_moment_tensor_example_early_window(golden_dir):
    cfg, stations = _cfg(golden_dir)
    assert cfg.sources[0].source_type == 2 and (cfg.sources[0].Mxz, cfg.sources[0].Mxx) == (1.0, 0.0)
    o = OracleSim(cfg, "sp")
    seis = o.forward()
    g = np.load(f"{golden_dir}/moment_tensor_ref_seis.npz")
    t = g["time"]
    w = t < WINDOW_END
    for c, ch in enumerate(("BXX", "BXZ")):
        ref = g[f"AA.S0001.{ch}"]
        peak = np.abs(ref).max()
        assert np.abs(ref[w]).max() == peak  # the peak is inside the window
        assert np.abs(seis[0, c][w] - ref[w]).max() < 1e-6 * peak


def test_moment_tensor_source_arrays_match_the_driver(golden_dir):
    cfg, _ = _cfg(golden_dir)
    sim, o = Simulation(cfg, device=False), OracleSim(cfg, "sp")
    si, sx, sg, sa, st = sim.sources()
    oi, ox, og, osa, ost = o.sources()
    np.testing.assert_array_equal(si, oi)
    np.testing.assert_array_equal(sa.astype(np.float64), osa)
    np.testing.assert_array_equal(st.astype(np.float64), ost)
    # a moment tensor acts through the gradient of the Lagrange weights: its arrays sum to zero (no net force)
    assert abs(osa[0, :, 0].sum()) < 1e-6 * np.abs(osa).max() and abs(osa[0, :, 1].sum()) < 1e-6 * np.abs(osa).max()
    sim.close()


def test_moment_tensor_in_a_fluid_element_is_refused():
    from specfem2d_amd import config
    from specfem2d_amd.config import Source
    cfg = config.layered_fluid_solid(8, 6, 2, 100.0, 10, 1e-3, sources=[
        Source(x=420.0, z=-50.0, f0=10.0, factor=1.0, source_type=2, Mxx=1.0, Mzz=1.0)])
    with pytest.raises(ValueError):
        Simulation(cfg, device=False)


@pytest.mark.parametrize("example,window_end", [("semi_infinite_homogeneous", 0.23), ("force_source_elastic", 0.26)])
def test_force_source_examples_early_window(golden_dir, example, window_end):
    """Two more reference examples with a force source (a half-space with a free surface; a whole space), set up
    from their own input files: within 1e-6 of their traces' peak until the waves reach the reference's CPML."""
    cfg, _ = io.config_from_par_file(os.path.join(golden_dir, "examples", example), cpml="stacey")
    o = OracleSim(cfg, "sp")
    seis = o.forward()
    g = np.load(f"{golden_dir}/{example}_ref_seis.npz")
    w = g["time"] < window_end
    assert g["time"][0] == pytest.approx(-o.t0, rel=1e-15)
    for c, ch in enumerate(("BXX", "BXZ")):
        ref = g[f"AA.S0001.{ch}"]
        assert np.abs(seis[0, c][w] - ref[w]).max() < 1e-6 * np.abs(ref).max()


def test_anisotropic_example_ref_seis(golden_dir):
    """Anisotropic elements (compute_forces_viscoelastic.F90:372-447, material table icodemat 2): the reference's
    EXAMPLES/anisotropic_isotropic_model (a 5-row anisotropic layer, c11 = 16.5e10, c13 = 5e10, c33 = 6.2e10,
    c55 = 3.96e10, in an isotropic solid) set up from its own input files, whole traces (1500 samples, Stacey on
    all sides) within 1e-5 of the peak; the same model with the layer made isotropic (vp = sqrt(c33/rho),
    vs = sqrt(c55/rho)) is more than 0.1 of the peak away."""
    cfg, _ = io.config_from_par_file(os.path.join(golden_dir, "examples", "anisotropic_isotropic_model"))
    assert cfg.aniso_elem.sum() == 250
    seis = OracleSim(cfg, "sp").forward()
    g = np.load(f"{golden_dir}/anisotropic_isotropic_model_ref_seis.npz")
    iso = OracleSim(cfg.replace(aniso_elem=None, aniso_c=None), "sp").forward()
    for c, ch in enumerate(("BXX", "BXZ")):
        ref = g[f"AA.S0001.{ch}"]
        assert np.abs(seis[0, c] - ref).max() < 1e-5 * np.abs(ref).max()
        assert np.abs(iso[0, c] - ref).max() > 0.1 * np.abs(ref).max()


def test_anisotropic_stress_reduces_to_isotropic():
    """c11 = c33 = lambda + 2 mu, c13 = lambda, c55 = mu, c15 = c35 = 0 is the isotropic stress: with those
    constants in every element the FP64 oracle's run equals the isotropic run to rounding; a nonzero c15 changes
    it."""
    cfg = config.small_test_mesh(nx=6, nz=5, nstep=200)
    rho, vp, vs = (np.asarray(a, float) for a in cfg.material_arrays())
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    n = cfg.nspec * 25
    c = np.stack([lam + 2 * mu, lam, np.zeros(n), lam + 2 * mu, np.zeros(n), mu], axis=1)
    ref = OracleSim(cfg, "dp").forward()
    got = OracleSim(cfg.replace(aniso_elem=np.ones(cfg.nspec, bool), aniso_c=c), "dp").forward()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    c2 = c.copy()
    c2[:, 2] = 0.2 * mu
    other = OracleSim(cfg.replace(aniso_elem=np.ones(cfg.nspec, bool), aniso_c=c2), "dp").forward()
    assert np.abs(other - ref).max() > 1e-3 * np.abs(ref).max()
