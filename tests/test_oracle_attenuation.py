"""Pins the oracle's viscoelastic attenuation (ATTENUATION_VISCOELASTIC, Newmark, N_SLS standard linear solids)
before it is trusted as the checker.

* relaxation times (the linear least-squares Zener fit of attenuation_model.f90 with USE_SOLVOPT false), the
  single-precision convolution coefficients of prepare_attenuation.f90, the memory-variable update of
  compute_attenuation_viscoelastic.f90 and the unrelaxed-moduli stress correction of
  compute_forces_viscoelastic.F90 -- the reference's committed trace
  ``EXAMPLES/check_absolute_amplitude_of_force_source_seismograms_viscoelastic/REF_SEIS`` (Qkappa 22.3, Qmu 20,
  3 solids, f0 18 Hz), set up from the example's own input files: within 1e-6 of its peak until the waves reach
  the reference's CPML layer (Stacey edges here), where the elastic trace of the same set-up is 150 % of the peak away;
* Q = 9999 everywhere: no point attenuates and the run is bit-identical to the elastic one.
"""
import os

import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config, io


def test_viscoelastic_example_ref_seis(golden_dir):
    cfg, _ = io.config_from_par_file(os.path.join(golden_dir, "examples", "force_source_viscoelastic"),
                                     cpml="stacey")
    assert cfg.attenuation and cfg.n_sls == 3 and cfg.attenuation_f0() == 18.0
    assert np.all(cfg.qkappa == 22.3) and np.all(cfg.qmu == 20.0)
    o = OracleSim(cfg, "sp")
    seis = o.forward()
    g = np.load(f"{golden_dir}/force_source_viscoelastic_ref_seis.npz")
    ge = np.load(f"{golden_dir}/force_source_elastic_ref_seis.npz")
    w = g["time"] < 0.26
    for c, ch in enumerate(("BXX", "BXZ")):
        ref = g[f"AA.S0001.{ch}"]
        err = np.abs(seis[0, c][w] - ref[w]).max() / np.abs(ref).max()
        el = np.abs(ge[f"AA.S0001.{ch}"][w] - ref[w]).max() / np.abs(ref).max()
        print(f"{ch}: max error before 0.26 s {err:.2e} of the peak (elastic trace: {el:.2f})")
        assert err < 1e-6 and el > 0.5


def test_q9999_is_elastic():
    cfg = config.tromp2005_kernel().replace(nstep=300)
    a = OracleSim(cfg, "sp").forward()
    b = OracleSim(cfg.replace(attenuation=True), "sp").forward()
    np.testing.assert_array_equal(a, b)


def test_attenuation_refuses_fluids_and_adjoint():
    cfg = config.layered_fluid_solid(8, 6, 2, 100.0, 10, 1e-3, sources=[config.Source(x=420.0, z=-450.0, f0=10.0)])
    with pytest.raises(ValueError):
        OracleSim(cfg.replace(attenuation=True, qkappa=30.0, qmu=20.0), "sp")
    o = OracleSim(config.tromp2005_kernel().replace(nstep=10, attenuation=True, qkappa=30.0, qmu=20.0), "sp")
    with pytest.raises(NotImplementedError):
        o.gradient(np.zeros((1, 2, 10)))
