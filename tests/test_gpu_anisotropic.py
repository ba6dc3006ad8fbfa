"""Anisotropic elements on the GPU (prepare_fields_elastic_device_ with ANISOTROPY and the c_ij arrays, the
per-point force kernel's anisotropic stress; Kernel_2_viscoelastic_impl.cu:616-853): the reference's
EXAMPLES/anisotropic_isotropic_model set up from its own input files (50 x 50 elements, a 5-row anisotropic layer,
Stacey on all sides, 1500 steps) against the FP32 oracle (1e-5 rel-L2) and the reference's traces (1e-5 of the
peak, whole length)."""
import os

import numpy as np
import pytest

from conftest import gpu_available
from oracle.oracle import OracleSim
from specfem2d_amd import io
from specfem2d_amd.driver import Simulation

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def test_anisotropic_example(golden_dir):
    cfg, _ = io.config_from_par_file(os.path.join(golden_dir, "examples", "anisotropic_isotropic_model"))
    sim = Simulation(cfg)
    try:
        seis, _ = sim.forward()
    finally:
        sim.close()
    ref = OracleSim(cfg, "sp").forward()
    err = float(np.linalg.norm(seis - ref) / np.linalg.norm(ref))
    print(f"anisotropic example: GPU vs FP32 oracle rel-L2 {err:.2e}", flush=True)
    assert err < 1e-5
    g = np.load(f"{golden_dir}/anisotropic_isotropic_model_ref_seis.npz")
    for c, ch in enumerate(("BXX", "BXZ")):
        r = g[f"AA.S0001.{ch}"]
        d = np.abs(seis[0, c] - r).max() / np.abs(r).max()
        print(f"anisotropic example {ch}: GPU vs reference {d:.2e} of the peak", flush=True)
        assert d < 1e-5
