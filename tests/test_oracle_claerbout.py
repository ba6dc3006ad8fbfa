"""The oracle's adjoint runs over fluid and coupled fluid-solid meshes, pinned by the reference's Claerbout kernel
benchmark (``EXAMPLES/BENCHMARK_CLAERBOUT_ADJOINT/{ACOUSTIC,ACOUSTIC_ELASTIC}``): the workflow of its
``run_this_example_kernel.sh`` (``oracle/claerbout.py``) over the FP32 oracle reproduces every figure its
``REF_KERNEL/output_run.log`` prints (``tests/golden/claerbout_*_kernel_log.json``): the waveform misfit of the
Gaussian-perturbed model, the adjoint source's peak, the min/max and sums of squares of the rho', alpha (and beta)
kernels, the model-space inner product and the misfit after the kernel's model update.

What a mistake would move: the fluid's Stacey storage and its backward subtraction, the reconstructed field's
sources and last frame, the adjoint source injection (sign, 1/kappa), the kernel accumulation (gradient products,
1/rho, 1/kappa, dt), the adjoint coupling terms (pressure = adjoint potential, displacement = minus the adjoint
acceleration), the solid kernels' rhop/alpha/beta combination -- each changes the kernels' extrema at the third
digit or earlier; the tolerance is two units in the printed last digit.
"""
import json
import os

import numpy as np
import pytest

from oracle.claerbout import OracleEngine, run
from specfem2d_amd import io

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _close(got, want, rel=2e-7, abs_=0.0):
    got, want = np.atleast_1d(got).astype(np.float64), np.atleast_1d(want).astype(np.float64)
    return np.all(np.abs(got - want) <= rel * np.abs(want) + abs_)


def _check(out, log, coupled):
    # figures printed with 8 significant digits (float32 repr) or 7 ({:e})
    assert _close(out["max_waveform_difference"][0], log["max_waveform_difference"], 3e-7)
    assert _close(out["max_adjoint_source"][0], log["max_adjoint_source"], 3e-7)
    assert _close(out["total_misfit"], log["total_misfit"], 2e-6)
    for k in ("kernel_rhop_minmax", "kernel_alpha_minmax", "kernel_beta_minmax"):
        assert _close(out[k], log[k], 3e-7), (k, out[k], log[k])
    for k in ("norm_rhop", "norm_alpha", "norm_beta"):
        assert _close(out[k], log[k], 3e-7), (k, out[k], log[k])
    assert _close(out["modelspace"], log["modelspace"], 2e-6), (out["modelspace"], log["modelspace"])
    assert _close(out["dataspace_new"], log["dataspace_new"], 2e-6), (out["dataspace_new"], log["dataspace_new"])


@pytest.mark.slow
def test_claerbout_acoustic_kernels_match_reference_log():
    cfg, _ = io.config_from_par_file(os.path.join(GOLDEN, "examples", "claerbout_acoustic"))
    log = json.load(open(os.path.join(GOLDEN, "claerbout_acoustic_kernel_log.json")))
    out = run(cfg, OracleEngine(), coupled=False)
    _check(out, log, False)


@pytest.mark.slow
def test_claerbout_coupled_kernels_match_reference_log():
    cfg, _ = io.config_from_par_file(os.path.join(GOLDEN, "examples", "claerbout_acoustic_elastic"))
    log = json.load(open(os.path.join(GOLDEN, "claerbout_acoustic_elastic_kernel_log.json")))
    out = run(cfg, OracleEngine(), coupled=True)
    _check(out, log, True)
