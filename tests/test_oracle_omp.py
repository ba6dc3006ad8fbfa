"""The OpenMP build of the oracle (liboracle_sp_omp.so: -O3, the elements of each colour of a greedy colouring in
parallel) is what the CPU baseline times; it must compute what the serial parity build computes.  Only the summation
order at shared points differs, so the traces agree to FP32 rounding (rel-L2 2.4e-6 on the 80 x 60 plain elastic
configuration, the same spread as the reference's own OpenMP build against its serial one, 2.55e-6)."""
import numpy as np

from oracle.oracle import OracleSim
from specfem2d_amd import config
from specfem2d_amd.config import Source


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_openmp_build_matches_serial_elastic():
    """(The thread count does not change the result: the elements of one colour share no point.)"""
    cfg = config.small_test_mesh(nx=20, nz=14, nstep=400, gll=True, seed=3)
    ref = OracleSim(cfg, "sp").forward()
    omp = OracleSim(cfg, "sp_omp").forward()
    assert rel(omp, ref) < 1e-5


def test_openmp_build_matches_serial_coupled_cpml():
    """Water over a random solid with a CPML frame: acoustic and elastic colours, layer memory variables."""
    nx, nz = 24, 16
    rng = np.random.default_rng(2)
    pert = np.repeat(0.05 * rng.standard_normal(nx * nz), 25)
    src = [Source(x=600.0, z=-500.0, f0=10.0, factor=1e9)]
    rec = np.array([[300.0, -60.0], [900.0, -400.0]])
    cfg = config.layered_fluid_solid(nx, nz, 4, 50.0, 600, 1.0e-3, pert=pert, sources=src, receivers=rec)
    cfg = cfg.replace(stacey=False, pml=True, nelem_pml=3)
    ref = OracleSim(cfg, "sp").forward()
    omp = OracleSim(cfg, "sp_omp").forward()
    assert rel(omp, ref) < 1e-5
