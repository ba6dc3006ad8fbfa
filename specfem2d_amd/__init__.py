"""MI355X-native time-stepping hot path of SPECFEM2D (elastic, Newmark, Stacey).

The product is the HIP library behind the reference's Fortran-to-C GPU entry points
(``include/specfem2d_gpu.h``) and the C++ host driver above it
(``include/specfem2d_driver.h``); this package builds and loads them.
"""
from .config import SimulationConfig, Source  # noqa: F401

__all__ = ["SimulationConfig", "Source", "build"]


def build(verbose: bool = False) -> None:
    """Compile the gfx950 libraries in-tree (hipcc, no GPU needed)."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    jobs = str(min(8, os.cpu_count() or 1))
    res = subprocess.run(["make", "-C", here, "-j", jobs], capture_output=not verbose, text=True)
    if res.returncode != 0:
        raise RuntimeError("building the HIP libraries failed:\n" + (res.stdout or "") + (res.stderr or ""))
