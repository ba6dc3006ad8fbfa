"""The C-ABI library loads and exports every entry point that include/*.h declares
(the reference's GPU layer, src/gpu/specfem2D_gpu_cuda_method_stubs.c, has 91)."""
import ctypes
import os
import re
import subprocess

import pytest

from specfem2d_amd import driver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(\w+)\s*\(", text)) - {"defined"})


def test_gpu_header_declares_the_reference_abi():
    names = declared("specfem2d_gpu.h")
    assert len(names) == 91, len(names)
    assert all(n.endswith("_") for n in names)


@pytest.mark.parametrize("header,lib", [("specfem2d_gpu.h", driver.GPU_LIB), ("specfem2d_gpu_ext.h", driver.GPU_LIB),
                                        ("specfem2d_driver.h", driver.HOST_LIB)])
def test_library_exports_every_declared_symbol(header, lib):
    assert os.path.exists(lib), "build the libraries first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared(header) if n not in exported]
    assert not missing, missing


def test_libraries_load_without_a_gpu():
    gpu, host = driver.load_libraries()
    assert isinstance(gpu, ctypes.CDLL) and isinstance(host, ctypes.CDLL)
    assert hasattr(gpu, "prepare_constants_device_") and hasattr(gpu, "compute_forces_viscoelastic_cuda_")
