"""Extract the reference's own committed golden traces into compact fixtures.

Reads ASCII ``time value`` trace files that the reference keeps under
``EXAMPLES/*/REF_SEIS`` and ``REF_KERNEL`` (data files of its own regression tests)
and stores them as float64 arrays in ``tests/golden/*.npz``.  Run in a container
that holds the reference tree; the GPU box only reads the .npz files.
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_trace(path):
    t, v = np.loadtxt(path, unpack=True)
    return t, v


def pack(example, sub, stations, out):
    arrays = {}
    for st in stations:
        for comp in ("BXX", "BXZ"):
            t, v = read_trace(os.path.join(REF, "EXAMPLES", example, sub, f"{st}.{comp}.semd"))
            arrays[f"{st}.{comp}"] = v
            arrays["time"] = t
    np.savez_compressed(os.path.join(HERE, out), **arrays)
    print("wrote", out, {k: a.shape for k, a in arrays.items()})


def pack_pressure(example, sub, station, out):
    t, v = read_trace(os.path.join(REF, "EXAMPLES", example, sub, f"{station}.PRE.semp"))
    np.savez_compressed(os.path.join(HERE, out), time=t, **{f"{station}.PRE": v})
    print("wrote", out, t.shape)


if __name__ == "__main__":
    pack("Tromp2005_kernel", "REF_SEIS", ["AA.S0001"], "tromp2005_kernel_ref_seis.npz")
    pack_pressure("check_absolute_amplitude_of_pressure_source_seismograms_acoustic", "REF_SEIS", "AA.S0001",
                  "pressure_source_acoustic_ref_seis.npz")
    pack("Tromp2005_kernel", "REF_KERNEL", ["AA.S0001"], "tromp2005_kernel_ref_kernel_traces.npz")
