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


def pack_float32(example, sub, stations, out):
    """As pack, the traces kept in float32 (long traces; the reference computed them in single precision)."""
    arrays = {}
    for st in stations:
        for comp in ("BXX", "BXZ"):
            t, v = read_trace(os.path.join(REF, "EXAMPLES", example, sub, f"{st}.{comp}.semd"))
            arrays[f"{st}.{comp}"] = v.astype(np.float32)
            arrays["time"] = t
    np.savez_compressed(os.path.join(HERE, out), **arrays)
    print("wrote", out, {k: a.shape for k, a in arrays.items()})


def pack_pressure(example, sub, station, out):
    t, v = read_trace(os.path.join(REF, "EXAMPLES", example, sub, f"{station}.PRE.semp"))
    np.savez_compressed(os.path.join(HERE, out), time=t, **{f"{station}.PRE": v})
    print("wrote", out, t.shape)


def pack_forward_potential(out):
    """The forward potential trace of EXAMPLES/LuoYang_fluid_solid_kernel: its adjoint-source tool
    (adj_source.f90) copies the forward run's AA.S0001.POT.semx into REF_KERNEL/AA.S0001.POT.adj unchanged up to
    t = 1.274 s (zero after)."""
    t, v = read_trace(os.path.join(REF, "EXAMPLES", "LuoYang_fluid_solid_kernel", "REF_KERNEL", "AA.S0001.POT.adj"))
    np.savez_compressed(os.path.join(HERE, out), time=t, **{"AA.S0001.POT": v}, window_end=np.array(1.274))
    print("wrote", out, t.shape)


def copy_inputs(example, files, out, subdir=""):
    """The example's input files (Par_file, SOURCE, interfaces, external mesh files: plain-text data) for the
    on-disk format tests; ``subdir`` keeps them in a subdirectory of the copy (the mesh files' own directory)."""
    import shutil
    dst = os.path.join(HERE, "examples", out, subdir)
    os.makedirs(dst, exist_ok=True)
    for rel in files:
        shutil.copyfile(os.path.join(REF, "EXAMPLES", example, rel), os.path.join(dst, os.path.basename(rel)))
    print("copied", files, "->", dst)


def claerbout_log_numbers(example, out):
    """The figures that EXAMPLES/BENCHMARK_CLAERBOUT_ADJOINT/<example>/REF_KERNEL/output_run.log prints for the
    reference's kernel benchmark (its run_this_example_kernel.sh): misfit of the synthetics against the data, the
    adjoint source's peak, the kernels' min/max and sums of squares, the model-space and data-space inner
    products; written as JSON, numbers as printed."""
    import json
    import re
    with open(os.path.join(REF, "EXAMPLES", "BENCHMARK_CLAERBOUT_ADJOINT", example, "REF_KERNEL",
                           "output_run.log")) as f:
        text = f.read()
    num = r"([-+0-9.eE]+)"

    def first(pattern, n=1):
        m = re.search(pattern, text)
        if not m:
            raise ValueError(f"{example}: '{pattern}' not in the log")
        return [float(m.group(k + 1)) for k in range(n)]

    vals = {
        "max_waveform_difference": first(r"maximum waveform difference \(syn - dat\) =\s+" + num)[0],
        "total_misfit": first(r"total misfit: sum\(s - d\)\^2 =\s+" + num)[0],
        "max_adjoint_source": first(r"maximum amplitude \|f\^adj\| =\s+" + num)[0],
        "kernel_rhop_minmax": first(r"kernel rhop : min/max =\s+" + num + r"\s+" + num, 2),
        "kernel_alpha_minmax": first(r"kernel alpha: min/max =\s+" + num + r"\s+" + num, 2),
        "kernel_beta_minmax": first(r"kernel beta : min/max =\s+" + num + r"\s+" + num, 2),
        "norm_rhop": first(r"norm of rhop kernel  =\s+" + num)[0],
        "norm_alpha": first(r"norm of alpha kernel =\s+" + num)[0],
        "norm_beta": first(r"norm of beta kernel  =\s+" + num)[0],
        "modelspace": first(r"MODELSPACE\s+=\s+" + num)[0],
        "dataspace": first(r"DATASPACE\s+=\s+" + num)[0],
        "dataspace_new": first(r"DATASPACE_NEW\s+=\s+" + num)[0],
    }
    vals["source"] = f"EXAMPLES/BENCHMARK_CLAERBOUT_ADJOINT/{example}/REF_KERNEL/output_run.log"
    with open(os.path.join(HERE, out), "w") as f:
        json.dump(vals, f, indent=1)
    print("wrote", out)


def claerbout_inputs(example, out):
    """Par_file, SOURCE and interfaces of a Claerbout benchmark, plus the STATIONS file its
    create_STATIONS_file.py writes for one station line in x and z (one station at x0 = 700, z0 = 500)."""
    copy_inputs(os.path.join("BENCHMARK_CLAERBOUT_ADJOINT", example),
                ["DATA/Par_file", "DATA/SOURCE", "DATA/interfaces.dat"], out)
    with open(os.path.join(HERE, "examples", out, "STATIONS"), "w") as f:
        f.write("S1\tAA\t700.0\t500.0\t0.0\t0.0\n")


def trace_digests(out):
    """sha256 of one ASCII trace per example and recording kind the reference keeps under EXAMPLES/*/REF_SEIS,
    with its parsed values, so the writer's byte compatibility is checked without copying the files."""
    import glob
    import hashlib
    arrays, names = {}, []
    seen = set()
    for path in sorted(glob.glob(os.path.join(REF, "EXAMPLES", "*", "REF_SEIS", "*.sem?"))):
        kind = (path.split(os.sep)[-3], path[-1])  # one trace per example and recording kind
        if kind in seen:
            continue
        seen.add(kind)
        try:
            t, v = read_trace(path)
        except ValueError:
            continue
        if np.ndim(t) == 0:
            continue
        k = len(names)
        names.append(os.path.relpath(path, os.path.join(REF, "EXAMPLES")))
        arrays[f"t{k}"], arrays[f"v{k}"] = t, v
        arrays[f"sha{k}"] = np.frombuffer(hashlib.sha256(open(path, "rb").read()).digest(), np.uint8)
    arrays["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, out), **arrays)
    print("wrote", out, len(names), "traces")


if __name__ == "__main__":
    pack("Tromp2005_kernel", "REF_SEIS", ["AA.S0001"], "tromp2005_kernel_ref_seis.npz")
    pack_pressure("check_absolute_amplitude_of_pressure_source_seismograms_acoustic", "REF_SEIS", "AA.S0001",
                  "pressure_source_acoustic_ref_seis.npz")
    copy_inputs("Tromp2005_kernel", ["DATA/Par_file", "DATA/SOURCE", "DATA/interfaces_Tromp2005.dat"], "Tromp2005_kernel")
    copy_inputs("check_absolute_amplitude_of_pressure_source_seismograms_acoustic",
                ["REF_SEIS/Par_file", "REF_SEIS/SOURCE", "DATA/interfaces_attenuation_analytic.dat"],
                "pressure_source_acoustic")
    trace_digests("ref_seis_ascii_digests.npz")
    pack_forward_potential("luoyang_fluid_solid_forward_potential.npz")
    pack("infinite_homogeneous_moment_tensor_vertical_dip_slip", "REF_SEIS", ["AA.S0001"],
         "moment_tensor_ref_seis.npz")
    pack("semi_infinite_homogeneous", "REF_SEIS", ["AA.S0001"], "semi_infinite_homogeneous_ref_seis.npz")
    copy_inputs("semi_infinite_homogeneous", ["DATA/Par_file", "DATA/SOURCE", "DATA/interfaces_elastic_analytic.dat"],
                "semi_infinite_homogeneous")
    pack("check_absolute_amplitude_of_force_source_seismograms_elastic", "REF_SEIS", ["AA.S0001"],
         "force_source_elastic_ref_seis.npz")
    copy_inputs("check_absolute_amplitude_of_force_source_seismograms_elastic",
                ["REF_SEIS/Par_file", "REF_SEIS/SOURCE", "DATA/interfaces_attenuation_analytic.dat"],
                "force_source_elastic")
    copy_inputs("infinite_homogeneous_moment_tensor_vertical_dip_slip",
                ["DATA/Par_file", "DATA/SOURCE", "DATA/interfaces_elastic_analytic.dat"], "moment_tensor")
    copy_inputs("LuoYang_fluid_solid_kernel", ["DATA/Par_file", "DATA/SOURCE", "DATA/interfaces.dat"],
                "LuoYang_fluid_solid_kernel")
    pack("Tromp2005_kernel", "REF_KERNEL", ["AA.S0001"], "tromp2005_kernel_ref_kernel_traces.npz")
    pack("check_absolute_amplitude_of_force_source_seismograms_viscoelastic", "REF_SEIS", ["AA.S0001"],
         "force_source_viscoelastic_ref_seis.npz")
    copy_inputs("check_absolute_amplitude_of_force_source_seismograms_viscoelastic",
                ["REF_SEIS/Par_file", "REF_SEIS/SOURCE", "DATA/interfaces_attenuation_analytic.dat"],
                "force_source_viscoelastic")
    ext = os.path.join("fluid_solid", "fluid_solid_external_mesh")
    copy_inputs(ext, ["DATA/Par_file", "DATA/SOURCE"], "fluid_solid_external_mesh")
    copy_inputs(ext, ["mesh/mesh", "mesh/nodes_coords", "mesh/mat", "mesh/surface_free", "mesh/surface_abs"],
                "fluid_solid_external_mesh", subdir="mesh")
    pack_float32(ext, "REF_SEIS", [f"AA.S{k:04d}" for k in range(1, 13)], "fluid_solid_external_mesh_ref_seis.npz")
    pack("anisotropic_isotropic_model", "REF_SEIS", ["AA.S0001"], "anisotropic_isotropic_model_ref_seis.npz")
    copy_inputs("anisotropic_isotropic_model", ["DATA/Par_file", "DATA/SOURCE", "DATA/interfaces.dat"],
                "anisotropic_isotropic_model")
    for example, name in (("ACOUSTIC", "claerbout_acoustic"), ("ACOUSTIC_ELASTIC", "claerbout_acoustic_elastic")):
        claerbout_inputs(example, name)
        claerbout_log_numbers(example, f"{name}_kernel_log.json")
