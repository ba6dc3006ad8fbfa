"""CPML absorbing layers on the GPU (PML_BOUNDARY_CONDITIONS, Newmark; elastic and acoustic elements), driven through
the Fortran-style C-ABI by the host driver (prepare_cpml_device_, pml_boundary_elastic_cuda_,
pml_boundary_acoustic_cuda_), against the FP32 oracle -- which four of the reference's CPML examples pin over their
whole traces -- and against those examples' committed traces.  The viscoelastic example (Qkappa 22.3, Qmu 20) runs
the viscoelastic layer (prepare_cpml_attenuation_device_: the Zener convolutions of the stretched gradients).

Tolerances: FP32 seismograms within rel-L2 1e-5 of the FP32 oracle, as for the other paths; against REF_SEIS the
oracle's own bar, 2e-5 of the peak over the whole trace (the fixtures come from earlier versions of the reference and
carry about 1e-6 of their peak of their own).
"""
import os

import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config, io
from specfem2d_amd.config import Source
from specfem2d_amd.driver import Simulation

pytestmark = pytest.mark.gpu
TOL = 1e-5
REF_BAR = 2e-5
EXAMPLES = [("semi_infinite_homogeneous", ("BXX", "BXZ"), 564), ("force_source_elastic", ("BXX", "BXZ"), 492),
            ("moment_tensor", ("BXX", "BXZ"), 564), ("pressure_source_acoustic", ("PRE",), 492),
            ("force_source_viscoelastic", ("BXX", "BXZ"), 492)]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _example(golden_dir, name):
    cfg, _ = io.config_from_par_file(os.path.join(golden_dir, "examples", name))
    assert cfg.pml and not cfg.stacey
    return cfg


@pytest.mark.parametrize("example,comps,nspec_pml", EXAMPLES)
def test_cpml_examples_full_length(golden_dir, example, comps, nspec_pml):
    cfg = _example(golden_dir, example)
    print(f"case: {example}, {cfg.nx}x{cfg.nz} elements, {cfg.nstep} steps, CPML {cfg.nelem_pml} elements", flush=True)
    sim = Simulation(cfg)
    seis, t = sim.forward()
    stats = sim.pml_stats()
    sim.close()
    ref = OracleSim(cfg, "sp").forward()
    g = np.load(f"{golden_dir}/{example}_ref_seis.npz")
    errs = [float(np.abs(seis[0, c] - g[f"AA.S0001.{ch}"]).max() / np.abs(g[f"AA.S0001.{ch}"]).max())
            for c, ch in enumerate(comps)]
    r = rel(seis, ref)
    print(f"{example}: rel-L2 vs FP32 oracle {r:.3e}; max |GPU - REF_SEIS| / peak over the whole trace "
          f"{max(errs):.2e}; layer {stats[0] + stats[1]} elements, {stats[2]} coefficient sets; "
          f"loop {t['loop_ms']:.1f} ms")
    assert stats[0] + stats[1] == nspec_pml
    assert r < TOL and max(errs) < REF_BAR


@pytest.mark.parametrize("mode", ["colour", "atomic", "patch"])
def test_cpml_with_each_assembly_scheme(golden_dir, mode, monkeypatch):
    """The layer's elements ride in every assembly plan with zero moduli: coloured elements, FP32 atomics and the
    general patch kernel (SF2D_NO_CHAIN) give the oracle's traces as the default chain kernel does."""
    if mode == "patch":
        monkeypatch.setenv("SF2D_NO_CHAIN", "1")
    else:
        monkeypatch.setenv("SF2D_ASSEMBLY", mode)
    cfg = _example(golden_dir, "moment_tensor")
    print(f"case: moment_tensor with {mode} assembly", flush=True)
    sim = Simulation(cfg)
    seis, _ = sim.forward()
    sim.close()
    ref = OracleSim(cfg, "sp").forward()
    r = rel(seis, ref)
    print(f"moment tensor, {mode} assembly: rel-L2 vs FP32 oracle {r:.3e}")
    assert r < TOL


def test_cpml_partial_runs_and_reset(golden_dir):
    """A run split in two calls equals the run in one call, and a reset restarts the memory variables from zero."""
    cfg = _example(golden_dir, "force_source_elastic")
    print("case: force_source_elastic, split run and reset", flush=True)
    sim = Simulation(cfg)
    whole, _ = sim.forward()
    sim.reset()
    a, _ = sim.forward(nstep=700)
    b, _ = sim.forward(nstep=cfg.nstep - 700)
    sim.reset()
    again, _ = sim.forward()
    sim.close()
    split = np.concatenate([a, b], axis=-1)
    print(f"split run vs whole run: max |diff| {np.abs(split - whole).max():.3e}; run after reset "
          f"{np.abs(again - whole).max():.3e}")
    assert np.array_equal(split, whole) and np.array_equal(again, whole)


def test_cpml_fluid_over_solid_against_oracle():
    """A water layer over a heterogeneous solid, the layer on the left, right and bottom sides (fluid and solid
    layer elements, corner elements in the solid, fluid-solid edges crossing the layer), the free surface on top;
    a pressure source in the water and a force in the solid; receivers in both media record displacement."""
    nx, nz, nf = 40, 30, 8
    rng = np.random.default_rng(5)
    pert = np.repeat(0.05 * rng.standard_normal(nx * nz), 25)
    src = [Source(x=1030.0, z=-120.0, f0=10.0, factor=1e4), Source(x=700.0, z=-1100.0, f0=12.0, factor=1e9,
                                                                     anglesource=25.0)]
    rec = np.array([[400.0, -60.0], [1600.0, -150.0], [900.0, -600.0], [1900.0, -1200.0], [200.0, -1400.0]])
    cfg = config.layered_fluid_solid(nx, nz, nf, 50.0, 2000, 1.0e-3, pert=pert, sources=src, receivers=rec)
    cfg = cfg.replace(stacey=False, pml=True, nelem_pml=4)
    print(f"case: fluid over solid {nx}x{nz}, {nf} fluid rows, CPML 4, 2000 steps", flush=True)
    sim = Simulation(cfg)
    seis, t = sim.forward()
    d, _, _ = sim.fields()
    pot, _, _ = sim.acoustic_fields()
    stats = sim.pml_stats()
    sim.close()
    o = OracleSim(cfg, "sp")
    ref = o.forward()
    od, _, _ = o.fields()
    opot, _, _ = o.acoustic_fields()
    exact = OracleSim(cfg, "dp").forward()
    gap = rel(ref, exact)  # the FP32 oracle's own distance from the FP64 one (1.2e-5: the water layer rings)
    r, r64 = rel(seis, ref), rel(seis, exact)
    print(f"fluid over solid with CPML: rel-L2 vs FP32 oracle {r:.3e}, vs FP64 oracle {r64:.3e} (FP32 oracle: "
          f"{gap:.3e}), final displacement {rel(d, od):.2e}, final potential {rel(pot, opot):.2e}; layer {stats[0]} "
          f"solid + {stats[1]} fluid elements; loop {t['loop_ms']:.1f} ms")
    assert stats[0] > 0 and stats[1] > 0
    # the GPU is an FP32 path of its own: within twice the FP32 oracle's own gap of either oracle
    bar = max(TOL, 2.0 * gap)
    assert r < bar and r64 < bar and rel(d, od) < 5 * bar and rel(pot, opot) < 5 * bar
