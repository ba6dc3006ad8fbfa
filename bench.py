"""Benchmark of the SPECFEM2D time-stepping hot path on one MI355X.

A step is one Newmark step of the wavefields, driven through the reference's Fortran-to-C GPU entry points
(predictors, element forces, CPML or Stacey edges, fluid-solid coupling, sources, correctors, seismograms).  The
metric is GLL-point updates per second: nglob x steps / time-loop wall time.

Default workload: the large fluid-solid forward configuration -- 2000 x 1000 elements of 10 m (20 km x 10 km,
32,012,001 GLL points), the top 1 km an acoustic water layer (vp 1500, rho 1020) under a free surface, below it a
seeded Gaussian random-field solid (correlation length 200 m, +-10 % around vp 3000, vs 1732, rho 2500), a CPML
layer 20 elements thick on the left, right and bottom sides, a Ricker 8 Hz force source, 200 receivers below the
sea floor, dt 2e-4 s.

The same line carries:
  * c2_elastic: the same mesh all elastic with Stacey edges, whose element-force launches give the roofline of the
    dominant kernel (HIP events on the library stream around the force launches of every step inside the timed
    region; one launch per colour, so a launch covers 1/ncolour of the elements);
  * c2_attenuation: the elastic mesh with viscoelastic attenuation everywhere;
  * the FWI gradient time per shot: the exact discrete gradient of an L2 waveform misfit (forward run with a
    checkpoint every 100 steps, segment recomputation and reverse sweep, all on the GPU) on the single-shot
    800 x 400-element, 2000-step adjoint-gradient workload; and batched against plain on the multi-shot
    600 x 300-element, 1500-step workload (shots along the free surface);
  * cpu_baseline: the C oracle's OpenMP build on the host cores, and on one core, over a bounded piece of the
    default workload, with its calibration against the reference solver.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def kernel_min_bytes(nspec: int, nglob: int, n_affine: int, chain: bool, rows: bool = False) -> int:
    """Minimal bytes of the force kernel's representation per step: per element 25 x (kappa, mu) = 200 B, the
    point indices (25 x 4 B for the chain kernels with per-point ids; one 16-byte record per GLL row of a chain of
    up to 12 elements, 80 B per chain, counted as 7 B per element, with row records; 25 x 2 B slot/rank words for
    the general patch kernel), the geometry (16 B for an affine element, 400 B otherwise); per unique point displ
    read and accel read-modify-write (24 B)."""
    per_elem = (207 if rows else 300) if chain else 250
    return per_elem * nspec + 16 * n_affine + 400 * (nspec - n_affine) + 24 * nglob


def force_kernel_bytes(nspec: int, nglob: int) -> int:
    """Algorithmic bytes of one element-force launch: per element 25 x (ibool + xix, xiz,
    gammax, gammaz, kappa, mu) = 700 B; per unique point displ read (8 B) and accel
    read-modify-write (16 B)."""
    return 700 * nspec + 24 * nglob


WORKLOADS = {
    # name: (nx, nz, element size m, dt, receivers, f0); dt keeps the Courant number below 0.45
    "c2_mesh_elastic_2000x1000": (2000, 1000, 10.0, 2.0e-4, 200, 8.0),
    "c3_elastic_800x400": (800, 400, 20.0, 5.0e-4, 200, 10.0),
    "c4_elastic_600x300": (600, 300, 20.0, 5.0e-4, 64, 10.0),
}


def make_config(name: str, nstep: int):
    from specfem2d_amd import config
    nx, nz, h, dt, nrec, f0 = WORKLOADS[name]
    return config.random_elastic(nx, nz, h, nstep, dt, nrec, seed=2024, f0=f0, corr_len=200.0, amplitude=0.10,
                                 vp0=3000.0, vs0=1732.0, rho0=2500.0, name=name)


def cpu_piece_config(nstep: int):
    """A 400 x 200-element piece of the default workload for the CPU baseline: the same element size, time step,
    water-layer share (the top 20 rows), random-field solid and source frequency, with a CPML frame 4 elements thick
    so that the layer holds the same share of the elements as the full mesh's 20-element frame (4 %)."""
    from specfem2d_amd import config
    return config.c2_mesh_coupled(nstep=nstep, nx=400, nz=200, nrec=20, seed=7, nelem_pml=4)


def read_calibration():
    """The committed calibration of the oracle's OpenMP build against the reference solver (oracle/calibrate.py)."""
    path = os.path.join(ROOT, "profiles", "cpu_calibration_c1.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        text = f.read()
    out = {}
    for key in ("calibration_1core", "calibration_allcores_vs_mpi8"):
        m = re.search(rf"^{key}:\s*([0-9.eE+-]+)", text, re.M)
        if m:
            out[key] = float(m.group(1))
    return out or None


def cpu_baseline(target_seconds: float = 10.0):
    """The oracle (C restatement of the reference CPU path, FP32, gcc -O3 with OpenMP: the elements of each colour
    in parallel) on a piece of the default workload, on all host cores the process may use and on one core."""
    from oracle import oracle
    from oracle.oracle import OracleSim
    nstep = 4000
    cfg = cpu_piece_config(nstep)
    want = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)

    def run(threads):
        got = oracle.set_threads(threads)
        orc = OracleSim(cfg, "sp_omp")
        orc.forward(nstep=2)  # warm caches
        steps, chunk = 0, 10
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < target_seconds and steps + chunk <= nstep - 2:
            orc.forward(nstep=chunk)
            steps += chunk
        dt_s = time.perf_counter() - t0
        return got, steps, dt_s, orc.nglob * steps / dt_s, orc.nglob

    threads, steps, dt_s, value, nglob = run(want)
    _, steps1, dt1, value1, _ = run(1)
    oracle.set_threads(want)
    return {
        "value": value,
        "unit": "GLL-pt updates/s",
        "cores": threads,
        "threads": threads,
        "value_1core": value1,
        "kind": "port",
        "sample": f"400x200-element piece of the default workload ({nglob} GLL points: water layer of 20 rows, "
                  f"random-field solid, CPML 4 elements thick), FP32 oracle OpenMP build (oracle/sem2d_oracle.c, "
                  f"gcc -O3 -fopenmp, elements of each colour in parallel): {steps} steps in {dt_s:.1f} s on "
                  f"{threads} threads, {steps1} steps in {dt1:.1f} s on 1",
        "calibration_vs_reference": read_calibration(),
    }


def fwi_gradient(repeats: int = 2):
    """Seconds per shot of the exact FWI gradient on the 800 x 400-element adjoint-gradient workload."""
    from specfem2d_amd import config
    from specfem2d_amd.driver import Simulation
    true_cfg = config.adjoint_gradient_800x400(anomaly=True)
    cfg = config.adjoint_gradient_800x400(anomaly=False)
    sim = Simulation(cfg)
    # observed data: the same simulation run in the anomaly model
    sim.set_model(true_cfg.rho, true_cfg.vp, true_cfg.vs)
    sim.reset()
    obs, _ = sim.forward()
    sim.set_model(cfg.rho, cfg.vp, cfg.vs)
    times = []
    for _ in range(repeats):  # the first run includes the allocation of the checkpoint buffers
        t0 = time.perf_counter()
        chi, _ = sim.gradient_forward(obs, chunk=100)
        g = sim.gradient_backward()
        times.append(time.perf_counter() - t0)
    gnorm = float(np.linalg.norm(g["vs"]))
    sim.close()
    if not (np.isfinite(chi) and chi > 0 and np.isfinite(gnorm) and gnorm > 0):
        raise RuntimeError("degenerate gradient")
    return {"workload": "adjoint_gradient_800x400", "s_per_shot": min(times), "nstep": cfg.nstep,
            "checkpoint_every": 100, "nglob": sim.nglob, "misfit": chi, "grad_vs_l2": gnorm,
            "method": "exact discrete adjoint, checkpointed recompute, FP64 gradient accumulation", "shots": 1}


def fwi_gradient_batched(nshot: int = 32, repeats: int = 2, nstep: int = 1500):
    """Seconds per shot of the exact FWI gradient on the multi-shot 600 x 300-element, 1500-step workload,
    plain (one shot through the reference's entry points, checkpoint every 100 steps) against batched (nshot
    shots advanced together, checkpoints every 39 steps, segments recomputed into their state slots).
    Observed data: the shots run in the random-field model; synthetics: the homogeneous background."""
    from specfem2d_amd import config
    from specfem2d_amd.driver import ShotBatch, Simulation
    from specfem2d_amd import autotune
    true_cfg = config.batched_fwi_600x300(nstep=nstep, true_model=True)
    cfg = config.batched_fwi_600x300(nstep=nstep)
    shots = config.surface_shots(cfg, nshot)
    b = ShotBatch(true_cfg, shots)
    obs = b.forward()
    b.close()
    # launch shapes tuned on this workload's mesh for each path (short forwards from rest), set per handle
    tuned_plain = autotune.tune(cfg.replace(sources=shots[0]), None, steps=20)
    tuned_batch = autotune.tune(cfg, shots, steps=20)
    sim = Simulation(cfg.replace(sources=shots[0]), **autotune.settings(tuned_plain))
    plain = []
    for _ in range(repeats):  # the first run includes the allocation of the checkpoint buffers
        t0 = time.perf_counter()
        sim.gradient_forward(obs[0], chunk=100)
        sim.gradient_backward()
        plain.append(time.perf_counter() - t0)
    sim.close()
    chunk = int(round(np.sqrt(cfg.nstep)))
    b = ShotBatch(cfg, shots, **autotune.settings(tuned_batch))
    batched = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        chi, _ = b.gradient_forward(obs, chunk=chunk)
        g = b.gradient_backward()
        batched.append(time.perf_counter() - t0)
    b.close()
    gnorm = float(np.linalg.norm(g["vs"]))
    if not (np.isfinite(chi) and chi > 0 and np.isfinite(gnorm) and gnorm > 0):
        raise RuntimeError("degenerate batched gradient")
    sp, sb = min(plain), min(batched) / nshot
    return {"workload": "batched_fwi_600x300", "nstep": cfg.nstep, "nglob": cfg.nglob, "shots_per_batch": nshot,
            "plain_s_per_shot": sp, "batched_s_per_shot": sb, "speedup_per_shot": sp / sb,
            "checkpoint_every": {"plain": 100, "batched": chunk}, "misfit_sum": chi, "grad_vs_l2": gnorm,
            "autotuned": {"plain": {k: tuned_plain[k] for k in ("chain_elems", "shots_per_wave")},
                          "batched": {k: tuned_batch[k] for k in ("chain_elems", "shots_per_wave")}},
            "note": "64 shots = 2 batches of 32 (checkpoint memory)"}


def batch_forward(nshot: int, nstep: int, shots_per_wave: int = 0):
    """A batched forward of the multi-shot FWI workload's mesh (600 x 300 elements, random-field model): nshot
    shots along the surface, nstep steps from rest, device time (profiling of the batched force kernel)."""
    from specfem2d_amd import config
    from specfem2d_amd.driver import ShotBatch
    cfg = config.batched_fwi_600x300(nstep=nstep, true_model=True)
    b = ShotBatch(cfg, config.surface_shots(cfg, nshot))
    if shots_per_wave:
        b.set_shots_per_wave(shots_per_wave)
    b.forward()  # warm-up
    seis, ms = b.forward(timing=True)
    out = {"workload": "batched_forward_600x300", "shots": nshot, "nstep": nstep, "nglob": b.nglob,
           "shots_per_wave": shots_per_wave, "ms_per_step": ms / nstep, "us_per_shot_step": 1e3 * ms / nstep / nshot,
           "finite": bool(np.all(np.isfinite(seis)))}
    b.close()
    return out


def read_traffic(name: str):
    """HBM bytes per force launch (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy) from the
    committed rocprofv3 PMC summary of this workload, if present."""
    path = os.path.join(ROOT, "profiles", f"elastic_forces_{name}_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def copy_calibration(mib: int, reps: int = 5):
    """A plain device-to-device copy of known size, for scaling the HBM byte counters: each launch reads
    and writes exactly ``mib`` MiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": reps}), file=sys.stderr)


def c2_coupled(steps: int = 100, warmup: int = 10):
    """Forward of the coupled C2 mesh (2000 x 1000 elements of 10 m, the top 100 rows a water layer under a free
    surface, a random-field solid below, fluid-solid coupling along the sea floor): GLL-point updates per second
    over all points (fluid and solid), wall clock around the timed steps."""
    import torch
    from specfem2d_amd import config
    from specfem2d_amd.driver import Simulation
    cfg = config.c2_mesh_coupled(nstep=warmup + steps)
    t = time.perf_counter()
    sim = Simulation(cfg)
    setup = time.perf_counter() - t
    sim.forward(nstep=warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seis, timing = sim.forward(nstep=steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {"workload": "c2_mesh_coupled_2000x1000", "nspec": sim.nspec, "nspec_fluid": sim.nspec_acoustic,
           "nglob": sim.nglob, "fluid_solid_edges": sim.ncoupling, "steps": steps, "warmup": warmup,
           "ms_per_step": wall * 1e3 / steps, "value": sim.nglob * steps / wall, "unit": "GLL-pt updates/s",
           "gpu_loop_ms_events": timing["loop_ms"], "finite": bool(np.all(np.isfinite(seis))), "setup_s": setup,
           "boundaries": "free surface on the water, Stacey bottom/left/right (no CPML)"}
    sim.close()
    return out


def c2_attenuation(steps: int = 50, warmup: int = 5):
    """Forward of the C2 elastic mesh (2000 x 1000 elements of 10 m, the random-field model) with viscoelastic
    attenuation everywhere (Qkappa 60, Qmu 30, 3 standard linear solids): GLL-point updates per second, and the
    element-force launches' share of the step with the survey's algorithmic bytes for this kernel, the elastic
    700 B per element plus 3600 B per element of attenuation, plus 24 B per point.  This build's chain kernel
    reads and writes 2723 B per element (25 points x (moduli 8 B, coefficient-set index 4 B, 9 memory variables and
    3 previous strains read and written 96 B), a 16-byte geometry record and about 7 B of point-id row records per
    element) plus 24 B per point."""
    import torch
    from specfem2d_amd.driver import Simulation
    cfg = make_config("c2_mesh_elastic_2000x1000", warmup + steps).replace(attenuation=True, qkappa=60.0, qmu=30.0)
    sim = Simulation(cfg)
    sim.forward(nstep=warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seis, timing = sim.forward(nstep=steps, phase_timing=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    traffic = (700 + 3600) * sim.nspec + 24 * sim.nglob
    own = 2723 * sim.nspec + 24 * sim.nglob
    forces_s = timing["forces_ms"] * 1e-3 / steps
    out = {"workload": "c2_mesh_elastic_2000x1000 + attenuation (Qkappa 60, Qmu 30, N_SLS 3)", "nspec": sim.nspec,
           "nglob": sim.nglob, "steps": steps, "ms_per_step": wall * 1e3 / steps, "value": sim.nglob * steps / wall,
           "unit": "GLL-pt updates/s", "forces_ms_per_step": forces_s * 1e3,
           "forces_roofline": {"bound": "hbm", "achieved": traffic / forces_s / 1e9, "peak": 8000.0, "unit": "GB/s",
                               "frac": traffic / forces_s / 8e12, "traffic": None,
                               "representation_bytes_per_step": own,
                               "representation_GBs": own / forces_s / 1e9},
           "finite": bool(np.all(np.isfinite(seis)))}
    sim.close()
    return out


def c2_elastic(steps: int, warmup: int):
    """The 2000 x 1000 mesh all elastic (the random-field model, Stacey bottom/left/right, free surface on top):
    GLL-point updates per second, and the roofline of the element-force kernel from HIP events around its launches
    (on the library's stream) in every timed step."""
    import torch
    from specfem2d_amd.driver import Simulation
    name = "c2_mesh_elastic_2000x1000"
    cfg = make_config(name, warmup + steps)
    t_setup = time.perf_counter()
    sim = Simulation(cfg)
    t_setup = time.perf_counter() - t_setup
    sim.forward(nstep=warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seis, timing = sim.forward(nstep=steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nglob, nspec = sim.nglob, sim.nspec
    # one force launch per colour; the events bracket all launches of a step
    launches = sim.colours(2)
    patch_mode = os.environ.get("SF2D_ASSEMBLY", "patch") == "patch"
    n_affine = sim.affine_elements() if patch_mode else 0
    chain = patch_mode and sim.chain_kernel()
    fk_ms = timing["forces_ms"] / steps / launches
    fk_bytes = force_kernel_bytes(nspec, nglob) / launches
    achieved = fk_bytes / (fk_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": (("k_elastic_forces_chain" if os.environ.get("SF2D_ROW_CHAIN") else "k_elastic_forces_col")
                   if chain else "k_elastic_forces_patch") if patch_mode
                  else "k_elastic_forces",
        "measured_on": name,
        "launches_per_step": launches,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": read_traffic(name),
        "algorithmic_bytes_per_launch": fk_bytes,
        "avg_launch_ms": fk_ms,
        # the bytes this kernel's own representation needs at minimum: affine elements carry one 16-byte
        # geometry record instead of 25 x 16 bytes (what it reads, not the formula above)
        "kernel_min_bytes_per_launch": kernel_min_bytes(nspec, nglob, n_affine, chain,
                                                        chain and sim.col_rows()) / launches,
        "chain_row_records": bool(chain and sim.col_rows()),
    }
    out = {"workload": name, "nspec": nspec, "nglob": nglob, "receivers": cfg.nrec, "dt": cfg.dt,
           "boundaries": "Stacey bottom/left/right, free surface top", "steps": steps, "warmup": warmup,
           "ms_per_step": wall * 1e3 / steps, "value": nglob * steps / wall, "unit": "GLL-pt updates/s",
           "gpu_loop_ms_events": timing["loop_ms"], "setup_s": t_setup,
           "assembly": os.environ.get("SF2D_ASSEMBLY", "patch"), "affine_elements": n_affine,
           "finite": bool(np.all(np.isfinite(seis)))}
    sim.close()
    return out, roofline


def c5_uneven(n: int = 48, seed: int = 5, max_groups: int = 6):
    """The uneven batch: ``n`` instances of log-uniform sizes (100..1000 x 100..500 elements of 20 m), half water
    with a pressure source, half viscoelastic random-field solids (3 solids, Q 30..100) with a force, CPML 10
    elements thick, NSTEP uniform in 1000..4000.  They run as unions of instances of similar NSTEP (at most
    ``max_groups``; each union one launch per kernel for all its instances), from rest through each union's longest
    NSTEP, the unions one after the other.  Useful throughput: the sum over the instances of GLL points x their own
    NSTEP, per second of all runs; ``padded_fraction``: the share of the unions' updates spent on instances that
    have already ended."""
    import torch
    from specfem2d_amd import config
    from specfem2d_amd.driver import Simulation, UnevenRuns
    cfgs = config.uneven_batch_instances(n=n, seed=seed)
    t = time.perf_counter()
    runs = UnevenRuns(cfgs, max_groups=max_groups)
    setup = time.perf_counter() - t
    print(f"uneven batch: {n} instances in {len(runs.groups)} unions set up in {setup:.0f} s", file=sys.stderr,
          flush=True)
    for b in runs.batches:
        Simulation.forward(b, nstep=5)  # first launches of every kernel, outside the timed run
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seis, timing = runs.forward()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    useful, union = runs.useful_updates(), runs.union_updates()
    groups = [{"instances": len(g), "nglob": b.nglob, "nstep": b.cfg.nstep,
               "ms_per_step": t["loop_ms"] / b.cfg.nstep} for g, b, t in zip(runs.groups, runs.batches,
                                                                          timing["groups"])]
    out = {"workload": f"c5_uneven_batch_{n}", "instances": n, "acoustic": sum(1 for c in cfgs if not c.attenuation),
           "viscoelastic": sum(1 for c in cfgs if c.attenuation), "nspec": sum(b.nspec for b in runs.batches),
           "nglob": sum(b.nglob for b in runs.batches), "nstep_range": [min(c.nstep for c in cfgs),
                                                                       max(c.nstep for c in cfgs)],
           "value": useful / wall, "unit": "GLL-pt updates/s (useful)", "wall_s": wall,
           "union_updates_per_s": union / wall, "padded_fraction": 1.0 - useful / union, "unions": groups,
           "gpu_loop_ms_events": timing["loop_ms"], "setup_s": setup,
           "finite": bool(all(np.all(np.isfinite(s)) for s in seis))}
    runs.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="seconds of each CPU baseline leg")
    ap.add_argument("--no-gradient", action="store_true", help="skip the FWI gradient time per shot")
    ap.add_argument("--fwi-only", action="store_true",
                    help="only the batched-vs-plain FWI gradient (profiling); prints its JSON object")
    ap.add_argument("--fwi-steps", type=int, default=1500, help="time steps of the batched FWI workload")
    ap.add_argument("--batch-shots", type=int, default=32,
                    help="shots per batch of the multi-shot FWI gradient (32 fill about 170 GB of checkpoints)")
    ap.add_argument("--no-attenuation", action="store_true", help="skip the C2 forward with attenuation")
    ap.add_argument("--attenuation-only", action="store_true",
                    help="only the C2 forward with viscoelastic attenuation (profiling); prints its JSON object")
    ap.add_argument("--coupled-only", action="store_true",
                    help="only the coupled fluid-solid C2 forward with Stacey edges (profiling); prints its JSON")
    ap.add_argument("--elastic-only", action="store_true",
                    help="only the all-elastic C2 forward and its force-kernel roofline (profiling)")
    ap.add_argument("--batch-forward-only", type=int, default=0, metavar="SHOTS",
                    help="only a batched forward of the 600x300 mesh with this many shots (profiling)")
    ap.add_argument("--shots-per-wave", type=int, default=0, help="shots per wave of --batch-forward-only (0: default)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the default workload (profiling); prints its JSON line without the other legs")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven batch of 48 instances")
    ap.add_argument("--uneven-only", type=int, default=0, metavar="N",
                    help="only the uneven batch, of N instances (48: the full workload)")
    ap.add_argument("--uneven-groups", type=int, default=6, help="most unions of the uneven batch (1: one union)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the timed region, run a plain copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU layout (shots per GPU, "
                                   "strip decomposition with halo exchange) is a design, not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    from specfem2d_amd import config
    from specfem2d_amd.driver import Simulation

    if args.fwi_only:
        print(json.dumps(fwi_gradient_batched(args.batch_shots, nstep=args.fwi_steps)))
        return
    if args.uneven_only:
        print(json.dumps(c5_uneven(args.uneven_only, max_groups=args.uneven_groups)))
        return
    if args.batch_forward_only:
        print(json.dumps(batch_forward(args.batch_forward_only, args.steps, args.shots_per_wave)))
        return
    if args.attenuation_only:
        print(json.dumps(c2_attenuation(min(args.steps, 50), min(args.warmup, 5))))
        return
    if args.coupled_only:
        print(json.dumps(c2_coupled(min(args.steps, 100), min(args.warmup, 10))))
        return
    if args.elastic_only:
        el, roof = c2_elastic(args.steps, args.warmup)
        el["roofline"] = roof
        print(json.dumps(el))
        if args.copy_calibration_mib:
            copy_calibration(args.copy_calibration_mib)
        return

    # the default workload: water layer, random-field solid, CPML 20 elements thick
    cfg = config.c2_mesh_coupled(nstep=args.warmup + args.steps, nelem_pml=20)
    t_setup = time.perf_counter()
    sim = Simulation(cfg)
    t_setup = time.perf_counter() - t_setup
    sim.forward(nstep=args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seis, timing = sim.forward(nstep=args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if not np.all(np.isfinite(seis)):
        print(json.dumps({"error": "non-finite seismograms"}))
        sys.exit(1)
    pml = sim.pml_stats()
    out = {
        "metric": "elastic GLL-pt updates/s",
        "value": sim.nglob * args.steps / wall,
        "unit": "GLL-pt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian random-field solid under a water layer)",
        "config": {
            "workload": "c2_fluid_solid_cpml_2000x1000",
            "nspec": sim.nspec,
            "nspec_fluid": sim.nspec_acoustic,
            "nglob": sim.nglob,
            "receivers": cfg.nrec,
            "dt": cfg.dt,
            "boundaries": "free surface on the water; CPML 20 elements thick on the left, right and bottom",
            "cpml_elements": {"solid": pml[0], "fluid": pml[1]},
            "fluid_solid_edges": sim.ncoupling,
            "parallelism": "1 GPU",
        },
        "gpu_loop_ms_events": timing["loop_ms"],
        "setup_s": t_setup,
    }
    sim.close()
    if args.headline_only:
        print(json.dumps(out))
        return
    el, roof = c2_elastic(args.steps, args.warmup)
    out["roofline"] = roof
    out["c2_elastic"] = el
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
    if not args.no_attenuation:
        out["c2_attenuation"] = c2_attenuation()
    if not args.no_gradient:
        out["fwi_gradient"] = fwi_gradient()
        out["fwi_gradient_batched"] = fwi_gradient_batched(args.batch_shots, nstep=args.fwi_steps)
    if not args.no_uneven:
        out["c5_uneven"] = c5_uneven(max_groups=args.uneven_groups)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
