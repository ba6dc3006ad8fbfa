"""Benchmark of the SPECFEM2D elastic time-stepping hot path on one MI355X.

A step is one Newmark step of the elastic wavefield, driven through the reference's
Fortran-to-C GPU entry points (predictor, element forces, Stacey, sources, corrector,
seismograms).  The metric is elastic GLL-point updates per second: nglob x steps / time.

Default workload: the 2000 x 1000-element mesh of the large forward configuration
(20 km x 10 km, 10 m elements, 32,012,001 GLL points) with a seeded Gaussian random-field
elastic model (correlation length 200 m, +-10 % around vp 3000, vs 1732, rho 2500), Stacey
on three sides and a free surface, Ricker 8 Hz source, 200 surface receivers, dt 2e-4 s
(Courant number 0.38 over the closest GLL points at the model's highest vp).  The water
layer, fluid-solid coupling and CPML of that configuration are not part of this path yet,
so the whole mesh is elastic with Stacey edges.

The same line carries the second headline figure, the FWI gradient time per shot: the exact
discrete gradient of an L2 waveform misfit (forward run with a checkpoint every 100 steps,
segment recomputation and reverse sweep, all on the GPU) on the single-shot 800 x 400-element,
2000-step adjoint-gradient workload, observed data from a +5 % vs anomaly; and batched against
plain on the multi-shot 600 x 300-element, 1500-step workload (shots along the free surface).

Prints one JSON line with the throughput, the roofline of the element-force kernel (HIP
events on the library stream around the force launches of every step inside the timed region;
one launch per element colour, so a launch covers 1/ncolour of the elements) and a CPU
baseline (the C oracle, one thread, on a bounded piece of the same workload).
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


def kernel_min_bytes(nspec: int, nglob: int, n_affine: int, chain: bool) -> int:
    """Minimal bytes of the force kernel's representation per step: per element 25 x (kappa, mu) = 200 B, the
    25 point indices (4 B each for the row-lane chain kernel, 2 B slot/rank words for the general patch kernel),
    the geometry (16 B for an affine element, 400 B otherwise); per unique point displ read and accel
    read-modify-write (24 B)."""
    per_elem = 300 if chain else 250
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


def cpu_baseline(name: str, target_seconds: float = 15.0):
    """Oracle (C restatement of the reference CPU path, FP32, 1 thread) on a piece of the workload:
    a 250 x 125-element mesh with the same element size, model statistics and boundaries."""
    from oracle.oracle import OracleSim
    from specfem2d_amd import config
    _, _, h, dt, _, f0 = WORKLOADS[name]
    nx, nz = 250, 125
    nstep = 2000
    cfg = config.random_elastic(nx, nz, h, nstep, dt, 20, seed=7, f0=f0, name="cpu_sample")
    orc = OracleSim(cfg, "sp")
    orc.forward(nstep=2)  # warm caches
    steps = 0
    t0 = time.perf_counter()
    chunk = 20
    while time.perf_counter() - t0 < target_seconds and steps + chunk <= nstep - 2:
        orc.forward(nstep=chunk)
        steps += chunk
    dt_s = time.perf_counter() - t0
    return {
        "value": orc.nglob * steps / dt_s,
        "unit": "GLL-pt updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{nx}x{nz}-element piece of the workload ({orc.nglob} GLL points), {steps} Newmark steps, "
                  f"FP32 oracle (oracle/sem2d_oracle.c, gcc -O2), {dt_s:.1f} s",
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
    saved = {k: os.environ.get(k) for k in ("SF2D_CHAIN_ELEMS", "SF2D_SHOTS_PER_WAVE")}
    # launch shapes tuned on this workload's mesh for each path (short forwards from rest)
    tuned_plain = autotune.tune(cfg.replace(sources=shots[0]), None, steps=20)
    tuned_batch = autotune.tune(cfg, shots, steps=20)
    autotune.apply(tuned_plain)
    sim = Simulation(cfg.replace(sources=shots[0]))
    plain = []
    for _ in range(repeats):  # the first run includes the allocation of the checkpoint buffers
        t0 = time.perf_counter()
        sim.gradient_forward(obs[0], chunk=100)
        sim.gradient_backward()
        plain.append(time.perf_counter() - t0)
    sim.close()
    for k, v in saved.items():
        os.environ.pop(k, None) if v is None else os.environ.__setitem__(k, v)
    autotune.apply(tuned_batch)
    chunk = int(round(np.sqrt(cfg.nstep)))
    b = ShotBatch(cfg, shots)
    batched = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        chi, _ = b.gradient_forward(obs, chunk=chunk)
        g = b.gradient_backward()
        batched.append(time.perf_counter() - t0)
    b.close()
    for k, v in saved.items():
        os.environ.pop(k, None) if v is None else os.environ.__setitem__(k, v)
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
    element-force launches' share of the step with their algorithmic traffic, 3300 B per element (25 points x
    (ibool 4 B, geometry 16 B, moduli 8 B, attenuation 104 B: a coefficient-set index read, 9 memory variables
    and 3 previous strains read and written)) plus 24 B per point (displacement read, acceleration
    read-modify-write)."""
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
    traffic = 3300 * sim.nspec + 24 * sim.nglob
    forces_s = timing["forces_ms"] * 1e-3 / steps
    out = {"workload": "c2_mesh_elastic_2000x1000 + attenuation (Qkappa 60, Qmu 30, N_SLS 3)", "nspec": sim.nspec,
           "nglob": sim.nglob, "steps": steps, "ms_per_step": wall * 1e3 / steps, "value": sim.nglob * steps / wall,
           "unit": "GLL-pt updates/s", "forces_ms_per_step": forces_s * 1e3,
           "forces_roofline": {"bound": "hbm", "achieved": traffic / forces_s / 1e9, "peak": 8000.0, "unit": "GB/s",
                               "frac": traffic / forces_s / 8e12, "traffic": None},
           "finite": bool(np.all(np.isfinite(seis)))}
    sim.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2_mesh_elastic_2000x1000", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-gradient", action="store_true", help="skip the FWI gradient time per shot")
    ap.add_argument("--fwi-only", action="store_true",
                    help="only the batched-vs-plain FWI gradient (profiling); prints its JSON object")
    ap.add_argument("--fwi-steps", type=int, default=1500, help="time steps of the batched FWI workload")
    ap.add_argument("--batch-shots", type=int, default=32,
                    help="shots per batch of the multi-shot FWI gradient (32 fill about 170 GB of checkpoints)")
    ap.add_argument("--no-coupled", action="store_true", help="skip the coupled fluid-solid C2 forward")
    ap.add_argument("--no-attenuation", action="store_true", help="skip the C2 forward with attenuation")
    ap.add_argument("--attenuation-only", action="store_true",
                    help="only the C2 forward with viscoelastic attenuation (profiling); prints its JSON object")
    ap.add_argument("--coupled-only", action="store_true",
                    help="only the coupled fluid-solid C2 forward (profiling); prints its JSON object")
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
    from specfem2d_amd.driver import Simulation

    if args.fwi_only:
        print(json.dumps(fwi_gradient_batched(args.batch_shots, nstep=args.fwi_steps)))
        return
    if args.attenuation_only:
        print(json.dumps(c2_attenuation(min(args.steps, 50), min(args.warmup, 5))))
        return
    if args.coupled_only:
        print(json.dumps(c2_coupled(min(args.steps, 100), min(args.warmup, 10))))
        return
    cfg = make_config(args.workload, args.warmup + args.steps)
    t_setup = time.perf_counter()
    sim = Simulation(cfg)
    t_setup = time.perf_counter() - t_setup
    # warmup
    sim.forward(nstep=args.warmup)
    torch.cuda.synchronize()
    # timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seis, timing = sim.forward(nstep=args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if not np.all(np.isfinite(seis)):
        print(json.dumps({"error": "non-finite seismograms"}))
        sys.exit(1)
    nglob, nspec = sim.nglob, sim.nspec
    value = nglob * args.steps / wall
    # one force launch per colour; the events bracket all launches of a step
    launches = sim.colours(2)
    patch_mode = os.environ.get("SF2D_ASSEMBLY", "patch") == "patch"
    n_affine = sim.affine_elements() if patch_mode else 0
    chain = patch_mode and sim.chain_kernel()
    fk_ms = timing["forces_ms"] / args.steps / launches
    fk_bytes = force_kernel_bytes(nspec, nglob) / launches
    achieved = fk_bytes / (fk_ms * 1e-3) / 1e9
    traffic = read_traffic(args.workload)
    out = {
        "metric": "elastic GLL-pt updates/s",
        "value": value,
        "unit": "GLL-pt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian random-field elastic model)",
        "config": {
            "workload": args.workload,
            "nspec": nspec,
            "nglob": nglob,
            "receivers": cfg.nrec,
            "dt": cfg.dt,
            "boundaries": "Stacey bottom/left/right, free surface top",
            "parallelism": "1 GPU",
            "assembly": os.environ.get("SF2D_ASSEMBLY", "patch"),
            "affine_elements": n_affine,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_elastic_forces_chain" if chain else "k_elastic_forces_patch") if patch_mode
                      else "k_elastic_forces",
            "launches_per_step": launches,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": fk_bytes,
            "avg_launch_ms": fk_ms,
            # the bytes this kernel's own representation needs at minimum: affine elements carry one 16-byte
            # geometry record instead of 25 x 16 bytes (what it reads, not the formula above)
            "kernel_min_bytes_per_launch": kernel_min_bytes(nspec, nglob, n_affine, chain) / launches,
        },
        "gpu_loop_ms_events": timing["loop_ms"],
        "setup_s": t_setup,
    }
    sim.close()
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
    if not args.no_coupled:
        out["c2_coupled"] = c2_coupled()
    if not args.no_attenuation:
        out["c2_attenuation"] = c2_attenuation()
    if not args.no_gradient:
        out["fwi_gradient"] = fwi_gradient()
        out["fwi_gradient_batched"] = fwi_gradient_batched(args.batch_shots, nstep=args.fwi_steps)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
