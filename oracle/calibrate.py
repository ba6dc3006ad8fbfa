"""Calibration of the CPU baseline against the reference solver itself (test infrastructure: it times the oracle).

Times the oracle's FP32 builds on the 80 x 60 plain elastic configuration (1600 steps, 77,361 GLL points) -- the
serial parity build (gcc -O2) and the OpenMP baseline build (gcc -O3, per-colour parallel elements) on 1 and on all
cores -- and prints them next to the reference's own time-loop wall times for the same configuration on the same
8-core Xeon @ 2.70 GHz (xspecfem2D built with amdflang -O3, FP32: 6.097 s serial, 1.549 s with 8 MPI ranks), so that
a baseline timed on another host relates to the reference.  Run on the 8-core build host:

    python oracle/calibrate.py > profiles/cpu_calibration_c1.md
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

REFERENCE_SERIAL_S = 6.097
REFERENCE_MPI8_S = 1.549


def _time(precision: str, threads: int) -> float:
    code = (f"import sys, time; sys.path.insert(0, {ROOT!r}); from oracle.oracle import OracleSim; "
            "from specfem2d_amd import config; cfg = config.plain_elastic_c1(); "
            f"o = OracleSim(cfg, {precision!r}); t = time.perf_counter(); o.forward(); "
            "print(time.perf_counter() - t)")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", code], env=env, check=True, capture_output=True, text=True)
    return float(out.stdout.strip().splitlines()[-1])


def main():
    cores = os.cpu_count() or 1
    best = lambda p, t: min(_time(p, t) for _ in range(3))  # noqa: E731
    serial = best("sp", 1)
    omp1 = best("sp_omp", 1)
    ompn = best("sp_omp", cores)
    nglob, nstep = 77361, 1600
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next(line.split(":", 1)[1].strip() for line in f if line.startswith("model name"))
    except (OSError, StopIteration):
        pass
    print("# CPU baseline calibration: oracle vs the reference solver, 80x60 plain elastic, 1600 steps\n")
    print(f"Host: {cpu}, {cores} cores.  Command: `python oracle/calibrate.py` (best of 3 runs each).\n")
    print("| Run | Time-loop wall | GLL-pt updates/s | Reference / run |")
    print("|---|---|---|---|")
    rows = [("reference xspecfem2D, serial (amdflang -O3)", REFERENCE_SERIAL_S),
            ("reference xspecfem2D, 8 MPI ranks", REFERENCE_MPI8_S),
            ("oracle serial parity build (gcc -O2), 1 thread", serial),
            ("oracle OpenMP baseline build (gcc -O3), 1 thread", omp1),
            (f"oracle OpenMP baseline build (gcc -O3), {cores} threads", ompn)]
    for name, t in rows:
        print(f"| {name} | {t:.3f} s | {nglob * nstep / t / 1e6:.1f} M | {REFERENCE_SERIAL_S / t:.3f} |")
    print()
    print(f"calibration_1core: {REFERENCE_SERIAL_S / omp1:.4f}")
    print(f"calibration_allcores_vs_mpi8: {REFERENCE_MPI8_S / ompn:.4f}")


if __name__ == "__main__":
    main()
