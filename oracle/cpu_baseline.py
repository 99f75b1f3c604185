"""The CPU reference path, timed: whole steps of the FP32 host port of the
step (the oracle source built with R = float and FP64 reductions, OpenMP) on
the host's cores.  TEST/BASELINE INFRASTRUCTURE -- it measures the CPU path the
GPU library is compared with; the product never imports it.

Workloads (same scenes as the GPU runs; Jacobi PCG to true rel-resid 1e-6,
monotone-cubic advection, CFL-adaptive dt capped at 0.01):
  c1       64^3 plume, source radius 6 at (32, 8, 32), confinement 2.0: 100 steps
  c2       256^3 plume, source radius 24 at (128, 32, 128), confinement 2.0
  c5       256^3 with 12 seeded solid spheres, confinement 2.5

`python oracle/cpu_baseline.py [c1] [c2] [c5] [--steps K] [--warmup W]`
prints one JSON line per workload (steps/s, MLUPS = cells * steps / us as in
benchmark.cu:213, per-step seconds and pressure iterations, threads used).
Steps 0..W-1 run untimed first; step 0 of these scenes has a zero velocity
field and a zero right-hand side.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from pathlib import Path

ORACLE_DIR = Path(__file__).resolve().parent
REPO = ORACLE_DIR.parent
sys.path.insert(0, str(ORACLE_DIR))
sys.path.insert(0, str(REPO))

import oracle as orc  # noqa: E402


def host_threads() -> dict:
    """The CPUs this process may use: OMP_NUM_THREADS when the host sets it
    (the GPU hosts grant a fixed share of their CPUs per GPU), else every CPU of
    the process's affinity mask; os.cpu_count() is reported beside it."""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else affinity
    return {"threads": min(threads, affinity), "affinity_cpus": affinity, "os_cpu_count": os.cpu_count(),
            "omp_num_threads_env": env or None}


def workload(name: str):
    import bench
    if name == "c1":
        n = 64
        cfg = bench.workload_config(n)
        import smoke_abi as abi
        src = abi.PlumeSource(32.0, 8.0, 32.0, 6.0, 1.0, 1.0, 0.05, 1)
        return cfg, src, None
    n = 256
    cfg = bench.workload_config(n)
    if name == "c5":
        cfg.vorticity_confinement = 2.5
        return cfg, bench.plume(n), bench.spheres_occupancy(n)
    return cfg, bench.plume(n), None


def options():
    return orc.options(mode=1, preconditioner=1, tol=1e-6, max_iterations=50000, max_refinements=4, warm_start=1,
                       adaptive_dt=1, cfl=0.5, dt_max=0.01, inner_reduction=1e-4)


def run(name: str, steps: int, warmup: int, threads: int | None = None) -> dict:
    cfg, src, occ = workload(name)
    th = host_threads()
    t = threads or th["threads"]
    secs, its = orc.f32_steps(cfg, options(), src, occ, warmup, steps, t)
    cells = cfg.nx * cfg.ny * cfg.nz
    total = sum(secs)
    return {"workload": name, "grid": [cfg.nx, cfg.ny, cfg.nz], "obstacles": occ is not None, "warmup_steps": warmup,
            "timed_steps": steps, "steps_per_s": steps / total, "mlups": cells * steps / (total * 1e6),
            "step_seconds": [round(s, 4) for s in secs], "pressure_iterations": its, "threads": t, "host": th,
            "kind": "port (FP32 host transcription of the step, FP64 reductions, OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=["c1", "c2", "c5"])
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 100 for c1, 5 for c2 and c5)")
    ap.add_argument("--warmup", type=int, default=1)
    args = ap.parse_args()
    os.environ.setdefault("SMOKE_ORACLE_PROGRESS", "1")
    for name in args.workloads:
        steps = args.steps or (100 if name == "c1" else 5)
        print(f"cpu baseline {name}: {steps} timed steps after {args.warmup} warm-up", file=sys.stderr, flush=True)
        print(json.dumps(run(name, steps, args.warmup)), flush=True)


if __name__ == "__main__":
    main()
