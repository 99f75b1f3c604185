"""Benchmark of the MI355X smoke step on the 256^3 plume workload.

Workload (the headline configuration): 256^3 cells, h = 1/256, closed no-slip
box, FIXED_VALUE 0 scalars, monotone-cubic advection, buoyancy 0.05 rho /
0.25 T (y up), vorticity confinement 2.0, CFL-adaptive dt (0.5 h / max|u|,
capped at 0.01), spherical source of radius 24 cells at (128, 32, 128)
injecting density 1 and temperature 1 +- 0.05 (seeded, regenerated each
step).  Every pressure solve is Jacobi-PCG with FP64 reductions and FP64
iterative refinement to a true relative residual ||b - Ax|| / ||b|| <= 1e-6.

One "step" is one full smoke step (forces, advection, pressure solve to
tolerance, projection, scalar advection) with all inputs resident in HBM.
The default run times 200 steps after 3 warm-up steps (the workload's length);
each solve starts from the linear extrapolation of the two previous pressures.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "visual-simulation-of-smoke_amd"))

import numpy as np  # noqa: E402

import smoke_abi as abi  # noqa: E402

METRIC = "steps/s @256³ to rel-resid 1e-6: learned-FPCG vs plain PCG; HBM GB/s"
HBM_PEAK_GBPS = 8000.0
# algorithmic HBM bytes per cell of the three passes of one PCG iteration
DIRECTION_BYTES_PER_CELL = 13   # pass 1: reads r, p (4+4), flags (1); writes p (4)
CURVATURE_BYTES_PER_CELL = 5    # pass 2: reads p (4), flags (1)
UPDATE_BYTES_PER_CELL = 21      # pass 3: reads p, e, r (4+4+4), flags (1); writes e, r (4+4)
UPDATE_KERNEL = "pcg_stencil_kernel<64, 8, 2, 4, 1, true>"


def workload_config(n: int) -> abi.Config:
    return abi.make_config(n, n, n, cell_size=1.0 / n, dt=0.01, pressure_iterations=64, ambient=0.0,
                           buoyancy_density=0.05, buoyancy_temperature=0.25, vorticity=2.0, cubic=True)


def plume(n: int) -> abi.PlumeSource:
    s = n / 256.0
    return abi.PlumeSource(128.0 * s, 32.0 * s, 128.0 * s, 24.0 * s, 1.0, 1.0, 0.05, 2)


def solver_options(ctx: abi.SmokeContext, check_interval: int, warm_start: int = 2, preconditioner: int = abi.PRECONDITIONER_JACOBI):
    ctx.set_solver_options(pressure_mode=abi.PRESSURE_TOLERANCE, preconditioner=preconditioner,
                           relative_tolerance=1e-6, max_iterations=50000, max_refinements=4, warm_start=warm_start,
                           adaptive_dt=1, cfl=0.5, dt_max=0.01, check_interval=check_interval)


def spheres_occupancy(n: int, count: int = 12, seed: int = 5) -> np.ndarray:
    """The obstacle scene: `count` seeded solid spheres of radius U(8, 24) *
    n / 256 cells, centred above the source (y in [0.3 n, 0.9 n])."""
    rng = np.random.default_rng(seed)
    s = n / 256.0
    z, y, x = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    occ = np.zeros((n, n, n), np.uint8)
    for _ in range(count):
        r = rng.uniform(8.0, 24.0) * s
        cx, cz = rng.uniform(r, n - r, size=2)
        cy = rng.uniform(0.3 * n, 0.9 * n - r)
        occ[(x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r] = 1
    return occ


def cpu_baseline(n: int, iterations_per_step: float, threads: int):
    """Bounded CPU sample of the same workload with the FP32 host port of the
    step (the oracle source built with R = float, FP64 reductions, OpenMP):
    one step's non-solve stages timed once plus 12 CG iterations timed, scaled
    to the GPU run's mean PCG iterations per step."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as orc
    cfg = workload_config(n)
    opts = orc.options(mode=1, preconditioner=1, tol=1e-6, max_refinements=4, warm_start=1, adaptive_dt=1, cfl=0.5,
                       dt_max=0.01, inner_reduction=1e-4)
    t0 = time.time()
    other, per_iter = orc.f32_step_sample(cfg, opts, plume(n), 12, threads)
    wall = time.time() - t0
    seconds_per_step = other + iterations_per_step * per_iter
    return {
        "value": 1.0 / seconds_per_step, "unit": "steps/s", "cores": threads, "kind": "port",
        "sample": (f"{n}^3 step from rest on {threads} host threads: non-solve stages {other:.2f} s once + 12 CG "
                   f"iterations at {per_iter * 1e3:.1f} ms each, scaled to {iterations_per_step:.0f} iterations/step "
                   f"(the GPU run's mean); {wall:.1f} s of CPU work"),
    }


def load_traffic():
    f = REPO / "profiles" / "pcg_update_256.traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--check-interval", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel event timing")
    ap.add_argument("--autotune", action="store_true", help="time candidate launch shapes on this grid first (after warm-up)")
    ap.add_argument("--warm-start", type=int, default=2, choices=(0, 1, 2),
                    help="initial guess of each solve: zero, previous pressure, extrapolation of the last two")
    ap.add_argument("--scene", choices=("plume", "spheres"), default="plume",
                    help="plume: the closed box; spheres: the same with 12 seeded solid spheres and confinement 2.5")
    ap.add_argument("--solver", choices=("jacobi", "learned"), default="jacobi",
                    help="Jacobi PCG, or flexible PCG with the learned preconditioner (Jacobi fallback)")
    ap.add_argument("--weights", default=str(REPO / "visual-simulation-of-smoke_amd" / "weights" / "preconditioner_plume256_crop64.bin"))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is design-only (see DESIGN.md)", file=sys.stderr)
        sys.exit(1)

    n = args.size
    hip = abi.Hip()
    copy_gbps, copy_ms = abi.measure_copy_bandwidth(1 << 30, 20)
    cfg = workload_config(n)
    if args.scene == "spheres":
        cfg.vorticity_confinement = 2.5
    src = plume(n)
    iters, resid, dts, learned_its, fallbacks = [], [], [], [], 0
    net = abi.LearnedPreconditioner(seed=0, path=args.weights) if args.solver == "learned" else None
    with abi.SmokeContext(cfg, hip=hip) as ctx:
        if args.scene == "spheres":
            ctx.set_occupancy(spheres_occupancy(n))
        if net is not None:
            ctx.set_learned_preconditioner(net)
            solver_options(ctx, 1, args.warm_start, abi.PRECONDITIONER_LEARNED)
        else:
            solver_options(ctx, args.check_interval, args.warm_start)
        step_index = 0
        for _ in range(args.warmup):
            ctx.plume_sources(src, step_index)
            ctx.step()
            step_index += 1
        hip.synchronize()
        launches = None
        if args.autotune:
            rep = ctx.autotune_launches()
            launches = {"pcg_curvature": rep.pcg_curvature.tuple(), "pcg_update": rep.pcg_update.tuple(),
                        "advect": [rep.advect[k].tuple() for k in range(4)],
                        "pcg_curvature_us": list(rep.pcg_curvature_us), "pcg_update_us": list(rep.pcg_update_us),
                        "advect_us": [list(rep.advect_us[k]) for k in range(4)], "candidates_timed": rep.candidates_timed}
        ctx.set_kernel_timing(not args.no_timing)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.plume_sources(src, step_index)
            ctx.step()
            st = ctx.stats()          # stream-synchronising; reads the solve's true FP64 residual
            iters.append(st.iterations)
            resid.append(st.relative_residual)
            dts.append(st.dt)
            learned_its.append(st.learned_iterations)
            fallbacks += st.learned_fallback
            step_index += 1
        hip.synchronize()
        elapsed = time.perf_counter() - t0
        timing = ctx.kernel_timing()
        div = ctx.view("FLOW_DIVERGENCE").astype(np.float64)
        fluid = ctx.view("OCCUPANCY") == 0
        divergence_l2 = float(np.sqrt(np.sum(div[fluid] ** 2)))
        if net is not None:
            ctx.set_learned_preconditioner(None)
    if net is not None:
        net.close()

    steps_per_s = args.steps / elapsed
    cells = n ** 3
    mean_iters = sum(iters) / len(iters)
    roofline = None
    if timing.pcg_timed_iterations > 0:
        k = timing.pcg_timed_iterations
        ms = {"direction": timing.pcg_direction_ms / k, "curvature": timing.pcg_curvature_ms / k,
              "update": timing.pcg_update_ms / k}
        per_cell = {"direction": DIRECTION_BYTES_PER_CELL, "curvature": CURVATURE_BYTES_PER_CELL,
                    "update": UPDATE_BYTES_PER_CELL}

        def gbps(name):
            return per_cell[name] * cells / (ms[name] * 1e-3) / 1e9

        achieved = gbps("update")
        reduction_ms = timing.pcg_reduction_ms / k
        iteration_ms = sum(ms.values()) + reduction_ms
        roofline = {
            "bound": "hbm", "kernel": UPDATE_KERNEL + " (PCG pass 3: e += alpha p, r -= alpha Ap, r.z, r.r, beta)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": load_traffic(), "algorithmic_bytes_per_launch": UPDATE_BYTES_PER_CELL * cells,
            "avg_launch_ms": round(ms["update"], 5), "launches_timed": k,
            "passes": {name: {"avg_launch_ms": round(ms[name], 5), "achieved_gbps": round(gbps(name), 1),
                              "algorithmic_bytes_per_launch": per_cell[name] * cells} for name in ms},
            "finalize_kernels_ms": round(reduction_ms, 5),
            "pcg_iteration": {"ms": round(iteration_ms, 5),
                              "achieved_gbps": round(sum(per_cell.values()) * cells / (iteration_ms * 1e-3) / 1e9, 1)},
            "stream_copy_gbps": round(copy_gbps, 1),
        }
    out = {
        "metric": METRIC, "value": round(steps_per_s, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"{n}^3 plume" + (" with 12 solid spheres" if args.scene == "spheres" else "") +
                                (", learned-preconditioner flexible PCG (Jacobi fallback)" if net is not None else ", plain Jacobi-PCG") +
                                " (FP64 reductions + refinement) to true rel-resid 1e-6, cubic advection, adaptive dt"),
                   "grid": [n, n, n], "scene": args.scene, "solver": "learned_fpcg" if net is not None else "jacobi_pcg",
                   "relative_tolerance": 1e-6, "warm_start": ["zero", "previous", "extrapolated"][args.warm_start],
                   "learned_fpcg": ({"weights": Path(args.weights).name} if net is not None else None)},
        "pcg_iterations_per_step": round(mean_iters, 1), "max_relative_residual": max(resid),
        "learned_iterations_per_step": round(sum(learned_its) / len(learned_its), 1), "learned_fallbacks": fallbacks,
        "divergence_l2_after_projection": divergence_l2,
        "dt_range": [min(dts), max(dts)], "roofline": roofline, "cpu_baseline": None,
    }
    if launches is not None:
        out["launches"] = launches
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(n, mean_iters, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
