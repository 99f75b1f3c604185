"""Benchmark of the MI355X smoke step on the 256^3 plume workload: learned
flexible PCG against plain (Jacobi) PCG.

Workload (the headline configuration): 256^3 cells, h = 1/256, closed no-slip
box, FIXED_VALUE 0 scalars, monotone-cubic advection, buoyancy 0.05 rho /
0.25 T (y up), vorticity confinement 2.0, CFL-adaptive dt (0.5 h / max|u|,
capped at 0.01), spherical source of radius 24 cells at (128, 32, 128)
injecting density 1 and temperature 1 +- 0.05 (seeded, regenerated each
step).  Every pressure solve reaches a true relative residual
||b - Ax|| / ||b|| <= 1e-6 (FP64 reductions and FP64 iterative refinement).

Solvers, each timed on its own context over the same steps:
  learned     flexible PCG preconditioned by a 5-level V(4,4)-cycle whose coarsest
              (16^3) level is solved by the trained network
              (weights/preconditioner_coarse16.bin) inside an exact deflation,
              then 8 Jacobi sweeps -- the headline `value`;
  jacobi      plain Jacobi PCG (the hand-written baseline solver);
  multilevel  the same flexible PCG with a 7-level V(3,3)-cycle and Jacobi sweeps
              on its 4^3 coarsest level (no network; the fastest classical
              cycle measured), reported beside them;
  smoothed    (on request) the learned solver's 5-level cycle with the
              network replaced by 16 Jacobi sweeps: what the network gains.

One "step" is one full smoke step (forces, advection, pressure solve to
tolerance, projection, scalar advection) with all inputs resident in HBM.
The default run times the workload's 200 steps after 3 warm-up steps for
each solver; each solve starts from the linear extrapolation of the two
previous pressures.  Prints one JSON line.
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
WEIGHTS = REPO / "visual-simulation-of-smoke_amd" / "weights" / "preconditioner_coarse16.bin"
# the grid-level network (64^3-crop training run), for the "network" solver
GRID_WEIGHTS = REPO / "visual-simulation-of-smoke_amd" / "weights" / "preconditioner_plume256_crop64.bin"
# damping of every V-cycle's Jacobi sweeps: 6/7 minimises damped Jacobi's smoothing factor on the 3D
# 7-point Laplacian (5/7 against 7/9 at 2/3); 150 steps of the 256^3 12-sphere scene, steps/s at
# 2/3, 0.75, 0.8, 6/7, 0.9, 0.95: learned V(4,4) 102.3, 104.3, 107.2, 112.2, 104.0, 92.0; classical
# V(3,3) 99.2, 107.7, 113.3, 116.5, 110.9, 97.6
OMEGA = 6.0 / 7.0
# damped-Jacobi sweeps after the network at the coarsest level of the learned solver (150 steps at
# 256^3 with OMEGA: 0, 4, 8, 16 sweeps gave 111.9, 112.3, 113.9, 106.0 steps/s on the 12-sphere scene)
LEARNED_COARSE_SWEEPS = 8
# pre- and post-sweeps: the learned cycle is fastest at V(4,4) (V(3,3): 112.5 against 113.9 steps/s),
# the classical one at V(3,3) (V(4,4): 110.4 against 116.4)
LEARNED_SWEEPS = 4
CLASSICAL_SWEEPS = 3
# algorithmic HBM bytes per cell
SWEEP_BYTES_PER_CELL = 13       # damped-Jacobi sweep: reads r, x (4+4), flags (1); writes x' (4)
RESTRICT_BYTES_PER_CELL = 9     # reads r, x, flags (+ 1/8 of a cell written)
SCALE_BYTES_PER_CELL = 9        # reads r, flags; writes x
PROLONG_BYTES_PER_CELL = 9      # reads x, flags (+ 1/8 of a cell); writes x
DIRECTION_BYTES_PER_CELL = 13   # PCG pass 1: reads r, p (4+4), flags (1); writes p (4)
CURVATURE_BYTES_PER_CELL = 5    # PCG pass 2: reads p (4), flags (1)
UPDATE_BYTES_PER_CELL = 21      # PCG pass 3: reads p, e, r (4+4+4), flags (1); writes e, r (4+4)
SWEEP_KERNEL = "mg_sweep4_kernel"
TIMING_STRIDE = 8               # kernel events on every 8th timed step
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


def configure(ctx: abi.SmokeContext, solver: str, net, warm_start: int, check_interval: int):
    if solver == "jacobi":
        solver_options(ctx, check_interval, warm_start)
    elif solver == "learned":
        # the network's coarse solve followed by damped-Jacobi sweeps at 16^3
        ctx.set_multilevel_options(levels=5, pre_sweeps=LEARNED_SWEEPS, post_sweeps=LEARNED_SWEEPS, omega=OMEGA,
                                   coarse_solver=abi.COARSE_NETWORK, coarse_sweeps=LEARNED_COARSE_SWEEPS)
        ctx.set_learned_preconditioner(net)
        solver_options(ctx, check_interval, warm_start, abi.PRECONDITIONER_MULTILEVEL)
    elif solver == "smoothed":
        # the learned solver's cycle with its network replaced by Jacobi sweeps (the ablation)
        ctx.set_multilevel_options(levels=5, pre_sweeps=LEARNED_SWEEPS, post_sweeps=LEARNED_SWEEPS, omega=OMEGA,
                                   coarse_solver=abi.COARSE_JACOBI, coarse_sweeps=16)
        solver_options(ctx, check_interval, warm_start, abi.PRECONDITIONER_MULTILEVEL)
    elif solver == "network":
        # the network as the preconditioner of the 256^3 grid itself (flexible PCG,
        # Jacobi-PCG fallback when the residual drops < 10 % over 5 iterations)
        ctx.set_learned_preconditioner(net)
        solver_options(ctx, check_interval, warm_start, abi.PRECONDITIONER_LEARNED)
    elif solver == "multilevel":
        ctx.set_multilevel_options(levels=7, pre_sweeps=CLASSICAL_SWEEPS, post_sweeps=CLASSICAL_SWEEPS, omega=OMEGA,
                                   coarse_solver=abi.COARSE_JACOBI, coarse_sweeps=8)
        solver_options(ctx, check_interval, warm_start, abi.PRECONDITIONER_MULTILEVEL)
    else:
        raise ValueError(solver)


def run_solver(solver: str, args, hip: abi.Hip, net, want_timing: bool, scene: str):
    n = args.size
    cfg = workload_config(n)
    if scene == "spheres":
        cfg.vorticity_confinement = 2.5
    src = plume(n)
    iters, resid, dts, fallbacks = [], [], [], 0
    out = {}
    with abi.SmokeContext(cfg, hip=hip) as ctx:
        if scene == "spheres":
            ctx.set_occupancy(spheres_occupancy(n))
        configure(ctx, solver, net, args.warm_start, args.check_interval)
        step_index = 0
        for _ in range(args.warmup):
            ctx.plume_sources(src, step_index)
            ctx.step()
            step_index += 1
        hip.synchronize()
        timed_kernels = want_timing and solver in ("jacobi", "learned")
        timing = None
        t0 = time.perf_counter()
        for k in range(args.steps):
            # kernel events on every TIMING_STRIDE-th timed step only: bracketing
            # every launch of every step would slow the timed steps by a few percent
            sample = timed_kernels and k % TIMING_STRIDE == 0
            if sample:
                ctx.set_kernel_timing(True)
            ctx.plume_sources(src, step_index)
            ctx.step()
            st = ctx.stats()          # stream-synchronising; reads the solve's true FP64 residual
            if sample:
                timing = add_timing(timing, ctx.kernel_timing())
                ctx.set_kernel_timing(False)
            iters.append(st.iterations)
            resid.append(st.relative_residual)
            dts.append(st.dt)
            fallbacks += st.learned_fallback
            step_index += 1
        hip.synchronize()
        elapsed = time.perf_counter() - t0
        div = ctx.view("FLOW_DIVERGENCE").astype(np.float64)
        fluid = ctx.view("OCCUPANCY") == 0
        out = {"steps_per_s": round(args.steps / elapsed, 4), "ms_per_step": round(1e3 * elapsed / args.steps, 3),
               "pcg_iterations_per_step": round(sum(iters) / len(iters), 2), "max_relative_residual": max(resid),
               "fallbacks": fallbacks, "divergence_l2_after_projection": float(np.sqrt(np.sum(div[fluid] ** 2))),
               "dt_range": [min(dts), max(dts)]}
        if timing is not None:
            out["kernel_timing"] = timing      # the sampled timed steps' kernels
        if want_timing and solver == "learned":
            out["multilevel_timing"] = ctx.time_multilevel(20)
        if solver != "jacobi":
            ctx.set_learned_preconditioner(None)
    return out


def add_timing(total, t):
    """Field-wise sum of two KernelTiming records (None: start from t)."""
    if total is None:
        return t
    for name, _ in abi.KernelTiming._fields_:
        setattr(total, name, getattr(total, name) + getattr(t, name))
    return total


def cpu_baseline(n: int, scene: str):
    """The CPU reference path on a bounded sample of the same workload: the
    FP32 host port of the step (the oracle source built with R = float, FP64
    reductions, OpenMP over the CPUs this process is given) runs step 0 of the
    scene untimed (zero velocity, zero right-hand side) and then one whole
    step, step 1 -- forces, advection, Jacobi-PCG to true rel-resid 1e-6 from
    the previous pressure, projection, scalars -- timed with steady_clock."""
    sys.path.insert(0, str(REPO / "oracle"))
    import cpu_baseline as cb
    r = cb.run("c5" if scene == "spheres" else "c2", steps=1, warmup=1)
    if n != 256:
        raise ValueError("the CPU sample is defined on the 256^3 workloads")
    return {
        "value": r["steps_per_s"], "unit": "steps/s", "cores": r["threads"], "kind": "port",
        "sample": (f"one whole step (step 1, after step 0 untimed) of the {n}^3 {'12-sphere' if scene == 'spheres' else 'plume'} "
                   f"workload on {r['threads']} host threads: {r['step_seconds'][0]:.1f} s with {r['pressure_iterations'][0]} "
                   f"Jacobi-PCG iterations to true rel-resid 1e-6 (no extrapolation; the GPU steps are later, warm-started ones)"),
        "mlups": round(r["mlups"], 3), "host": r["host"],
    }


def run_batch(count: int, steps: int, warmup: int, lanes: int = 16, modes=("fixed", "tolerance"), paths=("fused", "per_plume")):
    """The batched-plume workload (batch_plumes.plume_set(count, seed=3)):
    steps/s of the whole batch, fused (one launch per kernel over all plumes;
    fixed mode as one captured graph) against each plume's own step on
    `lanes` streams, in the reference's fixed-iteration mode (64 plain CG
    iterations) and to tolerance 1e-6 with Jacobi PCG (the plain method)."""
    import batch_plumes as bp
    plumes = bp.plume_set(count, seed=3)
    cells = bp.total_cells(plumes)
    out = {"plumes": count, "cells": cells, "grids": sorted({f"{c.nx}x{c.ny}x{c.nz}" for c, _ in plumes}), "steps": steps,
           "lanes": lanes}
    for mode in modes:
        for fused in [p == "fused" for p in paths]:
            with abi.SmokeBatch([c for c, _ in plumes], lanes=lanes) as batch:
                batch.set_fusion(fused)
                for p in batch.plumes:
                    if mode == "tolerance":
                        solver_options(p, 32, warm_start=2)
                srcs = [s for _, s in plumes]
                for k in range(warmup):
                    batch.step(srcs, k)
                batch.synchronize()
                iters = 0
                t0 = time.perf_counter()
                for k in range(warmup, warmup + steps):
                    batch.step(srcs, k)
                    if mode == "tolerance":
                        iters += sum(p.stats().iterations for p in batch.plumes)
                batch.synchronize()
                dt = time.perf_counter() - t0
                info = batch.info()
            key = f"{mode}_{'fused' if fused else 'per_plume'}"
            out[key] = {"batch_steps_per_s": round(steps / dt, 3), "mcells_per_s": round(cells * steps / dt / 1e6, 1)}
            if mode == "tolerance":
                out[key]["pcg_iterations_per_plume_step"] = round(iters / (steps * count), 1)
            if fused:
                out[key]["launches_per_step"] = int(info.launches_per_step)
                out[key]["launches_per_cg_iteration"] = int(info.launches_per_iteration)
                out[key]["graph"] = bool(info.graph)
            print(f"batch {key}: {out[key]}", file=sys.stderr, flush=True)
        if f"{mode}_fused" in out and f"{mode}_per_plume" in out:
            out[f"{mode}_fusion_speedup"] = round(out[f"{mode}_fused"]["batch_steps_per_s"] / out[f"{mode}_per_plume"]["batch_steps_per_s"], 3)
    return out


def load_traffic(name: str):
    f = REPO / "profiles" / name
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
    ap.add_argument("--warm-start", type=int, default=2, choices=(0, 1, 2),
                    help="initial guess of each solve: zero, previous pressure, extrapolation of the last two")
    ap.add_argument("--scene", choices=("plume", "spheres"), default="spheres",
                    help="spheres (default, the headline): the closed box with 12 seeded solid spheres and confinement 2.5; "
                         "plume: the closed box alone")
    ap.add_argument("--extra-scenes", default="plume",
                    help="comma-separated scenes also timed (same solvers, no kernel events), reported under 'scenes'")
    ap.add_argument("--solvers", default="learned,jacobi,multilevel",
                    help="comma-separated, from learned, jacobi, multilevel, smoothed; the first one is the reported value")
    ap.add_argument("--weights", default=str(WEIGHTS))
    ap.add_argument("--batch-plumes", type=int, default=0,
                    help="also time the batched-plume workload with this many plumes (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--batch-lanes", type=int, default=16)
    ap.add_argument("--batch-modes", default="fixed,tolerance", help="comma-separated, from fixed, tolerance")
    ap.add_argument("--batch-paths", default="fused,per_plume",
                    help="comma-separated, from fused (one launch per kernel over all plumes) and per_plume (each plume's own step)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is design-only (see DESIGN.md)", file=sys.stderr)
        sys.exit(1)

    n = args.size
    cells = n ** 3
    hip = abi.Hip()
    copy_gbps, _ = abi.measure_copy_bandwidth(1 << 30, 20)
    solvers = [s.strip() for s in args.solvers.split(",") if s.strip()]
    extra_scenes = [s.strip() for s in args.extra_scenes.split(",") if s.strip() and s.strip() != args.scene]
    nets = {}
    if "learned" in solvers:
        nets["learned"] = abi.LearnedPreconditioner(seed=0, path=args.weights)
    if "network" in solvers:
        nets["network"] = abi.LearnedPreconditioner(seed=0, path=GRID_WEIGHTS)
    results = {}
    for solver in solvers:
        results[solver] = run_solver(solver, args, hip, nets.get(solver), not args.no_timing, args.scene)
        print(f"bench {args.scene} {solver}: {results[solver]['steps_per_s']} steps/s", file=sys.stderr, flush=True)
    scenes = {}
    for scene in extra_scenes:
        scenes[scene] = {}
        for solver in solvers:
            scenes[scene][solver] = run_solver(solver, args, hip, nets.get(solver), False, scene)
            print(f"bench {scene} {solver}: {scenes[scene][solver]['steps_per_s']} steps/s", file=sys.stderr, flush=True)
        if "learned" in scenes[scene] and "jacobi" in scenes[scene]:
            scenes[scene]["learned_vs_plain_pcg_speedup"] = round(
                scenes[scene]["learned"]["steps_per_s"] / scenes[scene]["jacobi"]["steps_per_s"], 3)
    for network in nets.values():
        network.close()
    head = solvers[0]
    value = results[head]["steps_per_s"]

    roofline = None
    lt = results.get("learned", {}).pop("multilevel_timing", None)
    lk = results.get("learned", {}).pop("kernel_timing", None)
    if lt is not None and lk is not None and lk.mg_sweeps > 0:
        def gb(bytes_per_cell, us):
            return bytes_per_cell * cells / (us * 1e-6) / 1e9
        sweep_us = 1e3 * lk.mg_sweep_ms / lk.mg_sweeps
        achieved = gb(SWEEP_BYTES_PER_CELL, sweep_us)
        roofline = {
            "bound": "hbm", "kernel": SWEEP_KERNEL + " (damped-Jacobi sweep of the V-cycle on the 256^3 pressure grid, float4 rows)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": load_traffic("mg_sweep_256.traffic.json"),
            "traffic_source": ("profiles/mg_sweep_256.traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE of this kernel at 256^3, "
                               "scaled by the counters' reading for a copy kernel of known size; read from the file, not measured "
                               "in this run"),
            "algorithmic_bytes_per_launch": SWEEP_BYTES_PER_CELL * cells,
            "avg_launch_ms": round(sweep_us * 1e-3, 5), "launches_timed": int(lk.mg_sweeps),
            "timed": ("every sweep of the pressure grid in the V-cycles of every 8th timed step (7 per cycle: the 3 "
                      "pre-sweeps after x = omega D^-1 r and the 4 post-sweeps, the last of which also sums r.z); HIP "
                      "events on the context stream around each run of back-to-back sweeps, divided by its sweeps"),
            "isolated_launch_ms": round(lt.sweep_us * 1e-3, 5),
            "vcycle_kernels": {
                "restrict": {"avg_launch_ms": round(lt.restrict_us * 1e-3, 5), "achieved_gbps": round(gb(RESTRICT_BYTES_PER_CELL, lt.restrict_us), 1)},
                "first_sweep": {"avg_launch_ms": round(lt.scale_us * 1e-3, 5), "achieved_gbps": round(gb(SCALE_BYTES_PER_CELL, lt.scale_us), 1)},
                "prolong": {"avg_launch_ms": round(lt.prolong_us * 1e-3, 5), "achieved_gbps": round(gb(PROLONG_BYTES_PER_CELL, lt.prolong_us), 1)},
                "whole_vcycle_ms": round(lt.vcycle_us * 1e-3, 5),
            },
            "stream_copy_gbps": round(copy_gbps, 1),
        }
    timing = results.get("jacobi", {}).pop("kernel_timing", None)
    if timing is not None and timing.pcg_timed_iterations > 0:
        k = timing.pcg_timed_iterations
        ms = {"direction": timing.pcg_direction_ms / k, "curvature": timing.pcg_curvature_ms / k, "update": timing.pcg_update_ms / k}
        per_cell = {"direction": DIRECTION_BYTES_PER_CELL, "curvature": CURVATURE_BYTES_PER_CELL, "update": UPDATE_BYTES_PER_CELL}
        gbps = {name: per_cell[name] * cells / (ms[name] * 1e-3) / 1e9 for name in ms}
        reduction_ms = timing.pcg_reduction_ms / k
        iteration_ms = sum(ms.values()) + reduction_ms
        pcg = {
            "kernel": UPDATE_KERNEL + " (plain PCG pass 3: e += alpha p, r -= alpha Ap, r.z, r.r)",
            "achieved": round(gbps["update"], 1), "frac": round(gbps["update"] / HBM_PEAK_GBPS, 4),
            "traffic": load_traffic("pcg_update_256.traffic.json"),
            "traffic_source": "profiles/pcg_update_256.traffic.json (rocprofv3 counters, read from the file)",
            "algorithmic_bytes_per_launch": UPDATE_BYTES_PER_CELL * cells,
            "avg_launch_ms": round(ms["update"], 5), "launches_timed": k,
            "passes": {name: {"avg_launch_ms": round(ms[name], 5), "achieved_gbps": round(gbps[name], 1)} for name in ms},
            "finalize_kernels_ms": round(reduction_ms, 5),
            "pcg_iteration": {"ms": round(iteration_ms, 5),
                              "achieved_gbps": round(sum(per_cell.values()) * cells / (iteration_ms * 1e-3) / 1e9, 1)},
        }
        if roofline is None:
            roofline = {"bound": "hbm", "peak": HBM_PEAK_GBPS, "unit": "GB/s", "stream_copy_gbps": round(copy_gbps, 1), **pcg}
        else:
            roofline["plain_pcg_update_pass"] = pcg

    speedup = None
    if "learned" in results and "jacobi" in results:
        speedup = round(results["learned"]["steps_per_s"] / results["jacobi"]["steps_per_s"], 3)
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": results[head]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"{n}^3 plume" + (" with 12 seeded solid spheres (radius U(8,24)), confinement 2.5" if args.scene == "spheres" else "") +
                                ", every pressure solve to true rel-resid 1e-6 (FP64 reductions + refinement), cubic advection, "
                                f"adaptive dt; value = {head}"),
                   "grid": [n, n, n], "scene": args.scene, "relative_tolerance": 1e-6,
                   "warm_start": ["zero", "previous", "extrapolated"][args.warm_start],
                   "vcycle_omega": round(OMEGA, 6),
                   "solvers": {"learned": ("FPCG + 5-level V(4,4)-cycle, network coarse solver at 16^3 (" + Path(args.weights).name + f") + "
                                           f"{LEARNED_COARSE_SWEEPS} Jacobi sweeps"),
                               "jacobi": "plain Jacobi PCG", "multilevel": "FPCG + 7-level V(3,3)-cycle, Jacobi coarse solver (no network)",
                               "smoothed": "the learned solver's 5-level cycle, network replaced by 16 Jacobi sweeps",
                               "network": "FPCG preconditioned by the network on the 256^3 grid itself (" + GRID_WEIGHTS.name + ")"}},
        "solvers": results, "learned_vs_plain_pcg_speedup": speedup,
        "scenes": scenes,
        "roofline": roofline, "cpu_baseline": None,
    }
    if args.batch_plumes > 0:
        modes = tuple(m.strip() for m in args.batch_modes.split(",") if m.strip())
        paths = tuple(p.strip() for p in args.batch_paths.split(",") if p.strip())
        out["batched_plumes"] = run_batch(args.batch_plumes, args.batch_steps, 2, args.batch_lanes, modes, paths)
    if not args.no_cpu_baseline and n == 256:
        print("bench: CPU baseline (one whole 256^3 step on the host cores)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, args.scene)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
