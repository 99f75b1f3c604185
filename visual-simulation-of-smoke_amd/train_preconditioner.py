"""Trains the learned pressure preconditioner on pressure solves of the plume
workload, with the library's own kernels (BF16 matrix-core convolutions,
FP32 Adam) on one GPU.

1. runs the plume workload (`--grid`^3, the 256^3 benchmark plume scaled) for
   `--steps` steps with Jacobi-PCG to a true residual of 1e-6 and appends
   `--crops-per-step` random crops of `--crop`^3 cells of every solve's
   right-hand side and operator flags to a SMOKETD1 file
   (smoke_simulation_ext_write_training_crops);
2. trains the network (Adam, `--batch` crops per step, learning rate `--lr`)
   for `--epochs` passes over the crops or until `--time-limit` seconds,
   printing the loss ||A z - r||^2 / ||r||^2 as it goes;
3. saves the weights (SMOKENN1) to `--out`.

With `--device-crops` the crops never leave the GPU (the configuration the
north star's training run is specified on: the 256^3 plume's solves cut into
64^3 crops, batch 16, 30 epochs over `--samples` crops, Adam lr 1e-3): every
solve's crops are copied into one device array
(smoke_simulation_ext_copy_training_crops), each batch is gathered from it
by device-to-device copies in a per-epoch shuffled order, the last
`--validation` crops (from the latest solves) are held out and scored after
every epoch, and a JSON summary line ends the run.

With `--coarse` the network is trained as the coarse-grid solver of the
multilevel preconditioner instead: the plume (and, with `--spheres`, the
obstacle scene) runs with flexible PCG preconditioned by the V-cycle whose
coarsest level is smoothed only, every V-cycle's coarsest right-hand side is
captured on the device (smoke_simulation_ext_set_multilevel_capture), and the
network learns z = A_c^-1 r_c on those whole coarse volumes with the energy
loss (z^T A z / 2 - z^T r) / ||r||^2.

Usage: python train_preconditioner.py --out weights.bin [--grid 256 --steps 40 ...]
       python train_preconditioner.py --coarse --out coarse.bin [--grid 256 --levels 5 --steps 30 --spheres]
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import smoke_abi as abi  # noqa: E402
import training_data as td  # noqa: E402


def plume_config(n: int):
    cfg = abi.make_config(n, n, n, cell_size=1.0 / n, dt=0.01, pressure_iterations=64, ambient=0.0, buoyancy_density=0.05,
                          buoyancy_temperature=0.25, vorticity=2.0, cubic=True)
    s = n / 256.0
    src = abi.PlumeSource(128.0 * s, 32.0 * s, 128.0 * s, 24.0 * s, 1.0, 1.0, 0.05, 2)
    return cfg, src


def generate(path: Path, n: int, steps: int, crop: int, per_step: int, seed: int = 0, log=print):
    """Runs the plume and logs crops of every solve's right-hand side."""
    cfg, src = plume_config(n)
    rng = np.random.default_rng(seed)
    if path.exists():
        path.unlink()
    manifest = Path(str(path) + ".jsonl")
    if manifest.exists():
        manifest.unlink()
    with abi.SmokeContext(cfg) as ctx:
        ctx.set_solver_options(pressure_mode=abi.PRESSURE_TOLERANCE, preconditioner=abi.PRECONDITIONER_JACOBI, relative_tolerance=1e-6,
                               max_iterations=50000, max_refinements=4, warm_start=2, adaptive_dt=1, cfl=0.5, dt_max=0.01,
                               check_interval=32)
        for k in range(steps):
            ctx.plume_sources(src, k)
            ctx.step()
            if k < 2:
                continue   # the first solves have a zero or tiny right-hand side
            origins = rng.integers(0, n - crop + 1, size=(per_step, 3))
            ctx.write_training_crops(path, origins, (crop, crop, crop), step=k)
            if k % 10 == 0:
                log(f"data: step {k}, {ctx.stats().iterations} PCG iterations", flush=True)


def spheres_occupancy(n: int, count: int = 12, seed: int = 5) -> np.ndarray:
    """`count` seeded solid spheres of radius U(8, 24) * n / 256 cells,
    centred above the plume source (y in [0.3 n, 0.9 n])."""
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


def generate_coarse(n: int, levels: int, steps: int, capacity: int, scenes=("plume",), coarse_sweeps: int = 16, sweeps: int = 2,
                    omega: float = 2.0 / 3.0, log=print):
    """Coarsest-level right-hand sides of the V-cycles of `steps` multilevel
    solves per scene; returns {"r": (count, D, H, W) float32, "flags": uint8,
    "crop": (W, H, D)}."""
    hip = abi.Hip()
    rs, fs, dims = [], [], None
    for scene in scenes:
        cfg, src = plume_config(n)
        if scene == "spheres":
            cfg.vorticity_confinement = 2.5
        with abi.SmokeContext(cfg, hip=hip) as ctx:
            if scene == "spheres":
                ctx.set_occupancy(spheres_occupancy(n))
            ctx.set_multilevel_options(levels=levels, pre_sweeps=sweeps, post_sweeps=sweeps, omega=omega, coarse_solver=abi.COARSE_JACOBI,
                                       coarse_sweeps=coarse_sweeps)
            ctx.set_solver_options(pressure_mode=abi.PRESSURE_TOLERANCE, preconditioner=abi.PRECONDITIONER_MULTILEVEL,
                                   relative_tolerance=1e-6, max_iterations=50000, max_refinements=4, warm_start=2, adaptive_dt=1,
                                   cfl=0.5, dt_max=0.01, check_interval=16)
            level_dims, _ = ctx.multilevel_info()
            dims = level_dims[-1]
            cells = dims[0] * dims[1] * dims[2]
            buf = hip.malloc(4 * cells * capacity)
            try:
                ctx.set_multilevel_capture(buf, capacity)
                for k in range(steps):
                    ctx.plume_sources(src, k)
                    ctx.step()
                    if k % 10 == 0:
                        log(f"data: {scene} step {k}, {ctx.stats().iterations} iterations, "
                            f"{ctx.multilevel_info()[1]} coarse volumes", flush=True)
                ctx.synchronize()
                count = ctx.multilevel_info()[1]
                ctx.set_multilevel_capture(None, 0)
                flags = ctx.level_flags(len(level_dims) - 1)
                rs.append(hip.download(buf, (count, dims[2], dims[1], dims[0])))
                fs.append(np.broadcast_to(flags, (count, *flags.shape)))
            finally:
                hip.free(buf)
    return {"r": np.concatenate(rs), "flags": np.ascontiguousarray(np.concatenate(fs)), "crop": dims}


def generate_device_crops(hip: abi.Hip, n: int, samples: int, crop: int, per_step: int, seed: int = 0, log=print):
    """Runs the n^3 plume with Jacobi PCG to a true residual of 1e-6 and keeps
    `per_step` random crop^3 crops of every solve's unit-RMS right-hand side
    and flags on the device until `samples` crops exist (the first two solves
    are skipped: zero or tiny right-hand sides).  Returns (d_r, d_flags,
    steps run); the caller frees the two buffers."""
    cfg, src = plume_config(n)
    cells = crop ** 3
    d_r = hip.malloc(4 * cells * samples)
    d_f = hip.malloc(cells * samples)
    rng = np.random.default_rng(seed)
    have, k = 0, 0
    t0 = time.time()
    with abi.SmokeContext(cfg, hip=hip) as ctx:
        ctx.set_solver_options(pressure_mode=abi.PRESSURE_TOLERANCE, preconditioner=abi.PRECONDITIONER_JACOBI, relative_tolerance=1e-6,
                               max_iterations=50000, max_refinements=4, warm_start=2, adaptive_dt=1, cfl=0.5, dt_max=0.01,
                               check_interval=32)
        while have < samples:
            ctx.plume_sources(src, k)
            ctx.step()
            if k >= 2:
                m = min(per_step, samples - have)
                origins = rng.integers(0, n - crop + 1, size=(m, 3))
                ctx.copy_training_crops(origins, (crop, crop, crop), d_r + 4 * cells * have, d_f + cells * have)
                have += m
            if k % 20 == 0:
                log(f"data: step {k}, {ctx.stats().iterations} PCG iterations, {have} crops ({time.time() - t0:.0f} s)", flush=True)
            k += 1
        ctx.synchronize()
    return d_r, d_f, k


def train_device(net: abi.LearnedPreconditioner, hip: abi.Hip, d_r: int, d_f: int, count: int, validation: int, dims, batch: int, lr: float,
                 epochs: int, time_limit: float, seed: int = 0, log=print) -> dict:
    """Adam on the first count - validation device crops in batches of
    `batch` (a fresh shuffle per epoch, gathered by device copies); after each
    epoch the mean loss over the held-out crops (lr 0: no update).  Returns
    the run's record."""
    cells = int(np.prod(dims))
    n_train = count - validation
    d_rb = hip.malloc(4 * cells * batch)
    d_fb = hip.malloc(cells * batch)
    rng = np.random.default_rng(seed)
    rec = {"train_crops": n_train, "validation_crops": validation, "batch": batch, "lr": lr, "epochs_planned": epochs, "epochs": [],
           "steps": 0}
    t0 = time.time()
    last = t0
    try:
        for epoch in range(epochs):
            order = rng.permutation(n_train)
            losses = []
            for b0 in range(0, n_train - batch + 1, batch):
                for j, idx in enumerate(order[b0:b0 + batch]):
                    hip.copy(d_rb + 4 * cells * j, d_r + 4 * cells * int(idx), 4 * cells)
                    hip.copy(d_fb + cells * j, d_f + cells * int(idx), cells)
                losses.append(net.train_step(d_rb, d_fb, batch, dims, lr))
                rec["steps"] += 1
                now = time.time()
                if now - last > 10.0:
                    log(f"train: epoch {epoch} step {len(losses)} loss {np.mean(losses[-50:]):.4f} ({now - t0:.0f} s)", flush=True)
                    last = now
                if now - t0 > time_limit:
                    break
            val = []
            for v0 in range(n_train, count - batch + 1, batch):
                val.append(net.train_step(d_r + 4 * cells * v0, d_f + cells * v0, batch, dims, 0.0))
            rec["epochs"].append({"epoch": epoch, "steps": len(losses), "train_loss": float(np.mean(losses)) if losses else None,
                                  "validation_loss": float(np.mean(val)) if val else None, "seconds": round(time.time() - t0, 1)})
            log(f"epoch {epoch}: train loss {rec['epochs'][-1]['train_loss']:.4f}, validation loss "
                f"{rec['epochs'][-1]['validation_loss']:.4f} ({time.time() - t0:.0f} s)", flush=True)
            if time.time() - t0 > time_limit:
                break
    finally:
        hip.free(d_rb)
        hip.free(d_fb)
    rec["seconds"] = round(time.time() - t0, 1)
    rec["ms_per_step"] = round(1e3 * rec["seconds"] / max(rec["steps"], 1), 2)
    return rec


def train(net: abi.LearnedPreconditioner, data: dict, batch: int, lr: float, epochs: int, time_limit: float, seed: int = 0, log=print):
    """Adam on the crops (r scaled back from bfloat16, unit RMS); returns the loss history."""
    hip = abi.Hip()
    r = data["r"].astype(np.float32)
    flags = data["flags"]
    count = r.shape[0]
    dims = tuple(int(v) for v in data["crop"])
    d_r = hip.upload(np.ascontiguousarray(r))
    d_f = hip.upload(np.ascontiguousarray(flags))
    vol = int(np.prod(r.shape[1:]))
    rng = np.random.default_rng(seed)
    history = []
    t0 = time.time()
    last = t0
    try:
        # batches are contiguous runs of a shuffled copy, uploaded once per epoch
        for epoch in range(epochs):
            order = rng.permutation(count)
            hip.free(d_r)
            hip.free(d_f)
            d_r = hip.upload(np.ascontiguousarray(r[order]))
            d_f = hip.upload(np.ascontiguousarray(flags[order]))
            for b0 in range(0, count - batch + 1, batch):
                loss = net.train_step(d_r + 4 * vol * b0, d_f + vol * b0, batch, dims, lr)
                history.append(loss)
                now = time.time()
                if now - last > 10.0:
                    log(f"train: epoch {epoch} batch {b0 // batch} loss {np.mean(history[-20:]):.4f} ({now - t0:.0f} s)", flush=True)
                    last = now
                if now - t0 > time_limit:
                    return history
    finally:
        hip.free(d_r)
        hip.free(d_f)
    return history


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--data", default="/tmp/smoke_training_pairs.bin")
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--crops-per-step", type=int, default=16)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--time-limit", type=float, default=600.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--reuse-data", action="store_true")
    ap.add_argument("--coarse", action="store_true", help="train the multilevel preconditioner's coarse-grid solver")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--capacity", type=int, default=1500, help="coarse volumes captured per scene")
    ap.add_argument("--spheres", action="store_true", help="also capture the obstacle scene's solves")
    ap.add_argument("--sweeps", type=int, default=3, help="pre- and post-sweeps of the V-cycle that produces the coarse data")
    ap.add_argument("--omega", type=float, default=2.0 / 3.0, help="damping of that V-cycle's Jacobi sweeps")
    ap.add_argument("--device-crops", action="store_true", help="keep the crops on the GPU (the 20k-crop, 30-epoch run)")
    ap.add_argument("--samples", type=int, default=20000, help="training crops with --device-crops")
    ap.add_argument("--validation", type=int, default=512, help="held-out crops with --device-crops")
    ap.add_argument("--conv-kernels", choices=("auto", "gather", "staged"), default="auto",
                    help="convolution kernels (smoke_nn_set_conv_kernels; for measurements)")
    args = ap.parse_args()
    abi.set_conv_kernels(args.conv_kernels)
    if args.device_crops:
        import json
        hip = abi.Hip()
        total = args.samples + args.validation
        t0 = time.time()
        d_r, d_f, steps = generate_device_crops(hip, args.grid, total, args.crop, args.crops_per_step, args.seed)
        data_s = time.time() - t0
        try:
            with abi.LearnedPreconditioner(seed=args.seed) as net:
                rec = train_device(net, hip, d_r, d_f, total, args.validation, (args.crop,) * 3, args.batch, args.lr, args.epochs,
                                   args.time_limit, args.seed)
                net.save(args.out)
        finally:
            hip.free(d_r)
            hip.free(d_f)
        done = rec["steps"] * args.batch / (args.samples * args.epochs)
        rec.update({"grid": args.grid, "crop": args.crop, "solves": steps, "data_seconds": round(data_s, 1),
                    "fraction_of_samples_x_epochs": round(done, 4), "weights": str(args.out)})
        print(json.dumps(rec), flush=True)
        return
    if args.coarse:
        scenes = ("plume", "spheres") if args.spheres else ("plume",)
        data = generate_coarse(args.grid, args.levels, args.steps, args.capacity, scenes, sweeps=args.sweeps, omega=args.omega)
        print(f"data: {data['r'].shape[0]} coarse volumes of {data['crop']}", flush=True)
        with abi.LearnedPreconditioner(seed=args.seed) as net:
            net.set_loss(abi.NN_LOSS_ENERGY)
            hist = train(net, data, args.batch, args.lr, args.epochs, args.time_limit, args.seed)
            net.save(args.out)
        print(f"trained: {len(hist)} Adam steps, energy loss {np.mean(hist[:10]):.4f} -> {np.mean(hist[-10:]):.4f}; weights in {args.out}",
              flush=True)
        return
    path = Path(args.data)
    if not (args.reuse_data and path.exists()):
        generate(path, args.grid, args.steps, args.crop, args.crops_per_step, args.seed)
    data = td.read(path)
    print(f"data: {len(data['scale'])} crops of {data['crop']} from a {data['grid']} grid", flush=True)
    with abi.LearnedPreconditioner(seed=args.seed) as net:
        hist = train(net, data, args.batch, args.lr, args.epochs, args.time_limit, args.seed)
        net.save(args.out)
    print(f"trained: {len(hist)} Adam steps, loss {np.mean(hist[:10]):.4f} -> {np.mean(hist[-10:]):.4f}; weights in {args.out}", flush=True)


if __name__ == "__main__":
    main()
