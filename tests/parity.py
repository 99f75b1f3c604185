"""Shared GPU-vs-oracle parity harness.

Builds a seeded state (smooth velocity, density / temperature blobs, optional
solid spheres and plume sources), loads the same bytes into the HIP library
(through the C ABI: update_* with device pointers, import_state for the MAC
faces) and into the FP64 oracle, runs the same stages on both and reports
per-field relative L2 errors.
"""
from __future__ import annotations

import numpy as np

import smoke_abi as abi
import oracle as orc

W = (abi.NO_SLIP_WALL, 0.0, 0.0, 0.0, 0.0)
P = (abi.PERIODIC, 0.0, 0.0, 0.0, 0.0)
O = (abi.OUTFLOW, 0.0, 0.0, 0.0, 0.0)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return float(num / den) if den > 0 else float(num)


def user_force(cfg: abi.Config, rng, amplitude=2.0):
    """A smooth random cell force plus a small per-cell jitter, set in every
    cell (solid cells included: the reference keeps the user force there,
    B:902-908, and its scene caller sets it each frame, scene_plume.cpp:321-357)."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    zc, yc, xc = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx, indexing="ij")
    out = []
    for _ in range(3):
        k = rng.integers(1, 4, size=3)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        f = np.sin(np.pi * k[0] * xc + ph[0]) * np.cos(np.pi * k[1] * yc + ph[1]) * np.sin(np.pi * k[2] * zc + ph[2])
        f += 0.1 * rng.uniform(-1.0, 1.0, f.shape)
        out.append((amplitude * f).astype(np.float32))
    return out


def report(tag: str, errs: dict) -> None:
    """One parseable line per comparison: the measured errors behind each bound."""
    print(f"PARITY {tag} " + " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


def make_state(cfg: abi.Config, seed: int, amplitude=0.5, solids=0, sources=True, solid_motion=False, force=True):
    rng = np.random.default_rng(seed)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    h = float(np.float32(cfg.cell_size))
    st = {}
    for comp, name in enumerate(("u", "v", "w")):
        shape = abi.face_shape(cfg, comp)
        z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[comp] = 0.0
        px, py, pz = (x + off[0]) / nx, (y + off[1]) / ny, (z + off[2]) / nz
        f = np.zeros(shape)
        for _ in range(3):
            k = rng.integers(1, 4, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            f += rng.normal() * np.sin(np.pi * k[0] * px + ph[0]) * np.sin(np.pi * k[1] * py + ph[1]) * np.sin(np.pi * k[2] * pz + ph[2])
        st[name] = (amplitude * f / 3).astype(np.float32)
    zc, yc, xc = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    rho = np.zeros((nz, ny, nx))
    T = np.full((nz, ny, nx), cfg.ambient_temperature, np.float64)
    for _ in range(3):
        c = rng.uniform(0.2, 0.8, 3) * np.array([nx, ny, nz])
        r = rng.uniform(0.1, 0.25) * min(nx, ny, nz)
        d2 = ((xc - c[0]) ** 2 + (yc - c[1]) ** 2 + (zc - c[2]) ** 2) / r ** 2
        rho += np.exp(-d2) * rng.uniform(0.5, 1.5)
        T += np.exp(-d2) * rng.uniform(0.2, 1.0)
    st["density"] = rho.astype(np.float32)
    st["temperature"] = T.astype(np.float32)
    occ = np.zeros((nz, ny, nx), np.uint8)
    for _ in range(solids):
        c = rng.uniform(0.25, 0.75, 3) * np.array([nx, ny, nz])
        r = rng.uniform(0.08, 0.15) * min(nx, ny, nz)
        occ[((xc - c[0]) ** 2 + (yc - c[1]) ** 2 + (zc - c[2]) ** 2) <= r * r] = 1
    st["occupancy"] = occ
    st["solid_velocity"] = [(rng.normal(size=occ.shape) * 0.1 * solid_motion).astype(np.float32) for _ in range(3)]
    st["solid_temperature"] = np.full(occ.shape, cfg.ambient_temperature + 0.5, np.float32)
    if sources:
        src = (xc - nx / 2) ** 2 + (yc - ny * 0.15) ** 2 + (zc - nz / 2) ** 2 <= (0.12 * min(nx, ny, nz)) ** 2
        st["density_source"] = np.where(src, 1.0, 0.0).astype(np.float32)
        st["temperature_source"] = np.where(src, 1.0 + 0.05 * rng.uniform(-1, 1, src.shape), 0.0).astype(np.float32)
    else:
        st["density_source"] = np.zeros(occ.shape, np.float32)
        st["temperature_source"] = np.zeros(occ.shape, np.float32)
    st["force"] = user_force(cfg, np.random.default_rng(seed + 1000)) if force else [np.zeros(occ.shape, np.float32) for _ in range(3)]
    return st


def load_gpu(ctx: abi.SmokeContext, st):
    ctx.set_occupancy(st["occupancy"])
    ctx.update("solid_velocity", *st["solid_velocity"])
    ctx.update("solid_temperature", st["solid_temperature"])
    ctx.update("density_source", st["density_source"])
    ctx.update("temperature_source", st["temperature_source"])
    ctx.update("force", *st["force"])
    ctx.import_state(st["u"], st["v"], st["w"], st["density"], st["temperature"], st.get("pressure"))


def load_oracle(s: orc.OracleState, st):
    s["occupancy"] = st["occupancy"]
    for k, n in enumerate(("x", "y", "z")):
        s[f"solid_velocity_{n}"] = st["solid_velocity"][k]
        s[f"force_user_{n}"] = st["force"][k]
    s["solid_temperature"] = st["solid_temperature"]
    s["density_source"] = st["density_source"]
    s["temperature_source"] = st["temperature_source"]
    s["u"], s["v"], s["w"] = st["u"], st["v"], st["w"]
    s["density"], s["temperature"] = st["density"], st["temperature"]
    if "pressure" in st:
        s["pressure"] = st["pressure"]


def tolerance_options(ctx, tol=1e-6, adaptive=False, cfl=0.5, dt_max=0.01, warm=1, preconditioner=abi.PRECONDITIONER_JACOBI):
    ctx.set_solver_options(pressure_mode=abi.PRESSURE_TOLERANCE, preconditioner=preconditioner, relative_tolerance=tol,
                           max_iterations=20000, max_refinements=6, warm_start=warm, adaptive_dt=1 if adaptive else 0, cfl=cfl,
                           dt_max=dt_max, check_interval=16)


def one_step_parity(cfg: abi.Config, seed=1, steps_before=0, solids=0, tol=1e-6):
    """Full step on GPU (tolerance PCG) and oracle (FP64, tight tolerance)
    from the same imported state; relative L2 error per field."""
    st = make_state(cfg, seed, solids=solids)
    with abi.SmokeContext(cfg) as ctx:
        load_gpu(ctx, st)
        tolerance_options(ctx, tol=tol)
        ctx.step()
        g = ctx.export_state()
        stats = ctx.stats()
    s = orc.OracleState(cfg)
    load_oracle(s, st)
    s.step(orc.options(mode=1, tol=1e-10, inner_reduction=1e-12))
    errs = {
        "u": rel_l2(g["velocity_x"], s["u"]), "v": rel_l2(g["velocity_y"], s["v"]), "w": rel_l2(g["velocity_z"], s["w"]),
        "density": rel_l2(g["density"], s["density"]), "temperature": rel_l2(g["temperature"], s["temperature"]),
    }
    errs["solver_relative_residual"] = stats.relative_residual
    return errs
