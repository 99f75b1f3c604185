"""The batched-plume workload: many independent plumes of uneven grid sizes.

plume_set(count, seed) draws the scenes: sizes from {32^3, 48^3, 64^3,
96x64x64, 128x64x64, 96^3, 128^3} (nx x ny x nz, y vertical), cell size 1/64
for every plume, a source sphere of radius U(3, 12) cells at a random (x, z)
with its centre in the bottom 10% of y, temperature rate U(0.5, 2) and
vorticity confinement U(0, 4); closed no-slip boxes, monotone-cubic
advection, buoyancy as in the single-plume scene (scene_plume.cpp:271-296).
"""
from __future__ import annotations

import numpy as np

import smoke_abi as abi

SIZES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 64, 64), (128, 64, 64), (96, 96, 96), (128, 128, 128)]


def plume_set(count: int = 64, seed: int = 3, dt: float = 0.01):
    """[(config, plume source)] for `count` plumes drawn with `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx, ny, nz = SIZES[int(rng.integers(len(SIZES)))]
        r = float(rng.uniform(3.0, 12.0))
        cx = float(rng.uniform(r, nx - r))
        cz = float(rng.uniform(r, nz - r))
        cy = float(rng.uniform(0.0, 0.1 * ny))
        t_amp = float(rng.uniform(0.5, 2.0))
        eps = float(rng.uniform(0.0, 4.0))
        cfg = abi.make_config(nx, ny, nz, cell_size=1.0 / 64, dt=dt, pressure_iterations=64, ambient=0.0, buoyancy_density=0.05,
                              buoyancy_temperature=0.25, vorticity=eps, cubic=True)
        src = abi.PlumeSource(cx, cy, cz, r, 1.0, t_amp, 0.05, seed * 1000 + i)
        out.append((cfg, src))
    return out


def total_cells(plumes) -> int:
    return sum(cfg.nx * cfg.ny * cfg.nz for cfg, _ in plumes)
