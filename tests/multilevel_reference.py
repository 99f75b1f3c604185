"""Float64 numpy restatement of the multilevel preconditioner declared in
visual-simulation-of-smoke-3d-ext.h (SMOKE_SIMULATION_PRECONDITIONER_MULTILEVEL):
hierarchy, rediscretised 7-point operators, damped-Jacobi sweeps, the
piecewise-constant transfers and the V-cycle with a Jacobi coarse solver.
Test infrastructure only: the checker of csrc/multigrid.hip.

Arrays are (nz, ny, nx), x fastest, like the library's buffers."""
from __future__ import annotations

import numpy as np

IDENTITY = 0x40


def flags_of(solid: np.ndarray, anchor: int, periodic=(False, False, False)) -> np.ndarray:
    """Flag byte: bit d (x-, x+, y-, y+, z-, z+) = that neighbour exists
    (periodic axes wrap) and is not solid; identity for solid cells and the
    anchor."""
    fluid = ~solid
    flags = np.zeros(solid.shape, np.uint8)
    for d in range(6):
        ax = 2 - (d >> 1)            # array axis of x, y, z
        s = 1 if d & 1 else -1
        nb = np.roll(fluid, -s, axis=ax)
        if not periodic[d >> 1]:
            edge = [slice(None)] * 3
            edge[ax] = -1 if s == 1 else 0
            nb = nb.copy()
            nb[tuple(edge)] = False
        flags |= (nb.astype(np.uint8) << d)
    ident = solid.copy()
    if anchor < solid.size:
        ident.flat[anchor] = True
    flags[ident] = IDENTITY
    return flags


def coarsen(solid: np.ndarray, anchor: int):
    nz, ny, nx = solid.shape
    cz, cy, cx = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    pad = np.ones((2 * cz, 2 * cy, 2 * cx), bool)     # missing children count as solid
    pad[:nz, :ny, :nx] = solid
    csolid = pad.reshape(cz, 2, cy, 2, cx, 2).all(axis=(1, 3, 5))
    if anchor >= solid.size:
        return csolid, csolid.size
    z, y, x = np.unravel_index(anchor, solid.shape)
    return csolid, int(np.ravel_multi_index((z // 2, y // 2, x // 2), csolid.shape))


class Level:
    def __init__(self, solid, anchor, periodic, outflow):
        self.solid, self.anchor = solid, anchor
        self.periodic, self.outflow = periodic, outflow
        self.flags = flags_of(solid, anchor, periodic)
        self.ident = (self.flags & IDENTITY) != 0
        nz, ny, nx = solid.shape
        extra = np.zeros(solid.shape)
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        for f, (coord, val) in enumerate(((x, 0), (x, nx - 1), (y, 0), (y, ny - 1), (z, 0), (z, nz - 1))):
            if outflow[f]:
                extra += coord == val
        bits = sum(((self.flags >> d) & 1).astype(np.float64) for d in range(6))
        self.diag = np.where(self.ident, 1.0, np.maximum(bits + extra, 1.0))

    def apply(self, x):
        """A x on fluid rows (identity rows give 0); x must be 0 on identity rows."""
        out = self.diag * x
        for d in range(6):
            ax = 2 - (d >> 1)
            s = 1 if d & 1 else -1
            nb = np.roll(x, -s, axis=ax)
            out -= np.where((self.flags >> d) & 1, nb, 0.0)
        return np.where(self.ident, 0.0, out)

    def sweep(self, r, x, omega):
        return np.where(self.ident, 0.0, x + omega * (r - self.apply(x)) / self.diag)


def hierarchy(occupancy: np.ndarray, levels: int, periodic=(False, False, False), outflow=(False,) * 6):
    solid = occupancy != 0
    fluid = np.flatnonzero(~solid.ravel())
    anchor = int(fluid[0]) if fluid.size else solid.size
    out = [Level(solid, anchor, periodic, outflow)]
    for _ in range(levels - 1):
        solid, anchor = coarsen(solid, anchor)
        out.append(Level(solid, anchor, periodic, outflow))
    return out


def deflation_vector(L: Level):
    """v = 1 on the fluid rows and A v (nonzero only next to the anchor and on
    OUTFLOW faces); v^T A v."""
    v = np.where(L.ident, 0.0, 1.0)
    av = L.apply(v)
    return v, av, float(np.vdot(v, av))


def deflated_solve(L: Level, r, solve):
    """z = Q r + P^T solve(P r) with Q = v v^T / (v^T A v), P = I - A Q: exact
    when `solve` is A^-1, and the near-null mode of the anchored operator
    (v, nearly constant) never reaches `solve`."""
    v, av, vav = deflation_vector(L)
    if vav <= 0.0:
        return solve(r)
    c = np.vdot(v, r) / vav
    zp = solve(np.where(L.ident, 0.0, r - c * av))
    return np.where(L.ident, 0.0, zp + (c - np.vdot(av, zp) / vav) * v)


def restrict(res: np.ndarray, coarse_shape) -> np.ndarray:
    cz, cy, cx = coarse_shape
    pad = np.zeros((2 * cz, 2 * cy, 2 * cx))
    pad[:res.shape[0], :res.shape[1], :res.shape[2]] = res
    return 0.5 * pad.reshape(cz, 2, cy, 2, cx, 2).sum(axis=(1, 3, 5))


def prolong(e: np.ndarray, fine_shape) -> np.ndarray:
    up = np.repeat(np.repeat(np.repeat(e, 2, 0), 2, 1), 2, 2)
    return up[:fine_shape[0], :fine_shape[1], :fine_shape[2]]


def vcycle(levels, r, pre=2, post=2, omega=2.0 / 3.0, coarse_sweeps=8, l=0, coarse=None):
    """z = M(r); the coarsest level is solved by `coarse(level, r)` (default:
    `coarse_sweeps` damped-Jacobi sweeps from 0) inside the deflation of
    deflated_solve."""
    L = levels[l]
    r = np.where(L.ident, 0.0, r)
    if l == len(levels) - 1:
        def jacobi(rr):
            x = np.zeros_like(rr)
            for _ in range(coarse_sweeps):
                x = L.sweep(rr, x, omega)
            return x
        solve = (lambda rr: coarse(L, rr)) if coarse is not None else jacobi
        return deflated_solve(L, r, solve) if l > 0 else solve(r)
    x = np.zeros_like(r)
    for _ in range(pre):
        x = L.sweep(r, x, omega)
    C = levels[l + 1]
    rc = np.where(C.ident, 0.0, restrict(np.where(L.ident, 0.0, r - L.apply(x)), C.solid.shape))
    ec = vcycle(levels, rc, pre, post, omega, coarse_sweeps, l + 1, coarse)
    x = np.where(L.ident, 0.0, x + prolong(ec, L.solid.shape))
    for _ in range(post):
        x = L.sweep(r, x, omega)
    return x


def flexible_cg(level: Level, b, precond, tol=1e-6, maxit=5000):
    """Polak-Ribiere flexible CG from x = 0; returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    bn = np.linalg.norm(b)
    z = precond(r)
    p = z.copy()
    rz = np.vdot(r, z)
    for k in range(1, maxit + 1):
        ap = level.apply(p)
        a = rz / np.vdot(p, ap)
        x += a * p
        r -= a * ap
        if np.linalg.norm(r) <= tol * bn:
            return x, k
        zn = precond(r)
        rzn = np.vdot(r, zn)
        beta = (rzn - np.vdot(r, z)) / rz
        z, rz = zn, rzn
        p = z + beta * p
    return x, maxit
