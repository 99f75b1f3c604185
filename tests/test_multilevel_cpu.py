"""Pins the float64 restatement of the multilevel preconditioner
(tests/multilevel_reference.py), which is the checker of csrc/multigrid.hip.

* Level 0's operator equals the oracle's pressure operator (itself pinned
  against an independent graph-Laplacian assembly) on every fluid row, for
  closed, solid, outflow and periodic cases.
* The coarsening rule (solid when every existing child is solid, the anchor's
  parent is the coarse anchor) against a brute-force loop.
* With equal pre- and post-sweeps the V-cycle is a symmetric positive
  definite linear map: restriction and prolongation must be transposes (up to
  the factor 1/2) and the smoother self-adjoint for <M r, s> = <r, M s> to
  hold, so a wrong child index or transfer scaling fails here.
* Flexible CG with the V-cycle reaches the tolerance in a small fraction of
  Jacobi-PCG's iterations on an obstacle scene.
"""
import numpy as np
import pytest

import multilevel_reference as ml
import oracle as orc
import smoke_abi as abi

W = (abi.NO_SLIP_WALL, 0.0, 0.0, 0.0, 0.0)
O = (abi.OUTFLOW, 0.0, 0.0, 0.0, 0.0)
P = (abi.PERIODIC, 0.0, 0.0, 0.0, 0.0)

CASES = [
    ("closed box", (6, 5, 7), [W] * 6, 0.0),
    ("solids", (7, 6, 5), [W] * 6, 0.3),
    ("outflow top", (6, 8, 6), [W, W, W, O, W, W], 0.15),
    ("periodic x", (6, 5, 8), [P, P, W, W, W, W], 0.2),
]


def _level(shape, flow, occ):
    periodic = tuple(flow[2 * a][0] == abi.PERIODIC for a in range(3))
    outflow = tuple(f[0] == abi.OUTFLOW for f in flow)
    return ml.hierarchy(occ, 1, periodic, outflow)[0]


@pytest.mark.parametrize("name,shape,flow,solid_fraction", CASES)
def test_level0_operator_is_the_pressure_operator(oracle_mod, name, shape, flow, solid_fraction):
    rng = np.random.default_rng(len(name))
    occ = (rng.random(shape) < solid_fraction).astype(np.uint8)
    cfg = abi.make_config(shape[2], shape[1], shape[0], flow=flow)
    L = _level(shape, flow, occ)
    assert L.anchor == orc.pressure_anchor(cfg, occ)
    x = np.where(L.ident, 0.0, rng.normal(size=shape))
    ref = orc.pressure_apply(cfg, occ, x)
    fluid = ~L.ident
    np.testing.assert_allclose(L.apply(x)[fluid], ref[fluid], rtol=1e-14, atol=1e-13)


def test_coarsening_rule_brute_force():
    rng = np.random.default_rng(4)
    solid = rng.random((5, 7, 3)) < 0.6
    fluid = np.flatnonzero(~solid.ravel())
    anchor = int(fluid[0])
    cs, ca = ml.coarsen(solid, anchor)
    assert cs.shape == (3, 4, 2)
    for Z in range(3):
        for Y in range(4):
            for X in range(2):
                kids = [solid[z, y, x] for z in (2 * Z, 2 * Z + 1) for y in (2 * Y, 2 * Y + 1) for x in (2 * X, 2 * X + 1)
                        if z < 5 and y < 7 and x < 3]
                assert cs[Z, Y, X] == all(kids)
    z, y, x = np.unravel_index(anchor, solid.shape)
    assert ca == (z // 2) * 8 + (y // 2) * 2 + x // 2


def _spheres(n, count, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    occ = np.zeros((n, n, n), np.uint8)
    for _ in range(count):
        r = rng.uniform(0.08, 0.18) * n
        c = rng.uniform(r, n - r, size=3)
        occ[(x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r] = 1
    return occ


@pytest.mark.parametrize("pre,levels", [(1, 3), (2, 3), (2, 4)])
def test_vcycle_is_symmetric_positive_definite(pre, levels):
    occ = _spheres(18, 4, 1)
    occ = occ[:, :17, :15]            # odd extents: partial coarse cells
    hier = ml.hierarchy(occ, levels)
    rng = np.random.default_rng(pre + levels)
    mask = ~hier[0].ident
    r = np.where(mask, rng.normal(size=occ.shape), 0.0)
    s = np.where(mask, rng.normal(size=occ.shape), 0.0)
    M = lambda v: ml.vcycle(hier, v, pre=pre, post=pre, coarse_sweeps=5)
    mr, ms = M(r), M(s)
    assert abs(np.vdot(mr, s) - np.vdot(r, ms)) <= 1e-12 * np.linalg.norm(mr) * np.linalg.norm(s)
    assert np.vdot(mr, r) > 0 and np.vdot(ms, s) > 0
    np.testing.assert_allclose(M(2.0 * r - 3.0 * s), 2.0 * mr - 3.0 * ms, rtol=1e-12, atol=1e-12)
    assert np.all(mr[~mask] == 0.0)


def test_multilevel_fcg_needs_far_fewer_iterations_than_jacobi():
    occ = _spheres(32, 6, 2)
    hier = ml.hierarchy(occ, 4)
    L = hier[0]
    rng = np.random.default_rng(0)
    b = np.where(L.ident, 0.0, rng.normal(size=occ.shape))
    x_j, it_j = ml.flexible_cg(L, b, lambda r: np.where(L.ident, 0.0, r / L.diag))
    x_m, it_m = ml.flexible_cg(L, b, lambda r: ml.vcycle(hier, r, coarse_sweeps=16))
    for x in (x_j, x_m):
        assert np.linalg.norm(b - L.apply(x)) <= 1e-6 * np.linalg.norm(b)
    assert it_m * 5 < it_j, (it_m, it_j)


def test_deflated_coarse_solve_is_exact_for_an_exact_solver():
    """z = c v + P^T S(P r) equals A^-1 r when S = A^-1 (the identity
    A^-1 = Q + P^T A^-1 P), for an anchored level with solids; and with S = 0
    it is the v-component Q r alone."""
    occ = _spheres(10, 3, 5)
    L = ml.hierarchy(occ, 2)[1]
    fluid = ~L.ident
    n = L.solid.size
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        cols.append(L.apply(np.where(L.ident, 0.0, e.reshape(L.solid.shape))).ravel())
    A = np.array(cols).T
    idx = np.flatnonzero(fluid.ravel())
    Af = A[np.ix_(idx, idx)]

    def exact(r):
        out = np.zeros(n)
        out[idx] = np.linalg.solve(Af, r.ravel()[idx])
        return out.reshape(r.shape)

    rng = np.random.default_rng(9)
    r = np.where(fluid, rng.normal(size=L.solid.shape) + 0.3, 0.0)
    np.testing.assert_allclose(ml.deflated_solve(L, r, exact), exact(r), rtol=1e-10, atol=1e-10 * np.abs(exact(r)).max())
    v, av, vav = ml.deflation_vector(L)
    assert vav > 0
    np.testing.assert_allclose(ml.deflated_solve(L, r, np.zeros_like), v * (np.vdot(v, r) / vav), rtol=1e-12)
