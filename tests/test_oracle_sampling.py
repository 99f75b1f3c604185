"""Pins the oracle's interpolation, ghost rules and backtrace (B = backend-cuda.cu).

Every expectation comes from a closed form or invariant, not from the oracle's
own formulas: affine fields are reproduced exactly by both samplers, sampling
on grid nodes returns the node value, a uniform flow traces an exact straight
displacement, and clipped traces land on the closed-form bisection point.
"""
import math

import numpy as np
import pytest

import smoke_abi as abi
import oracle as orc


def test_monotonic_cubic_endpoints_and_linear_exactness(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(50):
        p = rng.normal(size=4)
        # endpoints: Hermite basis h00(0)=1, h01(1)=1 (B:378-392)
        assert orc.monotonic_cubic(*p, 0.0) == pytest.approx(p[1], abs=1e-15)
        assert orc.monotonic_cubic(*p, 1.0) == pytest.approx(p[2], abs=1e-14)
    for a, b in [(0.3, 1.7), (-2.0, -0.5), (5.0, 1e-3)]:
        p = [a + b * k for k in range(4)]
        for t in np.linspace(0, 1, 11):
            # linear data: both central slopes equal b, the cubic is exact
            assert orc.monotonic_cubic(*p, t) == pytest.approx(a + b * (1 + t), rel=1e-13, abs=1e-13)


def test_monotonic_cubic_slope_zeroing(oracle_mod):
    t = 0.3
    smooth = 3 * t * t - 2 * t ** 3          # h01; h00 = 1 - h01
    # |delta| < 1e-6: both slopes zero -> pure smoothstep between p1 and p2
    assert orc.monotonic_cubic(7.0, 1.0, 1.0 + 5e-7, -3.0, t) == pytest.approx(1.0 + 5e-7 * smooth, abs=1e-15)
    # local minimum at p1 (m1 * delta <= 0): m1 -> 0, m2 = (p3 - p1)/2 kept
    p0, p1, p2, p3 = 2.0, 0.0, 1.0, 2.0
    m2 = 0.5 * (p3 - p1)
    h11 = t ** 3 - t ** 2
    expect = (1 - smooth) * p1 + smooth * p2 + h11 * m2
    assert orc.monotonic_cubic(p0, p1, p2, p3, t) == pytest.approx(expect, abs=1e-14)
    # overshoot is possible (the result is not clamped): with m1 = 5.5, m2 = 0.5
    # the Hermite value at t = 0.5 is 0.6875 + 0.5 - 0.0625 = 1.125 > max(p1, p2)
    assert orc.monotonic_cubic(-10.0, 0.0, 1.0, 1.0, 0.5) == pytest.approx(1.125, abs=1e-15)


def _affine(shape, offsets, h, coef):
    """Field value a0 + a.x at the node positions of a (z,y,x) array."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    px, py, pz = (x + offsets[0]) * h, (y + offsets[1]) * h, (z + offsets[2]) * h
    return coef[0] + coef[1] * px + coef[2] * py + coef[3] * pz


@pytest.mark.parametrize("cubic", [False, True])
def test_samplers_reproduce_affine_fields(oracle_mod, cubic):
    h = 0.125                                     # exact in FP32 (the ABI's cell_size type)
    cfg = abi.make_config(9, 8, 10, cell_size=h)
    coef = (0.25, 1.5, -0.75, 2.0)
    rng = np.random.default_rng(1)
    kinds = [(orc.FIELD_DENSITY, (0.5, 0.5, 0.5), abi.cell_shape(cfg)),
             (orc.FIELD_U, (0.0, 0.5, 0.5), abi.face_shape(cfg, 0)),
             (orc.FIELD_V, (0.5, 0.0, 0.5), abi.face_shape(cfg, 1)),
             (orc.FIELD_W, (0.5, 0.5, 0.0), abi.face_shape(cfg, 2))]
    for kind, off, shape in kinds:
        f = _affine(shape, off, h, coef)
        for _ in range(25):
            # interior points: all 64 cubic taps stay inside the array
            p = np.array([rng.uniform(2.5, 5.5), rng.uniform(2.5, 4.5), rng.uniform(2.5, 6.5)]) * h
            got = orc.sample(cfg, kind, cubic, f, *p)
            assert got == pytest.approx(coef[0] + coef[1] * p[0] + coef[2] * p[1] + coef[3] * p[2], rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("cubic", [False, True])
def test_samplers_on_nodes_return_node_values(oracle_mod, cubic):
    cfg = abi.make_config(6, 7, 5, cell_size=0.5)
    rng = np.random.default_rng(2)
    f = rng.normal(size=abi.face_shape(cfg, 1))
    for (i, j, k) in [(0, 0, 0), (5, 7, 4), (2, 3, 1), (3, 6, 2)]:
        # v faces sit at ((i+.5)h, j h, (k+.5)h) (B:547-549)
        got = orc.sample(cfg, orc.FIELD_V, cubic, f, (i + 0.5) * 0.5, j * 0.5, (k + 0.5) * 0.5)
        assert got == pytest.approx(f[k, j, i], abs=1e-14)


def test_scalar_ghost_rules(oracle_mod):
    bc = [(abi.FIXED_VALUE, 3.5), (abi.ZERO_FLUX, 0.0), (abi.SCALAR_PERIODIC, 0.0), (abi.SCALAR_PERIODIC, 0.0),
          (abi.ZERO_FLUX, 0.0), (abi.FIXED_VALUE, -1.0)]
    cfg = abi.make_config(4, 5, 3, density_bc=bc)
    f = np.arange(60, dtype=np.float64).reshape(3, 5, 4)
    L = lambda i, j, k: orc.load(cfg, orc.FIELD_DENSITY, f, i, j, k)
    assert L(-1, 2, 1) == 3.5                     # FIXED_VALUE x-: the face value
    assert L(4, 2, 1) == f[1, 2, 3]               # ZERO_FLUX x+: clamp
    assert L(1, -1, 1) == f[1, 4, 1]              # periodic y: wrap
    assert L(1, 6, 1) == f[1, 1, 1]
    assert L(1, 2, -2) == f[0, 2, 1]              # ZERO_FLUX z-
    assert L(1, 2, 3) == -1.0                     # FIXED_VALUE z+
    assert L(-1, 2, 3) == 3.5                     # x is tested first (early return)
    assert L(4, 5, 3) == -1.0                     # x clamps, y wraps, z returns its value


def test_face_ghost_rules(oracle_mod):
    flow = [(abi.NO_SLIP_WALL, 0.3, 0.7, -0.2, 0.0), (abi.OUTFLOW, 0.0, 0.0, 0.0, 0.0),
            (abi.FREE_SLIP_WALL, 0.1, 0.9, 0.4, 0.0), (abi.NO_SLIP_WALL, 0.5, -0.6, 0.2, 0.0),
            (abi.PERIODIC, 0, 0, 0, 0), (abi.PERIODIC, 0, 0, 0, 0)]
    cfg = abi.make_config(3, 4, 5, flow=flow)
    rng = np.random.default_rng(3)
    u = rng.normal(size=abi.face_shape(cfg, 0))   # (5, 4, 4)
    v = rng.normal(size=abi.face_shape(cfg, 1))   # (5, 5, 3)
    Lu = lambda i, j, k: orc.load(cfg, orc.FIELD_U, u, i, j, k)
    Lv = lambda i, j, k: orc.load(cfg, orc.FIELD_V, v, i, j, k)
    # u, normal axis x: no-slip wall -> 2*vx_wall - u[face 0]; outflow -> u[face nx]
    assert Lu(-1, 1, 2) == pytest.approx(2 * 0.3 - u[2, 1, 0])
    assert Lu(4, 1, 2) == pytest.approx(u[2, 1, 3])
    # u, tangential axis y: free slip -> interior, no-slip -> reflect about vx_wall
    assert Lu(2, -1, 2) == pytest.approx(u[2, 0, 2])
    assert Lu(2, 4, 2) == pytest.approx(2 * 0.5 - u[2, 3, 2])
    # periodic z wraps the index
    assert Lu(2, 1, -1) == pytest.approx(u[4, 1, 2])
    assert Lu(2, 1, 5) == pytest.approx(u[0, 1, 2])
    # v, tangential axis x: no-slip x- reflects about vy of that face; outflow x+ copies
    assert Lv(-1, 2, 1) == pytest.approx(2 * 0.7 - v[1, 2, 0])
    assert Lv(3, 2, 1) == pytest.approx(v[1, 2, 2])
    # v, normal axis y: free-slip wall is NOT interior on the normal axis
    assert Lv(1, -1, 1) == pytest.approx(2 * 0.9 - v[1, 0, 1])
    assert Lv(1, 5, 1) == pytest.approx(2 * -0.6 - v[1, 4, 1])


def test_center_velocity_ghost_rules(oracle_mod):
    flow = [(abi.FREE_SLIP_WALL, 0.2, 0.3, 0.4, 0.0), (abi.NO_SLIP_WALL, 0.1, 0.2, 0.3, 0.0),
            (abi.OUTFLOW, 0, 0, 0, 0), (abi.NO_SLIP_WALL, 0, 0, 0, 0),
            (abi.PERIODIC, 0, 0, 0, 0), (abi.PERIODIC, 0, 0, 0, 0)]
    cfg = abi.make_config(3, 3, 4, flow=flow)
    f = np.random.default_rng(4).normal(size=abi.cell_shape(cfg))
    L = lambda c, x, y, z: orc.load_center_component(cfg, c, f, x, y, z)
    assert L(0, -1, 1, 2) == pytest.approx(2 * 0.2 - f[2, 1, 0])   # free slip, normal component reflects
    assert L(1, -1, 1, 2) == pytest.approx(f[2, 1, 0])             # free slip, tangential: interior
    assert L(2, 3, 1, 2) == pytest.approx(2 * 0.3 - f[2, 1, 2])    # no slip: reflect about prescribed
    assert L(1, 1, -1, 2) == pytest.approx(f[2, 0, 1])             # outflow: interior
    assert L(0, 1, 1, -1) == pytest.approx(f[3, 1, 1])             # periodic


def _uniform_flow_cfg(n, U):
    flow = [(abi.NO_SLIP_WALL, U[0], U[1], U[2], 0.0)] * 6
    return abi.make_config(n, n, n, cell_size=1.0 / n, flow=flow)


def test_trace_uniform_flow_is_exact(oracle_mod):
    n = 12
    U = (0.8, -0.4, 0.3)
    cfg = _uniform_flow_cfg(n, U)
    u = np.full(abi.face_shape(cfg, 0), U[0])
    v = np.full(abi.face_shape(cfg, 1), U[1])
    w = np.full(abi.face_shape(cfg, 2), U[2])
    occ = np.zeros(abi.cell_shape(cfg), np.uint8)
    dt = 0.05
    for start in ([0.5, 0.5, 0.5], [0.31, 0.77, 0.42], [0.95, 0.05, 0.6]):
        end = np.array(start) - dt * np.array(U)
        if np.all(end >= 0) and np.all(end <= 1):
            got = orc.trace(cfg, u, v, w, occ, dt, start)
            np.testing.assert_allclose(got, end, rtol=0, atol=1e-14)


def test_trace_clipped_by_wall_and_solid(oracle_mod):
    n = 10
    U = (2.0, 0.0, 0.0)
    cfg = _uniform_flow_cfg(n, U)
    u = np.full(abi.face_shape(cfg, 0), U[0])
    v = np.zeros(abi.face_shape(cfg, 1))
    w = np.zeros(abi.face_shape(cfg, 2))
    occ = np.zeros(abi.cell_shape(cfg), np.uint8)
    dt = 0.2
    s = np.array([0.13, 0.5, 0.5])
    e = s[0] - dt * U[0]                          # -0.27: outside x in [0, 1]
    lo = math.floor(1024 * s[0] / (s[0] - e)) / 1024
    got = orc.trace(cfg, u, v, w, occ, dt, s)
    np.testing.assert_allclose(got, [s[0] + (e - s[0]) * lo, 0.5, 0.5], atol=1e-14)
    # solid slab in cells x <= 2, i.e. x < 3h with h the FP32 cell size: the
    # bisection returns the largest k/1024 whose point is still in fluid
    occ[:, :, :3] = 1
    h = float(np.float32(1.0 / n))
    s = np.array([0.57, 0.5, 0.5])
    e = s[0] - dt * U[0]
    ks = [k for k in range(1025) if s[0] + (e - s[0]) * (k / 1024) >= 3 * h]
    lo = max(ks) / 1024
    got = orc.trace(cfg, u, v, w, occ, dt, s)
    np.testing.assert_allclose(got, [s[0] + (e - s[0]) * lo, 0.5, 0.5], atol=1e-14)
