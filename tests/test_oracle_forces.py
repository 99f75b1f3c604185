"""Pins the oracle's force stage, enforce rules, density fill and whole-step
invariants (B = backend-cuda.cu)."""
import numpy as np
import pytest

import smoke_abi as abi
import oracle as orc

W = (abi.NO_SLIP_WALL, 0.0, 0.0, 0.0, 0.0)
P = (abi.PERIODIC, 0.0, 0.0, 0.0, 0.0)
O = (abi.OUTFLOW, 0.0, 0.0, 0.0, 0.0)


def face_positions(cfg, comp):
    shape = abi.face_shape(cfg, comp)
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    h = float(np.float32(cfg.cell_size))
    off = [0.5, 0.5, 0.5]
    off[comp] = 0.0
    return (x + off[0]) * h, (y + off[1]) * h, (z + off[2]) * h


def test_solid_body_rotation_vorticity(oracle_mod):
    # u = -Om (y - c), v = Om (x - c): omega_z = 2 Om exactly away from the walls (B:870-900)
    n, om = 10, 1.75
    cfg = abi.make_config(n, n, n, cell_size=0.125)
    s = orc.OracleState(cfg)
    c = n * 0.125 / 2
    x, y, z = face_positions(cfg, 0)
    s["u"] = -om * (y - c)
    x, y, z = face_positions(cfg, 1)
    s["v"] = om * (x - c)
    s.diagnostics()
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(s["vorticity_z"][inner], 2 * om, rtol=1e-12)
    np.testing.assert_allclose(s["vorticity_x"][inner], 0, atol=1e-12)
    np.testing.assert_allclose(s["vorticity_y"][inner], 0, atol=1e-12)
    np.testing.assert_allclose(s["vorticity_magnitude"][inner], 2 * om, rtol=1e-12)
    # cell-centred velocity is the face average (B:859-868)
    xc = (np.arange(n) + 0.5) * 0.125
    np.testing.assert_allclose(s["center_y"][3, 4, :], om * (xc - c), rtol=1e-12, atol=1e-14)


def test_confinement_on_linear_vorticity_ramp(oracle_mod):
    # v = k x^2 / 2 gives omega_z = k x (exact central differences), so
    # grad|omega| = (k, 0, 0), N = x-hat, N x omega = (0, -omega_z, 0):
    # f_y = -eps h k x on cells two away from the walls (B:920-948)
    n, k, eps, h = 12, 3.0, 2.0, 0.125
    cfg = abi.make_config(n, n, n, cell_size=h, vorticity=eps, buoyancy_density=0.0, buoyancy_temperature=0.0)
    s = orc.OracleState(cfg)
    x, y, z = face_positions(cfg, 1)
    s["v"] = k * x * x / 2
    s.run_stages(0, 0, orc.options())
    xc = (np.arange(n) + 0.5) * h
    inner = (slice(2, -2),) * 3
    expect = -eps * h * k * np.broadcast_to(xc, (n, n, n))
    np.testing.assert_allclose(s["force_y"][inner], expect[inner], rtol=1e-10)
    np.testing.assert_allclose(s["force_x"][inner], 0, atol=1e-12)
    np.testing.assert_allclose(s["force_z"][inner], 0, atol=1e-12)


def test_buoyancy_and_force_to_face(oracle_mod):
    n, dt = 6, 0.02
    a, b, tamb = 0.3, 0.7, 0.1
    cfg = abi.make_config(n, n, n, dt=dt, buoyancy_density=a, buoyancy_temperature=b, ambient=tamb, vorticity=0.0)
    rng = np.random.default_rng(5)
    s = orc.OracleState(cfg)
    s["density"] = rng.random(s["density"].shape)
    s["temperature"] = rng.random(s["temperature"].shape)
    s["force_user_x"] = rng.normal(size=s["density"].shape)
    occ = np.zeros(abi.cell_shape(cfg), np.uint8)
    occ[3, 3, 3] = 1
    s["occupancy"] = occ
    s["solid_temperature"] = 9.0
    s.run_stages(0, 0, orc.options())
    a32, b32 = float(np.float32(a)), float(np.float32(b))   # the ABI carries FP32 factors
    fy = -a32 * s["density"] + b32 * (s["temperature"] - float(np.float32(tamb)))
    fy[occ == 1] = 0.0
    np.testing.assert_allclose(s["force_y"], fy, rtol=1e-12, atol=1e-15)
    assert s["temperature"][3, 3, 3] == 9.0      # solid temperature applied first (B:2001)
    np.testing.assert_allclose(s["force_x"], s["force_user_x"])
    # interior y faces: v = dt * mean of the two cells (B:969-986); walls -> 0
    v = s["v"]
    expect = float(np.float32(dt)) * 0.5 * (fy[:, :-1, :] + fy[:, 1:, :])
    mask = np.ones_like(expect, bool)
    mask[3, 2:4, 3] = False                      # faces touching the solid take its velocity (0)
    np.testing.assert_allclose(v[:, 1:-1, :][mask], expect[mask], rtol=1e-12, atol=1e-15)
    assert np.all(v[3, 3:5, 3] == 0.0)
    assert np.all(v[:, 0, :] == 0.0) and np.all(v[:, -1, :] == 0.0)


def test_enforce_outflow_periodic_and_solid_faces(oracle_mod):
    flow = [(abi.NO_SLIP_WALL, 0.4, 0, 0, 0), O, P, P, W, (abi.FREE_SLIP_WALL, 0, 0, -0.3, 0)]
    cfg = abi.make_config(5, 4, 6, flow=flow)
    rng = np.random.default_rng(6)
    s = orc.OracleState(cfg)
    occ = np.zeros(abi.cell_shape(cfg), np.uint8)
    occ[2, 1, 2] = 1
    occ[4, 3, 3] = 1                             # beside the outflow face? no: x = 3 of 5
    s["occupancy"] = occ
    s["solid_velocity_x"] = rng.normal(size=occ.shape)
    s["solid_velocity_y"] = rng.normal(size=occ.shape)
    for n in ("us", "vs", "ws"):
        s[n] = rng.normal(size=s[n].shape)
    us0 = s["us"].copy()
    s.run_stages(3, 3, orc.options())            # p = 0: projection applies only the boundary rules
    u, v, w = s["u"], s["v"], s["w"]
    assert np.all(u[:, :, 0] == np.float32(0.4))                 # wall x-: prescribed normal velocity
    np.testing.assert_array_equal(u[:, :, 5], u[:, :, 4])        # outflow x+: copies the final interior face
    np.testing.assert_array_equal(v[:, 4, :], v[:, 0, :])        # periodic y: face n = face 0
    assert np.all(w[0] == 0.0) and np.all(w[6] == np.float32(-0.3))
    # x face between the solid cell (2,1,2) and fluid (3,1,2) takes the solid velocity
    assert u[2, 1, 3] == s["solid_velocity_x"][2, 1, 2]
    assert u[2, 1, 2] == s["solid_velocity_x"][2, 1, 2]
    assert u[2, 1, 1] == us0[2, 1, 1]
    # y faces around (4,3,3) with periodic y: j = 3 and j = 4 (wrapped to 0)
    assert v[4, 3, 3] == s["solid_velocity_y"][4, 3, 3]
    assert v[4, 0, 3] == s["solid_velocity_y"][4, 3, 3]


def brute_nearest(occ, periodic):
    """Global search over all offsets, key (Chebyshev radius, squared distance,
    dz, dy, dx) -- the order the shell search of B:1243-1290 visits."""
    nz, ny, nx = occ.shape
    R = max(occ.shape)
    offs = np.array([(dz, dy, dx) for dz in range(-R, R + 1) for dy in range(-R, R + 1) for dx in range(-R, R + 1)])
    cheb = np.abs(offs).max(1)
    d2 = (offs ** 2).sum(1)
    order = np.lexsort((offs[:, 2], offs[:, 1], offs[:, 0], d2, cheb))
    offs, cheb = offs[order], cheb[order]
    offs, cheb = offs[cheb > 0], cheb[cheb > 0]
    out = np.arange(occ.size).reshape(occ.shape)
    for z, y, x in zip(*np.nonzero(occ)):
        best = -1
        for (dz, dy, dx) in offs:
            c = [z + dz, y + dy, x + dx]
            ok = True
            for a, n in enumerate((nz, ny, nx)):
                if periodic[2 - a]:
                    c[a] %= n
                elif not 0 <= c[a] < n:
                    ok = False
            if ok and occ[c[0], c[1], c[2]] == 0:
                best = (c[0] * ny + c[1]) * nx + c[2]
                break
        out[z, y, x] = best
    return out


@pytest.mark.parametrize("periodic", [(False, False, False), (True, False, True)])
def test_density_fill_nearest_fluid(oracle_mod, periodic):
    bc = []
    for p in periodic:
        bc += [(abi.SCALAR_PERIODIC, 0.0)] * 2 if p else [(abi.FIXED_VALUE, 0.0)] * 2
    cfg = abi.make_config(6, 5, 5, density_bc=bc)
    rng = np.random.default_rng(12)
    occ = (rng.random(abi.cell_shape(cfg)) < 0.45).astype(np.uint8)
    occ[1:4, 1:4, 1:4] = 1
    np.testing.assert_array_equal(orc.nearest_fluid(cfg, occ), brute_nearest(occ, periodic))
    full = np.ones(abi.cell_shape(cfg), np.uint8)
    assert np.all(orc.nearest_fluid(cfg, full) == -1)


def test_rest_state_is_stationary(oracle_mod):
    # zero velocity, no sources, T at ambient, rho = 0: nothing moves
    cfg = abi.make_config(8, 8, 8, ambient=0.3)
    s = orc.OracleState(cfg)
    for _ in range(2):
        st = s.step(orc.options(mode=1, tol=1e-10))
    for n in ("u", "v", "w"):
        assert np.all(s[n] == 0)
    assert np.all(s["temperature"] == np.float32(0.3)) and np.all(s["density"] == 0)
    assert st.rhs_norm == 0


def test_uniform_buoyancy_is_absorbed_by_pressure(oracle_mod):
    # a uniform body force in a closed box is a gradient: projection removes it
    cfg = abi.make_config(8, 8, 8, buoyancy_density=0.5, vorticity=0.0, cubic=False)
    s = orc.OracleState(cfg, initial_density=1.0)
    s.step(orc.options(mode=1, tol=1e-12, inner_reduction=1e-14))
    for n in ("u", "v", "w"):
        assert np.abs(s[n]).max() < 1e-10
    p = s["pressure"]
    # the pressure is a function of height only (hydrostatic layers)
    np.testing.assert_allclose(p, np.broadcast_to(p[:1, :, :1], p.shape), rtol=1e-8, atol=1e-10)
    assert np.all(np.diff(p, axis=1) < 0)


def test_plume_source_noise_statistics(oracle_mod):
    cfg = abi.make_config(16, 16, 16)
    src = abi.PlumeSource(8.0, 3.0, 8.0, 4.0, 1.0, 1.0, 0.05, 1)
    rs, ts = orc.plume_sources(cfg, src, 0)
    z, y, x = np.meshgrid(*[np.arange(16) + 0.5] * 3, indexing="ij")
    inside = (x - 8) ** 2 + (y - 3) ** 2 + (z - 8) ** 2 <= 16
    assert np.array_equal(rs > 0, inside) and np.all(rs[inside] == 1.0)
    noise = ts[inside] - 1.0
    assert np.abs(noise).max() <= 0.05 + 1e-7 and np.abs(noise.mean()) < 0.01
    rs2, ts2 = orc.plume_sources(cfg, src, 1)
    assert not np.array_equal(ts, ts2)            # regenerated every step
