"""Launch-shape table and autotuner (the counterpart of the reference's
occupancy-derived block choice, backend-cuda.cu:1881-1911).

The tuner times candidate tiles of the two PCG stencil passes and candidate
launches (gather-form blocks and the LDS-staged form) of the four advection
kernels on the context's own grid.  Checked here:
it leaves the simulation state untouched, the chosen shapes come from the
grid class's candidates, steps with tuned shapes agree with steps on the table
shapes (to the solve tolerance: a different PCG tile regroups the FP64
partial sums), a second context of the same class reuses the cached choice,
reset returns to the table, and set_launches applies a report's shapes and
rejects shapes that are not candidates.
"""
import numpy as np
import pytest

import smoke_abi as abi
from parity import load_gpu, make_state, rel_l2, tolerance_options

pytestmark = pytest.mark.gpu

FIELDS = ("velocity_x", "velocity_y", "velocity_z", "density", "temperature")


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    abi.load_library()
    if abi.Hip().device_count() == 0:
        pytest.skip("no GPU")


# (grid, candidate PCG shapes of its class, table shapes of passes 2 and 3)
CLASSES = {
    "float4_wide": ((256, 20, 12), {(64, 4, 2, 4), (64, 8, 2, 4), (32, 8, 2, 4), (64, 4, 4, 4)}, ((64, 4, 2, 4), (64, 8, 2, 4))),
    # nx not a multiple of 256: the table takes 32-thread (128-cell) rows
    "float4_128": ((128, 40, 24), {(64, 4, 2, 4), (64, 8, 2, 4), (32, 8, 2, 4), (64, 4, 4, 4)}, ((32, 8, 2, 4), (32, 8, 2, 4))),
    "float4_narrow": ((36, 28, 20), {(16, 16, 2, 4), (16, 8, 2, 4)}, ((16, 16, 2, 4), (16, 16, 2, 4))),
    "scalar_rows": ((45, 30, 22), {(64, 4, 2, 1), (64, 8, 2, 1), (32, 8, 2, 1)}, ((64, 4, 2, 1), (64, 4, 2, 1))),
}
# gather-form blocks (one output per thread), then the LDS-staged form
ADVECT_BLOCKS = {(64, 4, 1, 1), (32, 8, 1, 1), (16, 16, 1, 1), (32, 4, 2, 1), (64, 2, 2, 1), (128, 2, 1, 1), (8, 8, 4, 1), (8, 8, 8, 1),
                 (16, 4, 4, 1), (8, 16, 2, 1), (32, 8, 2, 2)}
ADVECT_TABLE = [(32, 8, 2, 2)] * 4


def _context(cfg, st, tolerance=True):
    ctx = abi.SmokeContext(cfg)
    load_gpu(ctx, st)
    if tolerance:
        tolerance_options(ctx)
    return ctx


@pytest.mark.parametrize("name", list(CLASSES))
def test_autotune_keeps_state_and_matches_table(name):
    dims, candidates, table = CLASSES[name]
    cfg = abi.make_config(*dims, cell_size=1.0 / dims[1], dt=0.01, vorticity=2.0, cubic=True)
    st = make_state(cfg, seed=11, solids=2)
    with _context(cfg, st) as tuned, _context(cfg, st) as plain:
        r0 = tuned.launches()
        assert (r0.pcg_curvature.tuple(), r0.pcg_update.tuple()) == table
        assert [r0.advect[k].tuple() for k in range(4)] == ADVECT_TABLE
        before = tuned.export_state()
        rep = tuned.autotune_launches()
        after = tuned.export_state()
        for f in FIELDS + ("pressure",):
            np.testing.assert_array_equal(before[f], after[f])
        assert rep.candidates_timed > 0 or rep.from_cache == 1
        assert rep.pcg_curvature.tuple() in candidates and rep.pcg_update.tuple() in candidates
        assert all(rep.advect[k].tuple() in ADVECT_BLOCKS for k in range(4))
        if rep.from_cache == 0:
            assert rep.pcg_curvature_us[1] > 0 and rep.pcg_curvature_us[1] <= rep.pcg_curvature_us[0] * 1.0001
            assert rep.pcg_update_us[1] > 0 and rep.pcg_update_us[1] <= rep.pcg_update_us[0] * 1.0001
            for k in range(4):
                assert 0 < rep.advect_us[k][1] <= rep.advect_us[k][0] * 1.0001
        for _ in range(2):
            tuned.step()
            plain.step()
        a, b = tuned.export_state(), plain.export_state()
        assert tuned.stats().relative_residual <= 1e-6
        for f in FIELDS:
            assert rel_l2(a[f], b[f]) < 1e-5, f


def test_autotune_cache_reset_and_fixed_mode_graph():
    dims = (128, 32, 16)
    cfg = abi.make_config(*dims, cell_size=1.0 / 32, dt=0.01, vorticity=2.0, cubic=True)
    st = make_state(cfg, seed=5, solids=1)
    with _context(cfg, st, tolerance=False) as a:
        ra = a.autotune_launches()
        with _context(cfg, st, tolerance=False) as b, _context(cfg, st, tolerance=False) as plain:
            rb = b.autotune_launches()
            assert rb.from_cache == 1 and rb.candidates_timed == 0
            assert (rb.pcg_curvature.tuple(), rb.pcg_update.tuple()) == (ra.pcg_curvature.tuple(), ra.pcg_update.tuple())
            assert [rb.advect[k].tuple() for k in range(4)] == [ra.advect[k].tuple() for k in range(4)]
            # fixed-iteration mode runs the captured graph, rebuilt with the tuned shapes
            b.step(3)
            plain.step(3)
            x, y = b.export_state(), plain.export_state()
            for f in FIELDS:
                assert rel_l2(x[f], y[f]) < 1e-5, f
            b.reset_launches()
            r = b.launches()
            assert (r.pcg_curvature.tuple(), r.pcg_update.tuple()) == ((32, 8, 2, 4), (32, 8, 2, 4))
            assert [r.advect[k].tuple() for k in range(4)] == ADVECT_TABLE
            b.step()


def test_set_launches_applies_and_rejects_shapes():
    dims = (128, 40, 24)
    cfg = abi.make_config(*dims, cell_size=1.0 / 40, dt=0.01, vorticity=2.0, cubic=True)
    st = make_state(cfg, seed=9, solids=2)
    with _context(cfg, st, tolerance=False) as a, _context(cfg, st, tolerance=False) as plain:
        r = a.launches()
        chosen = [(64, 4, 1, 1), (16, 16, 1, 1), (8, 8, 4, 1), (32, 8, 2, 2)]
        for k, shape in enumerate(chosen):
            r.advect[k].threads_x, r.advect[k].threads_y, r.advect[k].threads_z, r.advect[k].cells_per_thread = shape
        r.pcg_curvature.threads_x, r.pcg_curvature.threads_y, r.pcg_curvature.threads_z, r.pcg_curvature.cells_per_thread = (64, 4, 4, 4)
        a.set_launches(r)
        got = a.launches()
        assert [got.advect[k].tuple() for k in range(4)] == chosen
        assert got.pcg_curvature.tuple() == (64, 4, 4, 4) and got.pcg_update.tuple() == (32, 8, 2, 4)
        # the fixed-iteration graph is rebuilt with the new shapes; every
        # advection form gives the same bits
        a.step(2)
        plain.step(2)
        x, y = a.export_state(), plain.export_state()
        for f in FIELDS:
            assert rel_l2(x[f], y[f]) < 1e-5, f
        bad = a.launches()
        bad.advect[2].threads_x = 48
        with pytest.raises(abi.SmokeError):
            a.set_launches(bad)
        bad = a.launches()
        bad.pcg_update.threads_x, bad.pcg_update.threads_y, bad.pcg_update.threads_z, bad.pcg_update.cells_per_thread = (64, 4, 2, 1)
        with pytest.raises(abi.SmokeError):     # a scalar-row tile is not a candidate of a float4 grid
            a.set_launches(bad)
        assert [a.launches().advect[k].tuple() for k in range(4)] == chosen    # a rejected report changes nothing
