"""Batched plumes (csrc/batch.hip) through the C ABI: every plume of a batch
of uneven grids equals bitwise a context created alone from the same
description and driven the same way -- in the reference's fixed-iteration
mode (the fused step: one launch per kernel over all plumes, captured as one
graph), in the tolerance mode with Jacobi PCG (fused: per-plume convergence on
the device, gated refinement passes) and with the multilevel preconditioner
(each plume's own step in turn) -- on five small uneven grids and on the
64-plume set of the batched workload; and a network attached on two lanes is
refused."""
import numpy as np
import pytest

import batch_plumes as bp
import smoke_abi as abi
from parity import tolerance_options

pytestmark = pytest.mark.gpu

FIELDS = ("velocity_x", "velocity_y", "velocity_z", "density", "temperature", "pressure")


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    abi.load_library()
    if abi.Hip().device_count() == 0:
        pytest.skip("no GPU")


def _scenes():
    """Five plumes of uneven sizes (at most 48^3 cells).  In fixed-iteration
    mode they form three groups of the batched solve: float4-row grids with
    24 iterations, a scalar-row grid (x extent not a multiple of 4) and a
    plume with 20 iterations."""
    shapes = [(24, 24, 24), (40, 32, 24), (18, 16, 16), (48, 32, 32), (32, 24, 40)]
    iterations = [24, 24, 24, 20, 24]
    out = []
    for i, (nx, ny, nz) in enumerate(shapes):
        cfg = abi.make_config(nx, ny, nz, cell_size=1.0 / 32, dt=0.01, pressure_iterations=iterations[i], ambient=0.0, buoyancy_density=0.05,
                              buoyancy_temperature=0.25, vorticity=1.0 + 0.5 * i, cubic=i % 2 == 0)
        src = abi.PlumeSource(nx / 2 + 0.3 * i, 0.08 * ny, nz / 2 - 0.2 * i, min(nx, nz) / 6, 1.0, 0.8 + 0.2 * i, 0.05, 30 + i)
        out.append((cfg, src))
    return out


def _configure(ctx, mode):
    if mode == "tolerance":
        tolerance_options(ctx)
    elif mode == "multilevel":
        # the hierarchy first: the solver options check it against the grid
        ctx.set_multilevel_options(levels=3, pre_sweeps=2, post_sweeps=2, omega=2.0 / 3.0, coarse_solver=abi.COARSE_JACOBI, coarse_sweeps=8)
        tolerance_options(ctx, preconditioner=abi.PRECONDITIONER_MULTILEVEL)


FUSED = {"fixed": 1, "tolerance": 2, "multilevel": 0}


def _run_batch(scenes, mode, lanes, steps):
    with abi.SmokeBatch([c for c, _ in scenes], lanes=lanes) as batch:
        cells = [c.nx * c.ny * c.nz for c, _ in scenes]
        assert batch.lanes == abi.batch_lanes(cells, lanes)
        for p in batch.plumes:
            _configure(p, mode)
        for k in range(steps):
            batch.step([s for _, s in scenes], k)
        batch.synchronize()
        info = batch.info()
        got = [p.export_state() for p in batch.plumes]
        stats = [p.stats() for p in batch.plumes]
    assert info.fused == FUSED[mode] and info.steps == steps
    assert info.graph == (1 if mode == "fixed" else 0)
    if info.fused:
        print(f"batch {mode}: {info.launches_per_step} launches per step, {info.launches_per_iteration} per CG iteration", flush=True)
    return got, stats


@pytest.mark.parametrize("mode,lanes", [("fixed", 4), ("tolerance", 3), ("multilevel", 2)])
def test_batch_plumes_equal_standalone_contexts(mode, lanes):
    scenes = _scenes()
    steps = 3
    print(f"batch: mode {mode} lanes {lanes}", flush=True)
    got, stats = _run_batch(scenes, mode, lanes, steps)
    for i, (cfg, src) in enumerate(scenes):
        with abi.SmokeContext(cfg) as ctx:
            _configure(ctx, mode)
            for k in range(steps):
                ctx.plume_sources(src, k)
                ctx.step()
            ref = ctx.export_state()
            st = ctx.stats()
        for f in FIELDS:
            np.testing.assert_array_equal(got[i][f], ref[f], err_msg=f"plume {i} {f}")
        assert stats[i].iterations == st.iterations
        if mode != "fixed":
            assert st.relative_residual <= 1e-6


def test_batch_refuses_a_network_on_two_lanes():
    scenes = _scenes()[:2]
    with abi.SmokeBatch([c for c, _ in scenes], lanes=2) as batch, abi.LearnedPreconditioner(seed=0) as net:
        assert sorted(batch.lanes) == [0, 1]
        for p in batch.plumes:
            tolerance_options(p, preconditioner=abi.PRECONDITIONER_LEARNED)
            p.set_learned_preconditioner(net)
        with pytest.raises(abi.SmokeError):
            batch.step([s for _, s in scenes], 0)
        for p in batch.plumes:
            p.set_learned_preconditioner(None)
        batch.step([s for _, s in scenes], 0)
        batch.synchronize()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", ["fixed", "tolerance"])
def test_plume_set_64_fused_step_equals_lone_contexts(mode):
    """The 64-plume set of the batched workload (sizes 32^3 to 128^3, seed 3):
    two fused batch steps, each plume bitwise equal to its own context driven
    the same way."""
    plumes = bp.plume_set(64, seed=3)
    steps = 2
    print(f"plume set 64: mode {mode}", flush=True)
    got, stats = _run_batch(plumes, mode, 16, steps)
    for i, (cfg, src) in enumerate(plumes):
        with abi.SmokeContext(cfg) as ctx:
            _configure(ctx, mode)
            for k in range(steps):
                ctx.plume_sources(src, k)
                ctx.step()
            ref = ctx.export_state()
            st = ctx.stats()
        for f in FIELDS:
            np.testing.assert_array_equal(got[i][f], ref[f], err_msg=f"plume {i} {f}")
        assert stats[i].iterations == st.iterations and stats[i].refinements == st.refinements
        if mode == "tolerance":
            assert stats[i].relative_residual == st.relative_residual and st.relative_residual <= 1e-6


def test_plume_set_batch_runs_to_tolerance():
    """Twelve plumes drawn by the plume-set rule, two steps on four lanes."""
    plumes = bp.plume_set(12, seed=3)
    with abi.SmokeBatch([c for c, _ in plumes], lanes=4) as batch:
        for p in batch.plumes:
            p.set_multilevel_options(levels=4, pre_sweeps=4, post_sweeps=4, omega=2.0 / 3.0, coarse_solver=abi.COARSE_JACOBI, coarse_sweeps=8)
            tolerance_options(p, preconditioner=abi.PRECONDITIONER_MULTILEVEL)
        for k in range(2):
            batch.step([s for _, s in plumes], k)
        batch.synchronize()
        for p in batch.plumes:
            st = p.stats()
            assert st.relative_residual <= 1e-6 and st.learned_fallback == 0
            assert np.isfinite(p.export_state()["density"]).all()
