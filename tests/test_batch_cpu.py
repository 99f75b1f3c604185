"""Host logic of the batched plumes: the lane assignment rule exported by
the library (smoke_simulation_ext_batch_lanes) against a restatement of the
rule and against brute force -- longest-first greedy is within 4/3 - 1/(3m)
of the best makespan over all assignments -- plus its argument checks; and
the plume-set generator's draws."""
import itertools

import numpy as np
import pytest

import batch_plumes as bp
import smoke_abi as abi


def _lpt(cells, lanes):
    order = sorted(range(len(cells)), key=lambda i: (-cells[i], i))
    load = [0] * lanes
    out = [0] * len(cells)
    for i in order:
        best = min(range(lanes), key=lambda l: (load[l], l))
        out[i] = best
        load[best] += cells[i]
    return out


def _makespan(cells, lane, lanes):
    load = [0] * lanes
    for c, l in zip(cells, lane):
        load[l] += c
    return max(load)


@pytest.mark.parametrize("seed", range(6))
def test_lane_rule_matches_restatement(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 70))
    cells = [int(c) for c in rng.choice([32 ** 3, 48 ** 3, 64 ** 3, 96 * 64 * 64, 128 * 64 * 64, 96 ** 3, 128 ** 3], size=n)]
    for lanes in (1, 2, 3, 4, 8, 16):
        assert abi.batch_lanes(cells, lanes) == _lpt(cells, lanes)


@pytest.mark.parametrize("seed", range(4))
def test_lane_rule_makespan_bound_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    cells = [int(c) for c in rng.integers(1, 50, size=7)]
    for lanes in (2, 3):
        got = _makespan(cells, abi.batch_lanes(cells, lanes), lanes)
        best = min(_makespan(cells, a, lanes) for a in itertools.product(range(lanes), repeat=len(cells)))
        assert got * 3 * lanes <= best * (4 * lanes - 1), (got, best)


def test_lane_rule_ties_and_arguments():
    assert abi.batch_lanes([5, 5, 5, 5], 2) == [0, 1, 0, 1]
    assert abi.batch_lanes([1, 9], 4) == [1, 0]
    assert abi.batch_lanes([], 4) == []
    for bad in (0, 17, -1):
        with pytest.raises(abi.SmokeError):
            abi.batch_lanes([1, 2], bad)


def test_plume_set_draws():
    plumes = bp.plume_set(64, seed=3)
    assert len(plumes) == 64
    sizes = {(c.nx, c.ny, c.nz) for c, _ in plumes}
    assert sizes <= set(bp.SIZES) and len(sizes) >= 5
    for cfg, src in plumes:
        assert 3.0 <= src.radius <= 12.0
        assert src.radius <= src.center_x <= cfg.nx - src.radius and src.radius <= src.center_z <= cfg.nz - src.radius
        assert 0.0 <= src.center_y <= 0.1 * cfg.ny
        assert 0.5 <= src.temperature_rate <= 2.0 and 0.0 <= cfg.vorticity_confinement <= 4.0
    assert bp.plume_set(64, seed=3)[5][1].center_x == plumes[5][1].center_x
