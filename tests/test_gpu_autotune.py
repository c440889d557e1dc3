"""The brick-shape autotuner on the GPU: it times every candidate and picks
the fastest, and a grid whose configuration asks for brick_shape = "auto"
steps like the same grid with the rule's brick."""

import numpy as np
import pytest

from runko_amd.autotune import tune_brick
from gpu_common import build_gpu, canonical_lap, gpu_fields, gpu_particles, random_inputs
from helpers import rel_err

pytestmark = pytest.mark.gpu

Q = [-0.02, 0.02]
M = [1.0, 1.0]


def test_tuner_times_every_candidate_and_picks_the_fastest():
    cands = [(8, 8, 8), (4, 4, 4), (4, 4, 8)]
    best, table = tune_brick(8, ppc=2, candidates=cands, steps=2)
    assert set(table) == set(cands)
    assert all(v > 0 for v in table.values())
    assert best == min(table, key=table.get)


def test_auto_brick_grid_matches_rule_brick_grid():
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=31, ppc=2)
    res = []
    for shape in (None, "auto"):
        conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution="batched", brick_shape=shape)
        step = canonical_lap(sim)
        for _ in range(3):
            sim.for_one_lap(step)
        res.append(([gpu_particles(grid, s) for s in range(2)], gpu_fields(grid), grid.world.brick))
    (pa, fa, ba), (pb, fb, bb) = res
    assert bb in [(8, 8, 8), (4, 4, 4), (4, 4, 8), (4, 8, 4), (8, 4, 4), (4, 8, 8), (8, 4, 8), (8, 8, 4)]
    for (ta, sa, ia), (tb, sb, ib) in zip(pa, pb):
        assert np.array_equal(ia, ib) and np.array_equal(ta, tb)
        assert np.abs(sa - sb).max() < 1e-4
    assert rel_err(fb, fa) < 1e-5
