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
    from runko_amd.autotune import tune_shape
    b, wg, t2 = tune_shape(8, ppc=2, candidates=[((8, 8, 8), 512), ((8, 8, 8), 256)], steps=2)
    assert set(t2) == {((8, 8, 8), 512), ((8, 8, 8), 256)} and (b, wg) == min(t2, key=t2.get)


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


@pytest.mark.parametrize("walls,variant", [(False, "queue"), (True, "queue"), (False, "wg256"), (True, "wg256")])
def test_queue_launch_matches_grid_launch(walls, variant):
    # the persistent queue visits the same bricks in another order, and 256
    # threads take a brick's particles in chunks of 256 instead of 512; every
    # particle's work is independent of both (current windows summed in fixed
    # point, gathered in a fixed order), so the laps agree bit for bit
    import torch
    from runko_amd.init import shock_world, turbulence_world
    from runko_amd.stepper import BatchedLap, ShockLap
    from test_gpu_shock import world_particles

    res = []
    for mode in ("grid", variant):
        if walls:
            w, wl, bcs, _ = shock_world(nx=128, ny=32, nz=32, tile=16, ppc=2, ramp_len=40.0, seed=9)
            lap = ShockLap(w, wl, bcs)
        else:
            w, _ = turbulence_world(N=64, tile=16, ppc=4, seed=3)
            lap = BatchedLap(w)
        if mode == "queue":
            w.set_fused_launch("queue")
        elif mode == "wg256":
            w.set_fused_workgroup(256)
        for _ in range(3):
            lap()
            lap.settle()
        torch.cuda.synchronize()
        res.append(([world_particles(w, s) for s in range(2)], torch.stack([w.E, w.B]).cpu().numpy()))
        del lap, w
    (pg, fg), (pq, fq) = res
    for (sg, ig), (sq, iq) in zip(pg, pq):
        assert np.array_equal(ig, iq)
        if walls:  # the wall corrections are summed with float atomics: later laps differ in rounding
            assert np.abs(sg - sq).max() < 1e-4
        else:
            assert np.array_equal(sg, sq)
    if walls:
        assert rel_err(fq, fg) < 1e-5
    else:
        assert np.array_equal(fg, fq)
