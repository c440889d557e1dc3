"""Brick-shape autotuner for the fused particle step.

A world's particles are binned by bricks (sub-blocks of a tile); one
workgroup of the fused step owns one brick, staging its field means and
accumulating its current in LDS.  The brick shape sets the LDS footprint,
the occupancy, how many particles leave their brick each step and how much
halo each window re-reads, so the best shape depends on the tile size, the
particles per cell and the temperature.  `tune_shape` times the fused
kernel (HIP events of the library's own profile) for every candidate brick
shape, and for 8^3 bricks both workgroup sizes (512 and 256 threads), on a
small periodic world of the same tile size, ppc and theta (2 x 2 x 2 tiles,
seeded turbulence set-up), and returns the fastest per particle.
Results are cached per (tile, ppc, theta) for the process.

`tune_fused_launch` picks, on the live world, between the two launch forms
of the fused step (one workgroup per brick, or persistent workgroups taking
bricks from per-XCD queues) by timing a lap of each.

The shape is fixed when a world lays out its bins, so callers tune before
creating the world (`TileGrid` does when the configuration sets
`brick_shape = "auto"`).
"""

import itertools

import numpy as np
import torch

from . import _lib as L

# LDS bytes of one fused workgroup: E/B node means (float, 6 per node over
# (b + 2)^3) plus the current window (double, 3 per node over (b + 3)^3);
# at most 64 KiB, so two workgroups share a CU's 160 KiB
LDS_LIMIT = 64 * 1024
SIDES = (4, 8, 16)

_CACHE = {}


def fused_lds_bytes(brick):
    m = 4 * 6 * np.prod([b + 2 for b in brick])
    c = 8 * 3 * np.prod([b + 3 for b in brick])
    return int(m + c)


def candidate_bricks(n_cells, limit=LDS_LIMIT):
    """Shapes with every side in {4, 8, 16} dividing the tile and a fused LDS
    footprint within `limit`; the cubes first."""
    out = []
    for b in itertools.product(SIDES, repeat=3):
        if all(n % v == 0 for n, v in zip(n_cells, b)) and fused_lds_bytes(b) <= limit:
            out.append(tuple(int(v) for v in b))
    out.sort(key=lambda b: (len(set(b)) != 1, -min(b), b))
    return out


def candidate_shapes(n_cells, limit=LDS_LIMIT):
    """(brick, threads per workgroup) pairs: every candidate brick with 512
    threads, and 8^3 bricks with 256 threads as well."""
    out = [(b, 512) for b in candidate_bricks(n_cells, limit)]
    if out and out[0][0] == (8, 8, 8):
        out.insert(1, ((8, 8, 8), 256))
    return out


def time_brick(tile, brick, ppc, theta, steps=3, warmup=2, seed=1, workgroup=512):
    """ms of the fused kernel per step and particles of the trial world."""
    from .init import turbulence_world
    from .stepper import BatchedLap

    w, _ = turbulence_world(N=2 * tile, tile=tile, ppc=ppc, theta=theta, sigma=1.0, seed=seed, brick=brick)
    try:
        w.set_fused_workgroup(workgroup)
        lap = BatchedLap(w)
        for _ in range(warmup):
            lap()
            lap.settle()
        torch.cuda.synchronize()
        L.check(L.lib().rk_profile_enable(w.h, 1))
        for _ in range(steps):
            lap()
            lap.settle()
        torch.cuda.synchronize()
        ms = np.zeros(9)
        cnt = np.zeros(9, dtype=np.int64)
        L.check(L.lib().rk_profile_read(w.h, ms.ctypes.data, cnt.ctypes.data))
        L.check(L.lib().rk_profile_enable(w.h, 0))
        return float(ms[0] / max(cnt[0], 1)), w.total_particles()
    finally:
        del w
        torch.cuda.empty_cache()


def tune_shape(tile, ppc=16, theta=0.3, candidates=None, steps=3):
    """(best brick, best threads per workgroup, {(brick, threads): ns per
    particle}) for cubic tiles of side `tile`; candidates: (brick, threads)
    pairs (default candidate_shapes)."""
    key = (int(tile), int(ppc), float(theta))
    if key in _CACHE and candidates is None:
        return _CACHE[key]
    table = {}
    for b, wg in candidates or candidate_shapes((tile,) * 3):
        try:
            ms, npart = time_brick(tile, b, ppc, theta, steps=steps, workgroup=wg)
        except (L.RkError, RuntimeError):
            continue  # a shape the device or the library refuses
        table[(b, wg)] = 1e6 * ms / npart
    if not table:
        raise RuntimeError(f"autotune: no brick shape ran for tile {tile}")
    best = min(table, key=table.get)
    out = (best[0], best[1], table)
    if candidates is None:
        _CACHE[key] = out
    return out


def tune_brick(tile, ppc=16, theta=0.3, candidates=None, steps=3):
    """(best brick, {brick: ns per particle}) for cubic tiles of side `tile`
    (each brick at its best workgroup size); candidates: bricks."""
    pairs = None if candidates is None else [(b, 512) for b in candidates]
    _, _, table = tune_shape(tile, ppc, theta, pairs, steps)
    per = {}
    for (b, _), v in table.items():
        per[b] = min(v, per.get(b, v))
    return min(per, key=per.get), per


def tune_fused_launch(world, lap, laps_per_mode=1):
    """Times the fused particle step of `lap` (a lap object of `world` that
    calls the fused step: BatchedLap or a fused ShockLap) in both launch
    forms, `laps_per_mode` laps each (the laps advance the simulation), and
    leaves the world on the faster.  Returns (best, {mode: ms per launch})."""
    times = {}
    for mode in ("grid", "queue"):
        world.set_fused_launch(mode)
        torch.cuda.synchronize()
        L.check(L.lib().rk_profile_enable(world.h, 1))
        for _ in range(laps_per_mode):
            lap()
            lap.settle()
        torch.cuda.synchronize()
        ms = np.zeros(9)
        cnt = np.zeros(9, dtype=np.int64)
        L.check(L.lib().rk_profile_read(world.h, ms.ctypes.data, cnt.ctypes.data))
        L.check(L.lib().rk_profile_enable(world.h, 0))
        times[mode] = float(ms[0] / max(cnt[0], 1))
    best = min(times, key=times.get)
    world.set_fused_launch(best)
    return best, times
