"""The benchmark configurations at their stated sizes.

* 64^3 pair plasma (4^3 tiles of 16^3 cells, 8 ppc/species, 4.19e6
  particles, theta 0.3, sigma 1; projects/pic-turbulence/pic.py conventions):
  one lap of the batched form against the FP32 oracle (particles, ids,
  tiles bit-exact; fields to 1e-6), 100 laps of energy history against it
  (0.5%), and both against the FP64 oracle build (errors printed).
* 512^3 decaying turbulence (32^3 tiles, 16 ppc/species, 4.29e9 particles,
  theta 0.3, sigma 10, 8x8 dB modes): three batched laps keep every particle
  exactly once (count, and three checksums of the dense id index: its sum,
  its sum of squares and a multiplicative hash, all mod 2^64), leave the
  error counters at zero and change the total energy by less than 1e-3.
"""

import gc

import numpy as np
import pytest
import torch

from gpu_common import build_oracle, gpu_fields, gpu_particles, oracle_particles
from runko_amd.workloads import build_simulation, pair_plasma_inputs, turbulence_lap

pytestmark = pytest.mark.gpu

CFL = 0.45


def total_energy(e, charges):
    return e[0] + e[1] + sum(abs(q) * CFL**2 * e[2 + s] for s, q in enumerate(charges))


@pytest.fixture(scope="module")
def pair_inputs():
    return pair_plasma_inputs()


def test_config1_one_lap_matches_oracle(orc, pair_inputs):
    nt, n, Q, M, tiles = pair_inputs
    conf, grid, sim = build_simulation(nt, n, Q, M, tiles, execution="batched")
    w32 = build_oracle(orc, nt, n, Q, M, tiles)
    w64 = build_oracle(orc, nt, n, Q, M, tiles, prec=64)
    sim.for_one_lap(turbulence_lap(sim))
    w32.lap(0)
    w64.lap(0)
    total = 0
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w32, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(tg, to)  # per-tile membership (so per-tile counts)
        assert np.array_equal(sg, so)  # positions (so cell indices) and momenta
        _, s64, i64 = oracle_particles(w64, s)
        assert np.array_equal(i64, ig)
        dx = np.max(np.abs(sg[:, :3] - s64[:, :3]))
        du = np.max(np.abs(sg[:, 3:] - s64[:, 3:])) / np.max(np.abs(s64[:, 3:]))
        print(f"species {s}: vs FP64 oracle max |dx| {dx:.2e} cells, max |du|/max|u| {du:.2e}")
        assert dx < 1e-4 and du < 1e-5
        total += len(ig)
    assert total == 4 * 16 ** 3 * 64 * 2
    g, o, o64 = gpu_fields(grid), w32.global_fields(), w64.global_fields()
    err = np.max(np.abs(g - o)) / np.max(np.abs(o))
    err64 = np.max(np.abs(g - o64)) / np.max(np.abs(o64))
    err32_64 = np.max(np.abs(o - o64)) / np.max(np.abs(o64))
    print(f"fields: vs FP32 oracle {err:.2e}, vs FP64 oracle {err64:.2e} (FP32 oracle vs FP64 {err32_64:.2e})")
    assert err < 1e-6
    assert err64 < 1e-5


def test_config1_energy_history_100_laps(orc, pair_inputs):
    nt, n, Q, M, tiles = pair_inputs
    conf, grid, sim = build_simulation(nt, n, Q, M, tiles, execution="batched")
    w32 = build_oracle(orc, nt, n, Q, M, tiles)
    w64 = build_oracle(orc, nt, n, Q, M, tiles, prec=64)
    step = turbulence_lap(sim)
    e0 = total_energy(grid.world.energies(), Q)
    assert total_energy(w32.energies(), Q) == pytest.approx(e0, rel=1e-6)
    worst = worst64 = worst32_64 = 0.0
    for lap in range(100):
        sim.for_one_lap(step)
        w32.lap(lap)
        w64.lap(lap)
        if lap % 10 == 9:
            eg = total_energy(grid.world.energies(), Q)
            eo = total_energy(w32.energies(), Q)
            e64 = total_energy(w64.energies(), Q)
            worst = max(worst, abs(eg - eo) / abs(eo))
            worst64 = max(worst64, abs(eg - e64) / abs(e64))
            worst32_64 = max(worst32_64, abs(eo - e64) / abs(e64))
            assert abs(eg - eo) <= 5e-3 * abs(eo), (lap, eg, eo)
            assert abs(eg - e64) <= 5e-3 * abs(e64), (lap, eg, e64)
    print(f"100-lap total energy: worst vs FP32 oracle {worst:.2e}, vs FP64 oracle {worst64:.2e} "
          f"(FP32 oracle vs FP64 {worst32_64:.2e})")
    assert grid.world.total_particles() == 4 * 16 ** 3 * 64 * 2
    assert list(grid.world.error_counters()) == [0, 0]


def _id_checksums(w, s, per_tile):
    """(count, sum d, sum d^2, sum hash(d)) mod 2^64 over the alive particles
    of species s, d = tile tag * per_tile + ordinal."""
    K = 0x9E3779B97F4A7C15 - (1 << 64)  # odd 64-bit multiplier, as a signed int64
    acc = torch.zeros(4, dtype=torch.int64, device="cuda")
    for t in range(w.ntiles):
        _, _, ids = w.tile_particles(t, s)
        d = (ids >> 40) * per_tile + (ids & ((1 << 40) - 1))
        h = (d + 1) * K
        h = h ^ (h >> 29)
        acc += torch.stack([torch.tensor(d.numel(), device="cuda"), d.sum(), (d * d).sum(), h.sum()])
    return [int(v) for v in acc.cpu()]


def _expected_checksums(n, chunk=1 << 26):
    K = 0x9E3779B97F4A7C15 - (1 << 64)
    acc = torch.zeros(3, dtype=torch.int64, device="cuda")
    for a in range(0, n, chunk):
        d = torch.arange(a, min(n, a + chunk), dtype=torch.int64, device="cuda")
        h = (d + 1) * K
        h = h ^ (h >> 29)
        acc += torch.stack([d.sum(), (d * d).sum(), h.sum()])
    return [n] + [int(v) for v in acc.cpu()]


def test_config2_512_three_laps_keep_every_particle():
    from runko_amd.init import turbulence_world
    from runko_amd.stepper import BatchedLap

    gc.collect()
    torch.cuda.empty_cache()
    w, info = turbulence_world(N=512, tile=32, ppc=16, theta=0.3, sigma=10.0, seed=1)
    per_tile = w.bpt * (w.brick[0] * w.brick[1] * w.brick[2]) * 16
    n0 = w.total_particles()
    assert n0 == 512 ** 3 * 16 * 2
    Q = [info["q0"], info["q1"]]
    e0 = total_energy(w.energies(), Q)
    lap = BatchedLap(w)
    for i in range(3):
        lap()
        lap.settle()
        print(f"lap {i}: moves {w.bin_moves}, relayouts {w.relayouts}", flush=True)
    torch.cuda.synchronize()
    e1 = total_energy(w.energies(), Q)
    assert w.total_particles() == n0
    assert list(w.error_counters()) == [0, 0]
    print(f"total energy {e0:.6e} -> {e1:.6e} ({(e1 - e0) / e0:+.2e})")
    assert abs(e1 - e0) < 1e-3 * abs(e0)
    want = _expected_checksums(n0 // 2)
    for s in range(2):
        assert _id_checksums(w, s, per_tile) == want, f"species {s}"
    del lap, w
    gc.collect()
    torch.cuda.empty_cache()
