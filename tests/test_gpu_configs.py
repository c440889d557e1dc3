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
  error counters at zero and change the total energy by less than 0.5%.
  The energy measured with E and B at the same time level is not the
  leapfrog scheme's conserved one; the two differ by O((c k dt)^2) of the
  field energy, so with the dB modes laid over a box of side N the three-lap
  change scales as 1/N^2: the FP64 oracle gives +7.5% on a 64^3 box and
  +3.1% on a 128^3 box, about +0.2% at 512^3.
* The same workload on a 64^3 sub-box (2^3 tiles of 32^3 cells, 16
  ppc/species, the modes laid over the sub-box): one lap against the FP32
  oracle (particles, ids, tiles bit-exact; fields to 1e-6) and three laps of
  energy against the FP32 and FP64 oracles.
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
    assert total == 8 * 16 ** 3 * 64 * 2
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
    assert grid.world.total_particles() == 8 * 16 ** 3 * 64 * 2
    assert list(grid.world.error_counters()) == [0, 0]


def test_config2_subbox_matches_oracle(orc):
    from runko_amd.workloads import turbulence_sample_inputs

    nt, n, Q, M, tiles = turbulence_sample_inputs()
    conf, grid, sim = build_simulation(nt, n, Q, M, tiles, execution="batched")
    w32 = build_oracle(orc, nt, n, Q, M, tiles)
    w64 = build_oracle(orc, nt, n, Q, M, tiles, prec=64)
    step = turbulence_lap(sim)
    e0 = total_energy(grid.world.energies(), Q)
    sim.for_one_lap(step)
    w32.lap(0)
    w64.lap(0)
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w32, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(tg, to)
        assert np.array_equal(sg, so)
    g, o = gpu_fields(grid), w32.global_fields()
    err = np.max(np.abs(g - o)) / np.max(np.abs(o))
    print(f"one lap: fields vs FP32 oracle {err:.2e}")
    assert err < 1e-6
    for lap in range(1, 3):
        sim.for_one_lap(step)
        w32.lap(lap)
        w64.lap(lap)
    eg = total_energy(grid.world.energies(), Q)
    eo, e64 = total_energy(w32.energies(), Q), total_energy(w64.energies(), Q)
    print(f"three laps: total energy {e0:.6e} -> GPU {eg:.6e}, FP32 oracle {eo:.6e}, FP64 oracle {e64:.6e} "
          f"(GPU vs FP64 {abs(eg - e64) / e64:.2e})")
    assert abs(eg - eo) < 1e-4 * abs(eo)
    assert abs(eg - e64) < 1e-4 * abs(e64)
    assert grid.world.total_particles() == 8 * 32 ** 3 * 16 * 2
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
    assert abs(e1 - e0) < 5e-3 * abs(e0)
    want = _expected_checksums(n0 // 2)
    for s in range(2):
        assert _id_checksums(w, s, per_tile) == want, f"species {s}"
    del lap, w
    gc.collect()
    torch.cuda.empty_cache()


def _world_particles(w, s):
    """(tile per particle, states [k, 6] float64 of the float32 values, ids) sorted by id."""
    ts, sts, ids = [], [], []
    for t in range(w.ntiles):
        pos, vel, i = w.tile_particles(t, s)
        ts.append(np.full(i.numel(), t))
        sts.append(torch.cat([pos, vel]).t().cpu().numpy().astype(np.float64))
        ids.append(i.cpu().numpy().view(np.uint64))
    ts, sts, ids = np.concatenate(ts), np.concatenate(sts), np.concatenate(ids)
    o = np.argsort(ids, kind="stable")
    return ts[o], sts[o], ids[o]


def test_config3_reduced_box_matches_oracle(orc):
    """The reflecting-wall shock set-up on a 512 x 32 x 32 box of 32^3 tiles
    (gamma 10 flow, 8 ppc/species upstream, the 4x ramp over 256 cells at the
    wall, the plasma ending at an injector front inside the last tile, so
    tile occupancy varies about 10x as in the 2048 x 128 x 128 run): three
    fused shock laps (walls and edge conditions inside) against the FP32
    oracle issuing the reference's shock lap (projects/pic-shock/pic.py:250-299)
    tile method by tile method from the same particles and fields."""
    from runko_amd.init import shock_world
    from runko_amd.stepper import ShockLap

    w, walls, bcs, prm = shock_world(nx=512, ny=32, nz=32, tile=32, ppc=8, walloc=15.0, ramp_len=256.0,
                                     injloc=492.0, seed=5)
    counts = np.array([sum(w.tile_particles(t, s)[2].numel() for s in range(2)) for t in range(w.ntiles)])
    spread = counts.max() / counts[counts > 0].min()
    print(f"tile occupancy {counts.min()}..{counts.max()} ({spread:.1f}x)")
    assert spread > 8.0
    Q = [prm["q0"], prm["q1"]]
    o = orc.OracleWorld(w.nt, w.n, CFL, Q, [1.0, 1.0], prec=32)
    F = w.F[:3].cpu().numpy().astype(np.float64)
    for t in range(w.ntiles):
        sl = w.tile_interior(t)
        o.set_fields(t, F[:, :, sl[0], sl[1], sl[2]].reshape((9,) + tuple(w.n)))
        for s in range(2):
            pos, vel, ids = w.tile_particles(t, s)
            o.add_particles(t, s, torch.cat([pos, vel]).t().cpu().numpy().astype(np.float64),
                            ids.cpu().numpy().view(np.uint64))
    o.comm_local("emf_E")
    o.comm_local("emf_B")
    wall = walls[0]
    obcs = [dict(direction=b.direction, side=b.side, position=b.position, E=tuple(b.E), B=tuple(b.B),
                 J=tuple(b.J), E_components=b.E_components, B_components=b.B_components,
                 J_components=b.J_components) for b in bcs]

    def obc(mode):
        for b in obcs:
            o.apply_edge_bc(-1, mode, **b)

    lap = ShockLap(w, walls, bcs)
    for i in range(3):
        lap()
        lap.settle()
        o.push_half_b()
        obc("emf_B")
        o.comm_local("emf_B")
        o.push_particles()
        o.reflect_particles(-1, [(wall.walloc, wall.betawall, wall.gammawall)])
        o.pack_outgoing()
        o.comm_local("pic_particle")
        o.deposit_current()
        o.comm_local("emf_J_exchange")
        o.comm_local("emf_J")
        obc("emf_J")
        for _ in range(3):
            o.filter_current()
        obc("emf_J")
        o.push_half_b()
        obc("emf_B")
        o.comm_local("emf_B")
        o.push_e()
        obc("emf_E")
        o.add_current()
        obc("emf_E")
        o.comm_local("emf_E")
        if i == 0:
            for s in range(2):
                tg, sg, ig = _world_particles(w, s)
                to, so, io = oracle_particles(o, s)
                assert np.array_equal(ig, io), s
                assert np.array_equal(tg, to), s
                assert np.array_equal(sg, so), s
    torch.cuda.synchronize()
    G = torch.stack([w.E, w.B, w.J]).cpu().numpy()
    for t in range(w.ntiles):
        sl = w.tile_interior(t)
        g = G[:, :, sl[0], sl[1], sl[2]].reshape((9,) + tuple(w.n))
        ref = o.get_fields(t)
        err = np.max(np.abs(g - ref)) / max(np.max(np.abs(ref)), 1e-30)
        assert err < 1e-5, (t, err)
    for s in range(2):
        tg, sg, ig = _world_particles(w, s)
        to, so, io = oracle_particles(o, s)
        assert np.array_equal(ig, io), s
        assert np.array_equal(tg, to), s
        dx = np.max(np.abs(sg - so))
        print(f"species {s} after 3 laps: max |state difference| {dx:.2e}")
        assert dx < 1e-4
    assert list(w.error_counters()) == [0, 0]
