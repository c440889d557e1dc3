"""Stripe injection on the GPU tile (tests/py/test_pic_moving_injector.py:30-183
of the reference): Tile.batch_inject_in_x_stripe selects the cells whose x range
meets [x_left, x_right) and hands only their corners to the generator."""

import numpy as np
import pytest

import runko_amd as runko

pytestmark = pytest.mark.gpu


def make_config():
    config = runko.Configuration(None)
    config.n_tiles = [4, 1, 1]
    config.n_cells_per_tile = [10, 4, 4]
    config.cfl = 0.45
    config.field_propagator = "fdtd2"
    config.q0 = -1
    config.m0 = 1
    config.q1 = 1
    config.m1 = 1
    config.particle_pusher = "boris"
    config.field_interpolator = "linear_1st"
    config.current_depositer = "zigzag_1st"
    return config


def identity_pgen(x, y, z):
    return runko.pic.threeD.ParticleStateBatch(pos=(x, y, z), vel=(x, y, z))


@pytest.mark.parametrize("idx, stripe", [((0, 0, 0), (20.0, 25.0)), ((2, 0, 0), (5.0, 10.0))])
def test_stripe_no_overlap(idx, stripe):
    tile = runko.pic.threeD.Tile(idx, make_config())
    tile.batch_inject_in_x_stripe(0, identity_pgen, *stripe)
    assert len(tile.get_positions(0)[0]) == 0


@pytest.mark.parametrize("idx, stripe", [((0, 0, 0), (-1.0, 11.0)), ((1, 0, 0), (10.0, 20.0))])
def test_stripe_full_overlap(idx, stripe):
    config = make_config()
    tile = runko.pic.threeD.Tile(idx, config)
    tile.batch_inject_in_x_stripe(0, identity_pgen, *stripe)
    assert len(tile.get_positions(0)[0]) == np.prod(config.n_cells_per_tile)


def test_stripe_partial_overlap():
    config = make_config()
    tile = runko.pic.threeD.Tile((0, 0, 0), config)
    tile.batch_inject_in_x_stripe(0, identity_pgen, 3.0, 7.0)
    posx, _, _ = tile.get_positions(0)
    assert len(posx) == 4 * 4 * 4
    assert np.all(posx >= 3.0) and np.all(posx < 7.0)


def test_pgen_receives_only_stripe_coords():
    tile = runko.pic.threeD.Tile((0, 0, 0), make_config())
    received = []

    def capturing_pgen(x, y, z):
        received.append(np.array(x))
        return identity_pgen(x, y, z)

    tile.batch_inject_in_x_stripe(0, capturing_pgen, 5.0, 8.0)
    assert len(received) == 1
    xs = received[0]
    assert np.all(xs >= 5.0) and np.all(xs < 8.0)
    assert sorted(set(int(round(v)) for v in xs)) == [5, 6, 7]


def test_multiple_species():
    tile = runko.pic.threeD.Tile((0, 0, 0), make_config())
    tile.batch_inject_in_x_stripe(0, identity_pgen, 2.0, 5.0)
    tile.batch_inject_in_x_stripe(1, identity_pgen, 2.0, 5.0)
    assert len(tile.get_positions(0)[0]) == 3 * 4 * 4
    assert len(tile.get_positions(1)[0]) == 3 * 4 * 4


def test_batch_inject_to_cells_unchanged():
    config = make_config()
    a = runko.pic.threeD.Tile((0, 0, 0), config)
    b = runko.pic.threeD.Tile((0, 0, 0), config)
    a.batch_inject_to_cells(0, identity_pgen)
    b.batch_inject_in_x_stripe(0, identity_pgen, 0.0, 10.0)
    for u, v in zip(a.get_positions(0), b.get_positions(0)):
        np.testing.assert_array_equal(u, v)


def test_gap_free_injection_sequence():
    tile = runko.pic.threeD.Tile((0, 0, 0), make_config())
    beta_inj, beta_flow, stride = 0.5, 0.3, 2 * 0.45
    x_left_1, x_right_1 = 5.0, 5.0 + beta_inj * stride
    tile.batch_inject_in_x_stripe(0, identity_pgen, x_left_1, x_right_1)
    x_left_2, x_right_2 = x_right_1 - beta_flow * stride, x_right_1 + beta_inj * stride
    tile.batch_inject_in_x_stripe(0, identity_pgen, x_left_2, x_right_2)
    assert x_left_2 < x_right_1
    posx, _, _ = tile.get_positions(0)
    cells = sorted(set(int(round(p)) for p in posx))
    assert cells == list(range(int(round(x_left_1)), int(round(x_right_2))))


def test_moving_injector_fills_grid_tiles():
    # the injector over a configured grid: every lap's stripe lands in the
    # tiles it overlaps, and the particles are on the device
    config = make_config()
    config.tile_partitioning = "hilbert_curve"
    config.n_laps = 10
    grid = runko.TileGrid(config)
    for idx in grid.local_tile_indices():
        grid.add_tile(runko.pic.threeD.Tile(idx, config), idx)
    sim = grid.configure_simulation(config)
    inj = runko.MovingInjector(injloc=8.0, beta_inj=1.0, beta_flow=0.0, cfl=0.45, n_inj=2, Lx=40.0)
    sim._lap = 2
    inj.inject(sim, [(0, identity_pgen), (1, identity_pgen)], 1)
    # stripe [8, 8.9): cells 8 of tile 0 only
    counts = [t.number_of_particles(0) for t in grid.tiles()]
    assert counts == [16, 0, 0, 0]
    assert inj.injloc == pytest.approx(8.9)
    sim._lap = 4
    inj.inject(sim, [(0, identity_pgen)], 2)
    # stripe [8.9, 9.8): cell 8 and 9, twice
    assert [t.number_of_particles(0) for t in grid.tiles()] == [16 + 2 * 32, 0, 0, 0]
    assert grid.world.total_particles() == 16 + 2 * 32 + 16


def test_world_injector_fills_the_moving_stripes():
    # the world-level injector: MovingInjector's front bookkeeping, the stripe
    # filled on the GPU with ppc particles per touched cell and species
    import math

    import torch

    from runko_amd.init import WorldInjector, shock_world

    w, walls, bcs, par = shock_world(nx=128, ny=32, nz=32, tile=16, ppc=2, theta=1e-4, gamma_flow=10.0, sigma=0.1,
                                     walloc=15.0, ramp_len=40.0, ramp_factor=4.0, seed=3, injloc=60.0)
    # set-up: nothing at or beyond the front's cell ceil(60) = 60
    for s in range(2):
        for t in range(w.ntiles):
            p, _, _ = w.tile_particles(t, s)
            if p.numel():
                assert float(p[0].max()) < 60.0
    beta = math.sqrt(1 - 1 / 100.0)
    inj = WorldInjector(w, injloc=60.0, beta_flow=beta, ppc=2, theta=1e-4, gamma_flow=10.0, walloc=15.0, n_inj=5,
                        Lx=128.0)
    ref = runko.MovingInjector(injloc=60.0, beta_inj=1.0, beta_flow=beta, cfl=0.45, n_inj=5, walloc=15.0, Lx=128.0)

    class Rec:
        def __init__(self):
            self.calls = []

        def batch_inject_in_x_stripe(self, s, pgen, a, b):
            self.calls.append((a, b))

    class Sim:
        def __init__(self, lap, rec):
            self.lap, self.rec = lap, rec

        def local_tiles(self):
            return [self.rec]

    before = w.total_particles()
    for lap in range(0, 400):
        rec = Rec()
        ref.inject(Sim(lap, rec), [(0, None), (1, None)], 2)
        n = inj.inject(lap)
        if rec.calls:
            a, b = rec.calls[0]
            assert inj.stripes[-1] == (a, b)
            cells = math.ceil(b) - math.floor(a)
            assert n == cells * 32 * 32 * 2 * 2, (lap, a, b, n)
        else:
            assert n == 0
    assert not inj.moving and ref.injloc == inj.injloc and inj.events == len(inj.stripes) > 10
    assert w.total_particles() == before + inj.injected
    # the injected particles: ids above 2^39 ordinal, tile-tagged, unique;
    # one per cell and pass, species 1 on top of species 0, cold flow along -x
    got = {}
    for s in range(2):
        ps, vs, ids = [], [], []
        for t in range(w.ntiles):
            p, v, i = w.tile_particles(t, s)
            sel = (i & ((1 << 40) - 1)) >= (1 << 39)
            assert torch.all((i[sel] >> 40) == t)
            ps.append(p[:, sel].cpu().numpy())
            vs.append(v[:, sel].cpu().numpy())
            ids.append(i[sel].cpu().numpy())
        p, v, i = np.concatenate(ps, 1), np.concatenate(vs, 1), np.concatenate(ids)
        assert len(np.unique(i)) == len(i) == inj.injected // 2
        assert np.all(v[0] < 0) and abs(v[0].mean() + 10.0 * beta) < 0.01
        o = np.argsort(i)
        got[s] = (p[:, o], i[o])
        lo = min(a for a, b in inj.stripes)
        assert p[0].min() >= math.floor(lo)
    assert np.array_equal(got[0][1], got[1][1]) and np.array_equal(got[0][0], got[1][0])
    # per event and cell: exactly ppc particles per species
    cells = np.floor(got[0][0]).astype(np.int64)
    key = (cells[0] * 32 + cells[1]) * 32 + cells[2]
    cnt = np.bincount(key)
    expect = np.zeros_like(cnt)
    for a, b in inj.stripes:
        for x in range(math.floor(a), math.ceil(b)):
            expect[x * 1024:(x + 1) * 1024] += 2
    assert np.array_equal(cnt, expect[:len(cnt)]) and expect[len(cnt):].sum() == 0
