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
