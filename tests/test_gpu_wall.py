"""Reflector wall, edge boundary conditions and the shock lap on the GPU.

The first six tests are the reference's own reflector-wall behaviours
(tests/py/test_pic_reflector_wall.py:29-223) through the tile API; the rest
compare the HIP path with the oracle: particles bit-exact after push +
reflect, J within 1e-6 normwise relative after deposit and the exchange
(the correction current is summed with float32 atomics), edge bcs bit-exact,
and the pic-shock lap (projects/pic-shock/pic.py:250-299) in its plain and
batched forms."""

import numpy as np
import pytest

import runko_amd as runko
from gpu_common import build_gpu, build_oracle, gpu_fields, gpu_particles, oracle_particles, random_inputs
from helpers import rel_err

pytestmark = pytest.mark.gpu

P = runko.pic.threeD.ParticleState


def zero(x, y, z):
    return (0.0, 0.0, 0.0)


def make_reflector_tile(wall=None):
    """tests/py/test_pic_reflector_wall.py:9-26, with the wall registered."""
    config = runko.Configuration(None)
    config.n_tiles = [4, 4, 4]
    config.n_cells_per_tile = [10, 11, 13]
    config.cfl = 1
    config.field_propagator = "fdtd2"
    config.q0 = -1
    config.m0 = 1
    config.particle_pusher = "boris"
    config.field_interpolator = "linear_1st"
    config.current_depositer = "zigzag_1st"
    tile = runko.pic.threeD.Tile((0, 0, 0), config)
    tile.register_reflector_wall(wall or runko.pic.threeD.reflector_wall(walloc=5.0))
    tile.set_EBJ(zero, zero, zero)
    return tile


def by_id(tile, s=0):
    o = np.argsort(tile.get_ids(s), kind="stable")
    return [a[o] for a in tile.get_positions(s)], [a[o] for a in tile.get_velocities(s)]


def test_stationary_wall_reflects_ux():
    tile = make_reflector_tile()
    tile.inject(0, [P(pos=(5.5, 5.5, 6.5), vel=(-1.0, 0.0, 0.0))])
    tile.push_particles()
    tile.reflect_particles()
    (px, _, _), (vx, vy, vz) = by_id(tile)
    assert len(px) == 1
    assert vx[0] > 0 and abs(abs(vx[0]) - 1.0) < 1e-4
    assert px[0] > 5.0
    assert abs(vy[0]) < 1e-5 and abs(vz[0]) < 1e-5


def test_no_current_behind_wall_after_reflect_and_deposit():
    tile = make_reflector_tile()

    def gen(x, y, z):
        return [P(pos=(x + 0.5, y + 0.5, z + 0.5), vel=(-0.5, 0.0, 0.0))] if int(x) == 5 else []

    tile.inject_to_each_cell(0, gen)
    tile.push_particles()
    tile.reflect_particles()
    tile.deposit_current()
    _, _, J = tile.get_EBJ()
    for ix in range(4):
        for c in J:
            assert np.all(np.abs(c[ix]) < 1e-5), ix


def test_overflow_particle_gets_killed():
    tile = make_reflector_tile()
    tile.inject(0, [P(pos=(2.0, 5.5, 6.5), vel=(-0.1, 0.0, 0.0))])
    tile.push_particles()
    tile.reflect_particles()
    assert len(tile.get_ids(0)) == 0
    assert len(tile.get_positions(0)[0]) == 0
    assert tile.number_of_particles(0) == 0


def test_particle_in_front_of_wall_unchanged():
    tile = make_reflector_tile()
    tile.inject(0, [P(pos=(7.0, 5.5, 6.5), vel=(0.5, 0.2, -0.1))])
    tile.push_particles()
    pre = tile.get_positions(0) + tile.get_velocities(0)
    tile.reflect_particles()
    post = tile.get_positions(0) + tile.get_velocities(0)
    for a, b in zip(pre, post):
        assert np.array_equal(a, b)


def test_moving_wall_reflects_ux():
    beta = 0.5
    gamma = 1.0 / np.sqrt(1.0 - beta**2)
    tile = make_reflector_tile(runko.pic.threeD.reflector_wall(walloc=5.0, betawall=beta, gammawall=gamma))
    ux0 = -0.8
    tile.inject(0, [P(pos=(5.5, 5.5, 6.5), vel=(ux0, 0.0, 0.0))])
    tile.push_particles()
    tile.reflect_particles()
    (px, _, _), (vx, vy, vz) = by_id(tile)
    g = np.sqrt(1.0 + ux0**2)
    assert vx[0] == pytest.approx(gamma**2 * g * (2.0 * beta - ux0 / g * (1.0 + beta**2)), abs=1e-3)
    assert px[0] > 5.0
    assert abs(vy[0]) < 1e-5 and abs(vz[0]) < 1e-5


def test_mixed_particles_reflect_park_skip():
    tile = make_reflector_tile()
    tile.inject(0, [P(pos=(7.0, 5.5, 6.5), vel=(0.5, 0.0, 0.0)),
                    P(pos=(5.5, 5.5, 6.5), vel=(-1.0, 0.0, 0.0)),
                    P(pos=(2.0, 5.5, 6.5), vel=(-0.1, 0.0, 0.0))])
    tile.push_particles()
    tile.reflect_particles()
    (px, _, _), (vx, _, _) = by_id(tile)
    assert len(px) == 2
    assert abs(vx[0] - 0.5) < 1e-4 and px[0] > 5.0
    assert vx[1] > 0 and abs(abs(vx[1]) - 1.0) < 1e-4 and px[1] > 5.0


Q = [-0.02, 0.02]
M = [1.0, 1.0]


def wall_inputs(nt, n, seed, ux_shift=-1.5, umax=0.8, ppc=3, nsp=2):
    """Random plasma drifting towards -x (onto a wall near the left edge)."""
    inputs = random_inputs(nt, n, nsp, seed=seed, ppc=ppc, umax=umax)
    for F, parts in inputs.values():
        for st in parts:
            st[:, 3] += ux_shift
    return inputs


def assert_particles_equal(grid, w, nsp=2):
    for s in range(nsp):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert len(ig) == len(io), s
        assert np.array_equal(ig, io), s
        assert np.array_equal(tg, to), s
        assert np.array_equal(sg, so), s


@pytest.mark.parametrize("beta", [0.0, 0.25])
def test_reflect_parity_with_oracle(orc, beta):
    # particles start on both sides of the wall: some reflect, some are
    # parked and removed, the rest are untouched
    nt, n = (2, 2, 1), (8, 8, 8)
    walloc = 5.25
    inputs = wall_inputs(nt, n, seed=21)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    gam = 1.0 / np.sqrt(1.0 - beta**2)
    wall = runko.pic.threeD.reflector_wall(walloc=walloc, betawall=beta, gammawall=gam)
    for tile in grid.tiles():
        tile.register_reflector_wall(wall)
        tile.push_particles()
        tile.reflect_particles()
    w.push_particles()
    w.reflect_particles(-1, [(wall.walloc, wall.betawall, wall.gammawall)])
    n_before = sum(len(st) for _, parts in inputs.values() for st in parts)
    assert grid.world.total_particles() < n_before  # some were parked
    assert_particles_equal(grid, w)
    cm = runko.tools.comm_mode

    def lap(x):
        x.prtcl_deposit_current()
        x.comm_local(cm.emf_J_exchange)
        x.comm_local(cm.emf_J)

    sim.for_one_lap(lap)
    w.deposit_current()
    w.comm_local("emf_J_exchange")
    w.comm_local("emf_J")
    assert rel_err(gpu_fields(grid)[6:9], w.global_fields()[6:9]) < 1e-6


@pytest.mark.parametrize("mode", ["emf_E", "emf_B", "emf_J"])
def test_edge_bc_parity_with_oracle(orc, mode):
    nt, n = (3, 2, 1), (8, 6, 8)
    inputs = random_inputs(nt, n, 1, seed=22, ppc=1)
    for F, _ in inputs.values():
        F[6:9] = 0.3
    conf, grid, sim = build_gpu(nt, n, Q[:1], M[:1], inputs)
    w = build_oracle(orc, nt, n, Q[:1], M[:1], inputs)
    bcs = [dict(direction=0, side=0, position=10.5, Ex=1.0, Ey=2.0, Ez=3.0, Bx=4.0, By=5.0, Bz=6.0, Jx=7.0,
                Jy=8.0, Jz=9.0, E_components=0b110, B_components=0b001, J_components=0b111),
           dict(direction=1, side=1, position=7.2, Ey=-1.0, By=-2.0, Jz=-3.0),
           dict(direction=2, side=0, position=-1.0, Ex=5.0),      # left of the domain: no cells
           dict(direction=0, side=1, position=30.0, Bz=5.0)]      # right of the domain: no cells
    for b in bcs:
        bc = runko.emf.threeD.edge_bc(**b)
        for tile in grid.tiles():
            tile.register_edge_bc(bc)
    for tile in grid.tiles():
        tile.apply_edge_bcs(getattr(runko.tools.comm_mode, mode))
    for b in bcs:
        w.apply_edge_bc(-1, mode, direction=b["direction"], side=b["side"], position=b["position"],
                        E=[b.get(k, 0.0) for k in ("Ex", "Ey", "Ez")],
                        B=[b.get(k, 0.0) for k in ("Bx", "By", "Bz")],
                        J=[b.get(k, 0.0) for k in ("Jx", "Jy", "Jz")],
                        E_components=b.get("E_components", 7), B_components=b.get("B_components", 7),
                        J_components=b.get("J_components", 7))
    assert np.array_equal(gpu_fields(grid), w.global_fields())


def test_edge_bc_rejects_particle_mode():
    nt, n = (1, 1, 1), (8, 8, 8)
    inputs = random_inputs(nt, n, 0, seed=1)
    conf, grid, sim = build_gpu(nt, n, [], [], inputs)
    tile = grid.tiles()[0]
    with pytest.raises(RuntimeError, match="apply_edge_bc"):
        tile.apply_edge_bc(runko.emf.threeD.edge_bc(position=3.0), runko.tools.comm_mode.pic_particle)


def shock_lap(n_filter=3):
    """projects/pic-shock/pic.py:250-299 without IO and the injector."""
    cm = runko.tools.comm_mode

    def step(x):
        x.grid_push_half_b()
        x.grid_apply_edge_bcs(cm.emf_B)
        x.comm_local(cm.emf_B)
        x.prtcl_push()
        x.prtcl_reflect_particles()
        x.prtcl_pack_outgoing()
        x.comm_local(cm.pic_particle)
        x.prtcl_deposit_current()
        x.comm_local(cm.emf_J_exchange)
        x.comm_local(cm.emf_J)
        x.grid_apply_edge_bcs(cm.emf_J)
        for i in range(n_filter):
            if i > 0 and i % 3 == 0:
                x.comm_local(cm.emf_J)
            x.grid_filter_current()
        x.grid_apply_edge_bcs(cm.emf_J)
        x.grid_push_half_b()
        x.grid_apply_edge_bcs(cm.emf_B)
        x.comm_local(cm.emf_B)
        x.grid_push_e()
        x.grid_apply_edge_bcs(cm.emf_E)
        x.grid_add_current()
        x.grid_apply_edge_bcs(cm.emf_E)
        x.comm_local(cm.emf_E)
        x.prtcl_advance_reflector_walls()

    return step


def shock_world(execution, nt, n, walloc, beta, seed):
    inputs = wall_inputs(nt, n, seed=seed, ux_shift=-1.0, umax=0.3, ppc=2)
    for idx, (F, parts) in list(inputs.items()):
        F[0:3] = 0.0
        inputs[idx] = (F, [st[st[:, 0] > walloc + 0.5] for st in parts])
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution=execution)
    gam = 1.0 / np.sqrt(1.0 - beta**2)
    wall = runko.pic.threeD.reflector_wall(walloc=walloc, betawall=beta, gammawall=gam)
    bc = runko.emf.threeD.edge_bc(direction=0, side=0, position=walloc, E_components=0b110, B_components=0,
                                  J_components=0b111)
    for tile in grid.tiles():
        tile.register_reflector_wall(wall)
        tile.register_edge_bc(bc)
    return inputs, grid, sim


@pytest.mark.parametrize("beta", [0.0, 0.2])
def test_shock_lap_plain_matches_batched(beta):
    # plasma flowing onto the wall behind a conducting edge bc: nothing is
    # parked or lost, everything stays in front of the wall, and the batched
    # form gives the plain form's particles and fields
    nt, n = (3, 2, 1), (8, 8, 8)
    walloc = 4.0
    res = {}
    for ex in ("plain", "batched"):
        inputs, grid, sim = shock_world(ex, nt, n, walloc, beta, seed=23)
        n0 = grid.world.total_particles()
        for _ in range(6):
            sim.for_one_lap(shock_lap())
        assert grid.world.total_particles() == n0
        wall_now = grid.tiles()[0]._walls[0].walloc
        assert wall_now == pytest.approx(walloc + 6 * beta * 0.45, abs=1e-5)
        parts = [gpu_particles(grid, s) for s in range(2)]
        for _, sg, _ in parts:
            assert np.all(sg[:, 0] >= wall_now - 0.45 * beta - 1e-6)
        res[ex] = parts, gpu_fields(grid)
    (pp, fp), (pb, fb) = res["plain"], res["batched"]
    for (tp, sp, ip), (tb, sb, ib) in zip(pp, pb):
        assert np.array_equal(ip, ib) and np.array_equal(tp, tb)
        assert np.abs(sp - sb).max() < 1e-4
    assert rel_err(fb, fp) < 1e-5


def test_shock_lap_matches_oracle(orc):
    # one shock lap, tile by tile on the GPU and on the oracle
    nt, n = (3, 2, 1), (8, 8, 8)
    walloc = 4.0
    inputs, grid, sim = shock_world("plain", nt, n, walloc, 0.2, seed=24)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    wall = grid.tiles()[0]._walls[0]
    bc = dict(direction=0, side=0, position=walloc, E_components=0b110, B_components=0, J_components=0b111)
    sim.for_one_lap(shock_lap())

    w.push_half_b(); w.apply_edge_bc(-1, "emf_B", **bc); w.comm_local("emf_B")
    w.push_particles()
    w.reflect_particles(-1, [(walloc, wall.betawall, wall.gammawall)])
    w.pack_outgoing(); w.comm_local("pic_particle")
    w.deposit_current(); w.comm_local("emf_J_exchange"); w.comm_local("emf_J")
    w.apply_edge_bc(-1, "emf_J", **bc)
    for _ in range(3):
        w.filter_current()
    w.apply_edge_bc(-1, "emf_J", **bc)
    w.push_half_b(); w.apply_edge_bc(-1, "emf_B", **bc); w.comm_local("emf_B")
    w.push_e(); w.apply_edge_bc(-1, "emf_E", **bc)
    w.add_current(); w.apply_edge_bc(-1, "emf_E", **bc)
    w.comm_local("emf_E")
    assert_particles_equal(grid, w)
    g, o = gpu_fields(grid), w.global_fields()
    assert rel_err(g[6:9], o[6:9]) < 1e-5
    assert rel_err(g[0:6], o[0:6]) < 1e-5
