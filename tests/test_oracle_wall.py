"""Pins the oracle's reflector wall (src/runko/pic/reflector_wall.c++) and
edge boundary conditions (src/runko/emf/yee_lattice.c++:263-306,
src/runko/emf/tile.c++:808-847): the reference's own test behaviour
(tests/py/test_pic_reflector_wall.py), charge continuity of the corrected
current, and the edge-bc region rule.  CPU only."""

import numpy as np
import pytest

from helpers import HALO, cic_rho

N_CELLS = (10, 11, 13)


def wall_tile(orc, cfl=1.0, q=-1.0, prec=32):
    return orc.OracleWorld((1, 1, 1), N_CELLS, cfl, [q], [1.0], prec=prec)


def zero_fields(w):
    w.set_fields(0, np.zeros((9,) + N_CELLS))


def add(w, *states):
    w.add_particles(0, 0, np.array(states, dtype=np.float64))


def test_stationary_wall_reflects_ux(orc):
    w = wall_tile(orc)
    zero_fields(w)
    add(w, [5.5, 5.5, 6.5, -1.0, 0.0, 0.0])
    w.push_particles(0)
    w.reflect_particles(0, [(5.0, 0.0, 1.0)])
    p, _ = w.get_particles(0, 0)
    assert len(p) == 1
    assert p[0, 3] > 0 and abs(abs(p[0, 3]) - 1.0) < 1e-4
    assert p[0, 0] > 5.0
    assert abs(p[0, 4]) < 1e-5 and abs(p[0, 5]) < 1e-5


def test_no_current_behind_wall_after_reflect_and_deposit(orc):
    w = wall_tile(orc)
    zero_fields(w)
    st = [[5.5, y + 0.5, z + 0.5, -0.5, 0.0, 0.0] for y in range(N_CELLS[1]) for z in range(N_CELLS[2])]
    add(w, *st)
    w.push_particles(0)
    w.reflect_particles(0, [(5.0, 0.0, 1.0)])
    w.deposit_current(0)
    J = w.get_fields(0)[6:9]
    assert np.all(np.abs(J[:, 0:4]) < 1e-5)


def test_overflow_particle_gets_killed(orc):
    w = wall_tile(orc)
    zero_fields(w)
    add(w, [2.0, 5.5, 6.5, -0.1, 0.0, 0.0])
    w.push_particles(0)
    w.reflect_particles(0, [(5.0, 0.0, 1.0)])
    assert w.num_alive(0, 0) == 0


def test_particle_in_front_of_wall_unchanged(orc):
    w = wall_tile(orc)
    zero_fields(w)
    add(w, [7.0, 5.5, 6.5, 0.5, 0.2, -0.1])
    w.push_particles(0)
    before, _ = w.get_particles(0, 0)
    w.reflect_particles(0, [(5.0, 0.0, 1.0)])
    after, _ = w.get_particles(0, 0)
    assert np.array_equal(before, after)


def test_moving_wall_reflects_ux(orc):
    w = wall_tile(orc)
    zero_fields(w)
    beta = 0.5
    gamma = 1.0 / np.sqrt(1.0 - beta**2)
    ux0 = -0.8
    add(w, [5.5, 5.5, 6.5, ux0, 0.0, 0.0])
    w.push_particles(0)
    w.reflect_particles(0, [(5.0, beta, gamma)])
    p, _ = w.get_particles(0, 0)
    g = np.sqrt(1.0 + ux0**2)
    ux_expected = gamma**2 * g * (2.0 * beta - ux0 / g * (1.0 + beta**2))
    assert p[0, 3] == pytest.approx(ux_expected, abs=1e-3)
    assert p[0, 0] > 5.0
    assert abs(p[0, 4]) < 1e-5 and abs(p[0, 5]) < 1e-5


def test_mixed_particles_reflect_park_skip(orc):
    w = wall_tile(orc)
    zero_fields(w)
    add(w, [7.0, 5.5, 6.5, 0.5, 0.0, 0.0], [5.5, 5.5, 6.5, -1.0, 0.0, 0.0], [2.0, 5.5, 6.5, -0.1, 0.0, 0.0])
    w.push_particles(0)
    w.reflect_particles(0, [(5.0, 0.0, 1.0)])
    p, _ = w.get_particles(0, 0)
    assert len(p) == 2
    assert abs(p[0, 3] - 0.5) < 1e-4 and p[0, 0] > 5.0
    assert p[1, 3] > 0 and abs(abs(p[1, 3]) - 1.0) < 1e-4 and p[1, 0] > 5.0


@pytest.mark.parametrize("beta", [0.0, 0.3])
def test_reflected_current_conserves_charge(orc, beta):
    # particles moving along x (planar moves: the zigzag continuity is exact)
    # that cross the wall: the deposited current (regular deposit of the
    # reflected state + the +q / -q correction segments) carries the charge
    # from the pre-push position to the reflected one:
    # rho(after) - rho(before) + div J = 0 on every node, in FP64.
    w = wall_tile(orc, cfl=0.45, q=-1.0, prec=64)
    zero_fields(w)
    rng = np.random.default_rng(4)
    n = 40
    y = 1.0 + rng.random(n) * (N_CELLS[1] - 2)
    z = 1.0 + rng.random(n) * (N_CELLS[2] - 2)
    x = 5.0 + rng.random(n) * 0.3
    ux = -(0.3 + rng.random(n) * 2.0)
    st = np.stack([x, y, z, ux, np.zeros(n), np.zeros(n)], 1)
    w.add_particles(0, 0, st)

    pos0 = st[:, :3].copy()
    w.push_particles(0)
    g = 1.0 / np.sqrt(1.0 - beta**2)
    w.reflect_particles(0, [(5.0, beta, g)])
    p, _ = w.get_particles(0, 0)
    assert len(p) == n and np.all(p[:, 0] >= 5.0)
    w.deposit_current(0)
    J = w.get_fields_halo(0)[6:9]
    shape = J.shape[1:]
    lat = lambda a: a + HALO  # tile mins 0: lattice coordinate = x - (0 - 3)
    drho = cic_rho(lat(p[:, :3]), -1.0, shape) - cic_rho(lat(pos0), -1.0, shape)
    div = np.zeros(shape)
    div[1:, :, :] += J[0][1:] - J[0][:-1]
    div[:, 1:, :] += J[1][:, 1:] - J[1][:, :-1]
    div[:, :, 1:] += J[2][:, :, 1:] - J[2][:, :, :-1]
    inner = (slice(1, -1),) * 3
    assert np.max(np.abs(drho[inner] + div[inner])) < 1e-12



def _edge_region(n, side, position, direction, Lh):
    """Haloed-index ranges the reference writes for one tile at mins 0."""
    tmin, tmax = np.float32(0.0), np.float32(n[direction])
    pos = np.float32(position)
    Nd = n[direction]
    if side == 0:
        if pos <= tmin:
            return None
        wdt = Nd if pos >= tmax else int(np.float32(pos - tmin)) + 1
        rng = (0, HALO + min(wdt, Nd))
    else:
        if pos >= tmax:
            return None
        wdt = Nd if pos <= tmin else Nd - int(np.float32(pos - tmin))
        rng = (HALO + Nd - min(wdt, Nd), Lh[direction])
    return rng


@pytest.mark.parametrize("direction,side,position", [(0, 0, 4.3), (0, 1, 7.0), (1, 0, 0.0), (2, 1, -1.0),
                                                     (1, 1, 3.5), (2, 0, 20.0)])
def test_edge_bc_region(orc, direction, side, position):
    w = wall_tile(orc)
    F = np.ones((9,) + N_CELLS)
    w.set_fields(0, F)
    w.comm_local("emf_E")
    before = w.get_fields_halo(0)
    w.apply_edge_bc(0, "emf_E", direction=direction, side=side, position=position, E=(7.0, 8.0, 9.0),
                    E_components=0b101)
    after = w.get_fields_halo(0)
    Lh = before.shape[1:]
    rng = _edge_region(N_CELLS, side, position, direction, Lh)
    expect = before.copy()
    if rng is not None:
        sl = [slice(None)] * 3
        sl[direction] = slice(*rng)
        expect[0][tuple(sl)] = 7.0
        expect[2][tuple(sl)] = 9.0
    assert np.array_equal(after, expect)
