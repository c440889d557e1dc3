"""Pins the oracle's particle functions (interpolation + Boris, zigzag deposit,
pack/append/wrap, sort, kinetic energy) with closed forms, invariants and the
reference's own test behaviour.  CPU only."""

import numpy as np
import pytest

from helpers import HALO, cic_rho, div_staggered_E, rel_err, thermal_particles

CFL = 0.45
f32 = np.float32


def tile_world(orc, n=(8, 8, 8), q=(-1.0,), m=(1.0,), prec=32, nt=(1, 1, 1)):
    return orc.OracleWorld(nt, n, CFL, list(q), list(m), prec=prec)


def uniform_fields(w, E=(0, 0, 0), B=(0, 0, 0)):
    F = np.zeros((9,) + w.n)
    for c in range(3):
        F[c] = E[c]
        F[3 + c] = B[c]
    for t in range(w.ntiles):
        w.set_fields(t, F)


# staggered sample points (tile mins 0): E_x at (i+1/2, j, k), B_x at (i, j+1/2, k+1/2), ...
_STAGGER = [(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5), (0, 0.5, 0.5), (0.5, 0, 0.5), (0.5, 0.5, 0)]


@pytest.mark.parametrize("prec,tol", [(32, 2e-5), (64, 1e-12)])
def test_interpolation_of_affine_fields_is_exact(orc, prec, tol):
    # node means of a linear field sampled on the Yee stagger equal the field at
    # the node, and trilinear weights reproduce a linear function: the gathered
    # value is f(x_p) for every component.  A wrong neighbour in a mean or a
    # swapped weight shifts it by O(slope / 2).
    n = (8, 8, 8)
    w = tile_world(orc, n=n, prec=prec)
    rng = np.random.default_rng(5)
    coef = rng.uniform(-0.2, 0.2, (6, 4))
    coef[:, 0] += 1.0
    ii, jj, kk = np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij")
    F = np.zeros((9,) + n)
    for c in range(6):
        sx, sy, sz = _STAGGER[c]
        F[c] = coef[c, 0] + coef[c, 1] * (ii + sx) + coef[c, 2] * (jj + sy) + coef[c, 3] * (kk + sz)
    w.set_fields(0, F)
    # every corner and mean stays inside the interior: cells 1 .. n-2
    pos = 1.0 + rng.random((500, 3)) * (np.array(n) - 3.0)
    pos = pos.astype(np.float32).astype(np.float64)
    w.add_particles(0, 0, np.concatenate([pos, np.zeros_like(pos)], 1))
    got = w.interpolate(0, 0)
    want = coef[:, 0][None, :] + pos @ coef[:, 1:].T
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_boris_uniform_e_response(orc, axis):
    # tests/py/test_pic_particle_pusher.py:42-140: q < 0 at rest in E along one
    # axis moves against it; at B = 0 Boris gives u' = u + qm E / cfl exactly
    # up to the interpolation of a uniform field (one rounding per lerp)
    w = tile_world(orc)
    E = [0.0, 0.0, 0.0]
    E[axis] = 0.1
    uniform_fields(w, E=E)
    rng = np.random.default_rng(axis)
    st = np.zeros((40, 6))
    st[:, :3] = 2.0 + rng.random((40, 3)) * 4.0
    w.add_particles(0, 0, st)
    w.push_particles(0)
    out, _ = w.get_particles(0, 0)
    du = out[:, 3:]
    expected = -1.0 * f32(0.1) / f32(CFL)
    assert np.all(du[:, axis] < 0)
    assert np.allclose(du[:, axis], expected, rtol=1e-6, atol=0)
    others = [a for a in range(3) if a != axis]
    assert np.all(du[:, others] == 0)
    moved = out[:, :3] - st[:, :3].astype(np.float32)
    assert np.all(moved[:, axis] < 0) and np.all(moved[:, others] == 0)


def test_boris_gyration_phase_and_norm(orc):
    # pure B: the Boris rotation angle per step is 2 atan(|B0|),
    # B0 = 0.5 qm ginv B / cfl, ginv = cfl / sqrt(cfl^2 + |u0|^2); |u| is conserved
    w = tile_world(orc, n=(32, 32, 32))
    Bz = 0.05
    uniform_fields(w, B=(0, 0, Bz))
    u = 0.7  # Larmor radius ~10 cells: the orbit stays inside the tile
    w.add_particles(0, 0, np.array([[16.3, 16.6, 16.4, u, 0.0, 0.0]]))
    w.push_particles(0)
    st, _ = w.get_particles(0, 0)
    ginv = CFL / np.sqrt(CFL**2 + (u * CFL) ** 2)
    B0 = 0.5 * (-1.0) * ginv * Bz / CFL
    angle = np.arctan2(st[0, 4], st[0, 3])
    assert angle == pytest.approx(-2 * np.arctan(B0), rel=1e-5)
    for _ in range(200):
        w.push_particles(0)
    st, _ = w.get_particles(0, 0)
    assert np.linalg.norm(st[0, 3:]) == pytest.approx(u, rel=2e-6)


def test_boris_e_cross_b_drift(orc):
    # uniform E x B: the time-averaged velocity approaches the drift E x B / B^2
    # the orbit drifts 0.018 cells/step along x for 1200 steps: start near the
    # low edge of a 32-cell tile so it stays inside (x from 4 to about 26)
    w = tile_world(orc, n=(32, 24, 24))
    E, B = (0.0, 0.002, 0.0), (0.0, 0.0, 0.05)
    uniform_fields(w, E=E, B=B)
    w.add_particles(0, 0, np.array([[4.0, 12.0, 12.0, 0.0, 0.0, 0.0]]))
    xs = []
    for _ in range(1200):
        w.push_particles(0)
        xs.append(w.get_particles(0, 0)[0][0, 0])
    vdrift = CFL * E[1] / B[2]  # (E x B)_x / B^2 in code velocity units (x per step)
    # average over whole gyro periods: fit a line to x(t)
    t = np.arange(len(xs))
    assert 3.0 < min(xs) and max(xs) < 29.0
    slope = np.polyfit(t, np.array(xs), 1)[0]
    assert slope == pytest.approx(vdrift, rel=2e-2)


def lattice_x1(pos, vel, origo, q=None):
    """x1, x2 of the zigzag in float32, same expression order as the deposit."""
    u = vel.astype(f32)
    dot = (u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]) + u[:, 2] * u[:, 2]
    invgam = f32(1) / np.sqrt(f32(1) + dot)
    s = f32(CFL) * invgam
    x2 = pos.astype(f32) - np.asarray(origo, dtype=f32)
    x1 = x2 - u * s[:, None]
    return x1, x2


def zigzag_segments(x1, x2):
    """Relay points of the zigzag in float32 (particle_current_zigzag_1st.c++:263-269)."""
    fi1, fi2 = np.floor(x1), np.floor(x2)
    a = np.minimum(fi1, fi2) + f32(1)
    b = np.maximum(np.maximum(fi1, fi2), f32(0.5) * (x1 + x2))
    xr = np.minimum(a, b)
    return fi1, fi2, xr


def continuity_defect(x1, x2, q, shape):
    """First-order zigzag with midpoint weights is charge conserving per axis
    pair but not for the trilinear CIC charge in 3D: a straight segment D in
    one cell leaves rho2 - rho1 + div J = (q/4) Dx Dy Dz (-1)^(1+a+b+c) on the
    cell's node (a,b,c) (from a2b2c2 - a1b1c1 = sum of midpoint terms
    + Da Db Dc / 4).  Returns that field for all particles (float64)."""
    fi1, fi2, xr = zigzag_segments(x1, x2)
    out = np.zeros(shape)
    for lo, d, fi in ((x1, xr - x1, fi1), (xr, x2 - xr, fi2)):
        dd = d.astype(np.float64)
        vol = q / 4 * dd[:, 0] * dd[:, 1] * dd[:, 2]
        i = fi.astype(np.int64)
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    np.add.at(out, (i[:, 0] + a, i[:, 1] + b, i[:, 2] + c), vol * (-1.0) ** (1 + a + b + c))
    return out


def test_zigzag_charge_continuity(orc):
    # rho(x2) - rho(x1) + div J equals the closed-form 3D defect of the scheme
    # on every node, up to FP32 rounding; moves along at most two axes conserve exactly
    q = -0.37
    w = tile_world(orc, q=(q,))
    rng = np.random.default_rng(7)
    st = thermal_particles(rng, 300, (2.0, 2.0, 2.0), (6.0, 6.0, 6.0), umax=3.0)
    w.add_particles(0, 0, st)
    w.deposit_current(0)
    F = w.get_fields_halo(0)
    p, _ = w.get_particles(0, 0)
    x1, x2 = lattice_x1(p[:, :3], p[:, 3:], (-3.0, -3.0, -3.0))
    shape = F.shape[1:]
    r1 = cic_rho(x1.astype(np.float64), q, shape)
    r2 = cic_rho(x2.astype(np.float64), q, shape)
    div = div_staggered_E(F[6], F[7], F[8])
    defect = continuity_defect(x1, x2, q, shape)
    assert np.max(np.abs(div)) > 0.1 * abs(q)
    assert np.max(np.abs(defect)) > 1e-4 * abs(q)  # the 3D term is real
    assert np.max(np.abs(r2 - r1 + div - defect)) < 1e-5 * abs(q)


def test_zigzag_planar_moves_conserve_charge(orc):
    q = 0.8
    w = tile_world(orc, q=(q,))
    rng = np.random.default_rng(17)
    st = thermal_particles(rng, 200, (2.0, 2.0, 2.0), (6.0, 6.0, 6.0), umax=3.0)
    st[:, 5] = 0.0  # no motion along z
    w.add_particles(0, 0, st)
    w.deposit_current(0)
    F = w.get_fields_halo(0)
    p, _ = w.get_particles(0, 0)
    x1, x2 = lattice_x1(p[:, :3], p[:, 3:], (-3.0, -3.0, -3.0))
    shape = F.shape[1:]
    res = cic_rho(x2.astype(np.float64), q, shape) - cic_rho(x1.astype(np.float64), q, shape) + \
        div_staggered_E(F[6], F[7], F[8])
    assert np.max(np.abs(res)) < 1e-5 * abs(q)


def test_zigzag_total_current(orc):
    # sum over nodes of J_c = q * sum over particles of (x2 - x1)_c
    q = 0.5
    w = tile_world(orc, q=(q,))
    rng = np.random.default_rng(8)
    st = thermal_particles(rng, 200, (2.0, 2.0, 2.0), (6.0, 6.0, 6.0), umax=2.0)
    w.add_particles(0, 0, st)
    w.deposit_current(0)
    F = w.get_fields_halo(0)
    p, _ = w.get_particles(0, 0)
    x1, x2 = lattice_x1(p[:, :3], p[:, 3:], (-3.0, -3.0, -3.0))
    for c in range(3):
        expected = q * np.sum(x2[:, c].astype(np.float64) - x1[:, c].astype(np.float64))
        assert np.sum(F[6 + c]) == pytest.approx(expected, rel=1e-5, abs=1e-6)


def test_zigzag_reference_behaviour(orc):
    # tests/py/test_pic_current_depositer_zigzag_1st.py:48-248
    def run(vel, charges=(-1.0,), extra=None):
        w = tile_world(orc, q=charges, m=[1.0] * len(charges))
        cells = np.array([[x + 0.5, y + 0.5, z + 0.5] for x in range(2, 6) for y in range(2, 6) for z in range(2, 6)])
        st = np.concatenate([cells, np.tile(vel, (len(cells), 1))], axis=1)
        w.add_particles(0, 0, st)
        if extra is not None:
            extra(w, st)
        w.deposit_current(0)
        return w.get_fields_halo(0)[6:9]

    J = run((0, 0, 0))
    assert np.all(J == 0)  # static particles
    J = run((0.1, 0, 0))
    assert np.all(J[0] <= 0) and np.any(J[0] < 0) and np.all(J[1] == 0) and np.all(J[2] == 0)
    J = run((0, -0.1, 0))
    assert np.all(J[1] >= 0) and np.any(J[1] > 0) and np.all(J[0] == 0) and np.all(J[2] == 0)
    # opposite charges on top of each other cancel (the two species are summed
    # separately, so up to float32 rounding of the partial sums)
    J = run((0.2, -0.1, 0.3), charges=(-1.0, 1.0), extra=lambda w, st: w.add_particles(0, 1, st))
    assert np.max(np.abs(J)) < 1e-6


def _ref_cancel_world(orc, vels_by_species):
    """tests/py/test_pic_current_depositer_zigzag_1st.py:10-45: tile (0,0,0) of
    10x11x13 cells, cfl = 1, q = (-1, 1, 2), one particle per cell at
    cell + 0.5 + 0.1 (sin x, cos y, sin(x+z))."""
    n = (10, 11, 13)
    w = orc.OracleWorld((4, 4, 4), n, 1.0, [-1.0, 1.0, 2.0], [1.0, 1.0, 1.0], prec=32)
    i, j, k = np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij")
    x, y, z = i.ravel(), j.ravel(), k.ravel()
    pos = np.stack([x + 0.5 + 0.1 * np.sin(x), y + 0.5 + 0.1 * np.cos(y), z + 0.5 + 0.1 * np.sin(x + z)], 1)
    for s, v in vels_by_species.items():
        w.add_particles(0, s, np.concatenate([pos, np.tile(v, (len(pos), 1))], 1))
    w.deposit_current(0)
    J = w.get_fields(0)[6:9]
    return J[:, 1:-1, 1:-1, 1:-1]  # interior_index_space: boundaries need neighbours


def test_zigzag_reference_cancellations(orc):
    v = np.array((-0.2, 0.1, 0.09))
    J = _ref_cancel_world(orc, {0: v, 1: v})  # opposite charges, same motion
    assert np.all(np.round(J, 2) == 0)
    v = np.array((0.1, -0.2, -0.03))
    J = _ref_cancel_world(orc, {0: v, 1: -v, 2: v})  # three particles: -v + (-v) + 2v
    assert np.all(np.round(J, 2) == 0)


def test_fp64_and_fp32_builds_agree(orc):
    # the same 5 laps in both precisions: positions agree to FP32 round-off growth
    out = {}
    for prec in (32, 64):
        w = orc.OracleWorld((2, 2, 2), (6, 6, 6), CFL, [-0.01, 0.01], [1.0, 1.0], prec=prec)
        rng = np.random.default_rng(11)
        for t in range(w.ntiles):
            F = np.zeros((9, 6, 6, 6))
            F[0:6] = 0.02 * rng.standard_normal((6, 6, 6, 6))
            w.set_fields(t, F)
            i, j, k = t // 4, (t // 2) % 2, t % 2
            st = thermal_particles(rng, 100, (6 * i, 6 * j, 6 * k), (6 * i + 6, 6 * j + 6, 6 * k + 6), umax=1.0)
            w.add_particles(t, 0, st)
            w.add_particles(t, 1, st)
        w.comm_local("emf_E")
        w.comm_local("emf_B")
        for lap in range(5):
            w.lap(lap)
        _, st, ids = w.all_particles(0)
        order = np.argsort(ids)
        out[prec] = (st[order], ids[order], w.global_fields())
    assert np.all(out[32][1] == out[64][1])
    assert np.max(np.abs(out[32][0][:, :3] - out[64][0][:, :3])) < 1e-4
    assert rel_err(out[32][2][0:3], out[64][2][0:3]) < 1e-3


def _wrap(x, L):
    x = np.asarray(x, dtype=f32)
    return np.where(x < 0, f32(L), f32(0)) + np.fmod(x, f32(L))


def test_migration_conservation_and_wrap(orc):
    # tests/py-multirank/test_pic_simulation.py:100-691 behaviour: every particle
    # lands in the tile that contains it, ids stay unique, coordinates that
    # leave the box are wrapped to (x<0 ? max : min) + fmod(x, L)
    nt, n = (2, 2, 2), (6, 6, 6)
    w = orc.OracleWorld(nt, n, CFL, [-0.01, 0.01], [1.0, 1.0], prec=32)
    rng = np.random.default_rng(5)
    for t in range(w.ntiles):
        w.set_fields(t, np.zeros((9,) + n))
        i, j, k = t // 4, (t // 2) % 2, t % 2
        for s in range(2):
            st = thermal_particles(rng, 150, (6 * i, 6 * j, 6 * k), (6 * i + 6, 6 * j + 6, 6 * k + 6), umax=4.0)
            w.add_particles(t, s, st)
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    w.push_particles()
    before = {}
    for s in range(2):
        _, st, ids = w.all_particles(s)
        before[s] = dict(zip(ids.tolist(), st))
    w.pack_outgoing()
    leavers = sum(len(w.get_outgoing(t)[1]) for t in range(w.ntiles))
    assert leavers > 50
    w.comm_local("pic_particle")
    Lg = 12.0
    for s in range(2):
        ts, st, ids = w.all_particles(s)
        assert len(ids) == len(before[s]) and len(np.unique(ids)) == len(ids)
        assert set(ids.tolist()) == set(before[s].keys())
        for t_, p, i_ in zip(ts, st, ids):
            x = before[s][int(i_)][:3]
            assert np.all(p[:3].astype(f32) == _wrap(x, Lg))
            ti = (t_ // 4, (t_ // 2) % 2, t_ % 2)
            for a in range(3):
                assert 6 * ti[a] <= p[a] <= 6 * ti[a] + 6


def test_pack_directions(orc):
    # direction = [x >= min] - [x < max] per axis, index 9(dx+1)+3(dy+1)+(dz+1)
    w = orc.OracleWorld((3, 3, 3), (4, 4, 4), CFL, [-1.0], [1.0], prec=32)
    t = w.tile_index(1, 1, 1)
    w.set_fields(t, np.zeros((9, 4, 4, 4)))
    pts = np.array([[3.5, 5.0, 5.0], [8.0, 5.0, 5.0], [4.0, 4.0, 4.0], [7.99, 8.0, 3.0], [5, 5, 5]])
    st = np.concatenate([pts, np.zeros((5, 3))], 1)
    w.add_particles(t, 0, st)
    w.pack_outgoing(t)
    out, ids, dirs, sp = w.get_outgoing(t)
    got = {int(i) & ((1 << 40) - 1): int(d) for i, d in zip(ids, dirs)}
    assert got == {0: 9 * 0 + 3 * 1 + 1, 1: 9 * 2 + 3 * 1 + 1, 3: 9 * 1 + 3 * 2 + 0}
    assert w.num_alive(t, 0) == 2


def test_sort_is_a_permutation_by_cell(orc):
    # tests/py/test_pic_particle_sorting.py:47-85 + the key order of pic/tile.c++:417-438
    n = (6, 7, 5)
    w = tile_world(orc, n=n)
    rng = np.random.default_rng(9)
    st = thermal_particles(rng, 500, (0, 0, 0), n, umax=1.0)
    w.add_particles(0, 0, st)
    a, ia = w.get_particles(0, 0)
    w.sort_particles(0)
    b, ib = w.get_particles(0, 0)
    assert sorted(ia.tolist()) == sorted(ib.tolist())
    ka = {int(i): tuple(r) for i, r in zip(ia, a)}
    for i, r in zip(ib, b):
        assert ka[int(i)] == tuple(r)
    cells = (b[:, :3].astype(f32) + f32(3)).astype(np.uint32)
    L = [v + 6 for v in n]
    key = (cells[:, 0] * L[1] + cells[:, 1]) * L[2] + cells[:, 2]
    assert np.all(np.diff(key.astype(np.int64)) >= 0)


def test_kinetic_energy_closed_form(orc):
    w = tile_world(orc, q=(-1.0, 1.0), m=(1.0, 1.0))
    u = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0], [1.0, 1.0, 1.0]])
    st = np.concatenate([np.full((3, 3), 4.0), u], 1)
    w.add_particles(0, 0, st)
    e = w.energies()
    expected = 0.0 + (np.sqrt(26.0) - 1) + (2.0 - 1)
    assert e[2] == pytest.approx(expected, rel=1e-6)
    assert e[3] == 0.0


def test_ids_and_dead_slot_counts(orc):
    # ids = tile_tag << 40 | ordinal (pic/tile.c++:468-481); slots include dead ones
    w = orc.OracleWorld((2, 3, 4), (4, 4, 4), CFL, [-1.0], [1.0], prec=32)
    t = w.tile_index(1, 2, 3)
    st = np.concatenate([np.array([[5.0, 9.0, 13.0]] * 4), np.zeros((4, 3))], 1)
    w.add_particles(t, 0, st)
    _, ids = w.get_particles(t, 0)
    assert list(ids) == [(t << 40) | k for k in range(4)]
    w.add_particles(t, 0, np.array([[0.5, 9.0, 13.0, 0, 0, 0]]))  # outside the tile
    w.pack_outgoing(t)
    assert w.num_alive(t, 0) == 4 and w.num_slots(t, 0) == 5
