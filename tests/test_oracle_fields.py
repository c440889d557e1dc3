"""Pins the oracle's field functions (fdtd2, add_current, binomial filters,
halo copy / J_exchange) with closed forms and the reference's own test
behaviour.  CPU only."""

import numpy as np
import pytest

from helpers import HALO, periodic_div_staggered_B, rel_err

CFL = 0.45


def one_tile(orc, n=(8, 8, 8), prec=32, filt="binomial2"):
    return orc.OracleWorld((1, 1, 1), n, CFL, [], [], prec=prec, current_filter=filt)


def affine(n, c0, grad, shift):
    """c0 + grad . (x + shift) on an [n] interior grid (integer values)."""
    i, j, k = np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij")
    return c0 + grad[0] * (i + shift[0]) + grad[1] * (j + shift[1]) + grad[2] * (k + shift[2])


def test_push_half_b_affine_curl_exact(orc):
    # tests/py/test_emf_fdtd2.py:44-186 checks the sign; here the value is exact:
    # the fdtd2 difference of an affine field is its gradient, so
    # B += dt (dEy/dz - dEz/dy, dEz/dx - dEx/dz, dEx/dy - dEy/dx), dt = float32(cfl/2)
    n = (8, 7, 9)
    w = one_tile(orc, n)
    gE = [(0, 2, -1), (3, 0, 1), (-2, 1, 0)]
    F = np.zeros((9,) + n)
    for c in range(3):
        F[c] = affine(n, 1 + c, gE[c], (0, 0, 0))
    w.set_fields(0, F)
    w.push_half_b(0)
    out = w.get_fields_halo(0)
    dt = np.float32(CFL / 2)
    exp = [gE[1][2] - gE[2][1], gE[2][0] - gE[0][2], gE[0][1] - gE[1][0]]
    sl = tuple(slice(HALO, HALO + v - 1) for v in n)  # the +1 neighbour must be interior
    for c in range(3):
        got = out[3 + c][sl]
        assert np.all(got == np.float32(dt * np.float32(exp[c]))), (c, np.unique(got))


def test_push_e_affine_curl_exact(orc):
    n = (9, 8, 7)
    w = one_tile(orc, n)
    gB = [(0, 1, 2), (-1, 0, 3), (2, -2, 0)]
    F = np.zeros((9,) + n)
    for c in range(3):
        F[3 + c] = affine(n, 5 - c, gB[c], (0, 0, 0))
    w.set_fields(0, F)
    w.push_e(0)
    out = w.get_fields_halo(0)
    dt = np.float32(CFL)
    # E += dt (dBz/dy - dBy/dz, dBx/dz - dBz/dx, dBy/dx - dBx/dy)
    exp = [gB[2][1] - gB[1][2], gB[0][2] - gB[2][0], gB[1][0] - gB[0][1]]
    sl = tuple(slice(HALO + 1, HALO + v) for v in n)  # the -1 neighbour must be interior
    for c in range(3):
        got = out[c][sl]
        assert np.all(got == np.float32(dt * np.float32(exp[c]))), (c, np.unique(got))


def _periodic_world(orc, nt=(2, 2, 2), n=(6, 6, 6), prec=32, filt="binomial2"):
    return orc.OracleWorld(nt, n, CFL, [], [], prec=prec, current_filter=filt)


def _set_global(w, G):
    """G: [9, Nx, Ny, Nz] global interior arrays -> tiles."""
    nx, ny, nz = w.n
    for i in range(w.n_tiles[0]):
        for j in range(w.n_tiles[1]):
            for k in range(w.n_tiles[2]):
                w.set_fields(w.tile_index(i, j, k),
                             G[:, i * nx:(i + 1) * nx, j * ny:(j + 1) * ny, k * nz:(k + 1) * nz])


def test_div_b_preserved(orc):
    rng = np.random.default_rng(1)
    w = _periodic_world(orc)
    N = tuple(a * b for a, b in zip(w.n_tiles, w.n))
    G = np.zeros((9,) + N)
    G[0:6] = rng.standard_normal((6,) + N)
    _set_global(w, G)
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    g0 = w.global_fields()
    d0 = periodic_div_staggered_B(g0[3], g0[4], g0[5])
    for _ in range(6):
        w.push_half_b()
        w.comm_local("emf_B")
        w.push_e()
        w.comm_local("emf_E")
    g1 = w.global_fields()
    d1 = periodic_div_staggered_B(g1[3], g1[4], g1[5])
    assert np.max(np.abs(g1[3:6] - g0[3:6])) > 0.1  # B did change
    assert np.max(np.abs(d1 - d0)) < 2e-5


def test_vacuum_dispersion_relation(orc):
    # sin^2(omega/2) = cfl^2 sin^2(k/2) for a plane wave along x (dt = dx = 1)
    nt, n = (4, 1, 1), (8, 3, 3)
    w = orc.OracleWorld(nt, n, CFL, [], [], prec=32)
    N = nt[0] * n[0]
    m = 3
    k = 2 * np.pi * m / N
    G = np.zeros((9, N, 3, 3))
    x = np.arange(N)
    G[1] = np.cos(k * x)[:, None, None]  # Ey at (i, j+1/2, k)
    nx = n[0]
    for t in range(nt[0]):
        w.set_fields(t, G[:, t * nx:(t + 1) * nx])
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    hist = []
    for _ in range(60):
        w.push_half_b()
        w.comm_local("emf_B")
        w.push_half_b()
        w.comm_local("emf_B")
        w.push_e()
        w.comm_local("emf_E")
        hist.append(w.global_fields()[1][:, 1, 1].copy())
    H = np.array(hist)  # [t, x]
    # E_{n+1} + E_{n-1} = 2 cos(omega) E_n for the eigenmode: least squares over all points
    num = np.sum((H[2:] + H[:-2]) * H[1:-1])
    den = np.sum(2 * H[1:-1] ** 2)
    cosw = num / den
    omega = np.arccos(cosw)
    expected = 2 * np.arcsin(CFL * np.sin(k / 2))
    assert abs(omega - expected) / expected < 1e-4


@pytest.mark.parametrize("filt", ["binomial2", "binomial2_unrolled"])
def test_filter_preserves_constant(orc, filt):
    # tests/py/test_emf_current_filter_binomial2.py:49-80
    w = _periodic_world(orc, filt=filt)
    N = tuple(a * b for a, b in zip(w.n_tiles, w.n))
    G = np.zeros((9,) + N)
    for c in range(3):
        G[6 + c] = c + 1
    _set_global(w, G)
    w.comm_local("emf_J")
    w.filter_current()
    g = w.global_fields()
    for c in range(3):
        assert np.all(g[6 + c] == c + 1)


@pytest.mark.parametrize("filt", ["binomial2", "binomial2_unrolled"])
def test_filter_annihilates_nyquist(orc, filt):
    # checkerboard: strictly between 0 and 1 in the reference test; (1,2,1) kills it exactly
    w = _periodic_world(orc, filt=filt)
    N = tuple(a * b for a, b in zip(w.n_tiles, w.n))
    i, j, k = np.meshgrid(*(np.arange(v) for v in N), indexing="ij")
    G = np.zeros((9,) + N)
    G[6] = (-1.0) ** (i + j + k)
    _set_global(w, G)
    w.comm_local("emf_J")
    w.filter_current()
    assert np.all(w.global_fields()[6] == 0.0)


@pytest.mark.parametrize("filt", ["binomial2", "binomial2_unrolled"])
def test_filter_transfer_function(orc, filt):
    w = _periodic_world(orc, nt=(2, 2, 2), n=(8, 8, 8), filt=filt, prec=64)
    N = tuple(a * b for a, b in zip(w.n_tiles, w.n))
    i, j, k = np.meshgrid(*(np.arange(v, dtype=np.float64) for v in N), indexing="ij")
    kv = [2 * np.pi * 1 / N[0], 2 * np.pi * 3 / N[1], 2 * np.pi * 2 / N[2]]
    G = np.zeros((9,) + N)
    G[7] = np.cos(kv[0] * i + kv[1] * j + kv[2] * k)
    _set_global(w, G)
    w.comm_local("emf_J")
    w.filter_current()
    gain = np.prod([np.cos(v / 2) ** 2 for v in kv])
    assert rel_err(w.global_fields()[7], gain * G[7]) < 1e-12


def test_add_current(orc):
    rng = np.random.default_rng(3)
    n = (5, 6, 7)
    w = one_tile(orc, n)
    F = rng.standard_normal((9,) + n).astype(np.float32).astype(np.float64)
    w.set_fields(0, F)
    w.add_current(0)
    out = w.get_fields(0)
    for c in range(3):
        assert np.all(out[c] == (F[c].astype(np.float32) - F[6 + c].astype(np.float32)))


def j_exchange_counts(n, h=3):
    """Geometric multiplicities of J0 after J copy + J_exchange on a tile whose
    every face has a neighbour: core J0, 3-deep face slabs 2J0, edges 4J0,
    corners 8J0.  (tests/py-multirank/test_emf_simulation.py:159-179 writes
    these counts without the slab depth, so its numbers do not add up to the
    tile volume; the counts below do.)"""
    nx, ny, nz = n
    c = [v - 2 * h for v in n]
    e1 = c[0] * c[1] * c[2]
    e2 = 2 * h * (c[0] * c[1] + c[1] * c[2] + c[0] * c[2])
    e4 = 4 * h * h * (c[0] + c[1] + c[2])
    e8 = 8 * h ** 3
    assert e1 + e2 + e4 + e8 == nx * ny * nz
    return [e1, e2, e4, e8]


def test_j_exchange_multiplicity(orc):
    # tests/py-multirank/test_emf_simulation.py:116-215 (2x2x2 tiles of 10x11x13)
    nt, n = (2, 2, 2), (10, 11, 13)
    w = orc.OracleWorld(nt, n, CFL, [], [], prec=32)
    J0 = np.array([1.0, 2.0, 3.0])
    for t in range(w.ntiles):
        F = np.zeros((9,) + n)
        for c in range(3):
            F[6 + c] = J0[c]
        w.set_fields(t, F)
    w.comm_local("emf_J")
    w.comm_local("emf_J_exchange")
    expected = j_exchange_counts(n)
    for t in range(w.ntiles):
        J = w.get_fields(t)[6:9]
        for c in range(3):
            vals, counts = np.unique(J[c], return_counts=True)
            assert list(vals) == [J0[c], 2 * J0[c], 4 * J0[c], 8 * J0[c]]
            assert list(counts) == expected


def test_field_energy_closed_form(orc):
    n = (4, 5, 6)
    w = one_tile(orc, n)
    F = np.zeros((9,) + n)
    F[0], F[1], F[2] = 0.5, -0.25, 0.125
    F[3], F[4], F[5] = 1.0, 2.0, 0.0
    w.set_fields(0, F)
    e = w.energies()
    cells = np.prod(n)
    assert e[0] == pytest.approx(cells * (0.25 + 0.0625 + 0.015625) / 2, rel=1e-12)
    assert e[1] == pytest.approx(cells * 5.0 / 2, rel=1e-12)
