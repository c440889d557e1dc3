"""Pins the oracle's Gauss-law residual (the FP64 checker of the GPU
library's rk_gauss_residual) with closed forms: the cloud-in-cell charge of
a particle on a node, its image through the binomial passes, total charge,
and the discrete divergence of a periodic sine field.  CPU only."""

import numpy as np
import pytest

CFL = 0.45


def world(orc, nt=(2, 1, 2), n=(8, 6, 8), q=(0.25, -0.25), prec=64):
    return orc.OracleWorld(nt, n, CFL, list(q), [1.0] * len(q), prec=prec)


def add_one(w, tile, s, pos):
    w.add_particles(tile, s, np.array([[*pos, 0.0, 0.0, 0.0]]))


@pytest.mark.parametrize("passes", [0, 1, 2])
def test_particle_on_a_node(orc, passes):
    # a particle exactly on a node puts all its charge there; each binomial
    # pass spreads a node's value as (1,2,1)^3/64, whose centre weight is 1/8
    w = world(orc)
    add_one(w, w.tile_index(1, 0, 0), 0, (11.0, 3.0, 5.0))
    mx, rms_rho, rms_div = w.gauss_residual(passes)
    V = 16 * 6 * 16
    # F^p of a spike q is q times the p-fold separable kernel k x k x k
    k1 = np.array([1.0, 2.0, 1.0]) / 4.0
    k = np.array([1.0])
    for _ in range(passes):
        k = np.convolve(k, k1)
    assert mx == pytest.approx(0.25 * np.max(k) ** 3, rel=1e-14)
    assert rms_rho == pytest.approx(0.25 * np.sqrt(np.sum(k ** 2) ** 3 / V), rel=1e-12)
    assert rms_div == 0.0


def test_neutral_pair_and_cic_weights(orc):
    # a positron on top of an electron: rho = 0 everywhere; an electron alone
    # between nodes spreads -q over the 8 corners with trilinear weights
    w = world(orc)
    add_one(w, 0, 0, (2.25, 1.5, 3.75))
    add_one(w, 0, 1, (2.25, 1.5, 3.75))
    assert w.gauss_residual(0)[0] == 0.0
    w2 = world(orc)
    add_one(w2, 0, 0, (2.25, 1.5, 3.75))
    mx, rms_rho, _ = w2.gauss_residual(0)
    wx, wy, wz = np.array([0.75, 0.25]), np.array([0.5, 0.5]), np.array([0.25, 0.75])
    weights = np.einsum("a,b,c->abc", wx, wy, wz).ravel()
    assert mx == pytest.approx(0.25 * weights.max(), rel=1e-14)
    assert rms_rho == pytest.approx(0.25 * np.sqrt(np.sum(weights ** 2) / (16 * 6 * 16)), rel=1e-14)


def test_divergence_of_a_sine_field(orc):
    # Ex(i) = A sin(2 pi i / N): the backward difference is
    # 2 A sin(pi / N) cos(2 pi (i - 1/2) / N), whose rms is sqrt(2) A sin(pi / N)
    w = world(orc, q=())
    nt, n = (2, 1, 2), (8, 6, 8)
    N = nt[0] * n[0]
    A = 0.3
    for ti in range(nt[0]):
        for tk in range(nt[2]):
            F = np.zeros((9,) + n)
            i = np.arange(ti * n[0], (ti + 1) * n[0])
            F[0] = A * np.sin(2 * np.pi * i / N)[:, None, None]
            w.set_fields(w.tile_index(ti, 0, tk), F)
    mx, rms_rho, rms_div = w.gauss_residual(3)
    d = 2 * A * np.sin(np.pi / N) * np.cos(2 * np.pi * (np.arange(N) - 0.5) / N)
    assert rms_rho == 0.0
    assert rms_div == pytest.approx(np.sqrt(2) * A * np.sin(np.pi / N), rel=1e-12)
    assert mx == pytest.approx(np.max(np.abs(d)), rel=1e-12)
