"""Shared helpers for the tests (numpy only)."""

import numpy as np

HALO = 3


def rel_err(a, b):
    """Normwise relative error max|a-b| / max|b| (elementwise relative error is
    meaningless for sums with cancellation such as J)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)


def periodic_div_staggered_B(Bx, By, Bz):
    """div B at cell centres for Bx at (i, j+1/2, k+1/2) etc., periodic."""
    return ((np.roll(Bx, -1, 0) - Bx) + (np.roll(By, -1, 1) - By) + (np.roll(Bz, -1, 2) - Bz))


def cic_rho(pos_lat, q, shape):
    """First-order (cloud-in-cell) charge of particles at lattice coordinates
    pos_lat [n,3] on integer nodes of an array of `shape` (float64)."""
    rho = np.zeros(shape)
    f = np.floor(pos_lat)
    w = pos_lat - f
    i = f.astype(np.int64)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                wt = (w[:, 0] if a else 1 - w[:, 0]) * (w[:, 1] if b else 1 - w[:, 1]) * (w[:, 2] if c else 1 - w[:, 2])
                np.add.at(rho, (i[:, 0] + a, i[:, 1] + b, i[:, 2] + c), q * wt)
    return rho


def div_staggered_E(Ex, Ey, Ez):
    """div at integer nodes of fields staggered at +1/2 (index i <-> i+1/2):
    (Ex[i]-Ex[i-1]) + (Ey[j]-Ey[j-1]) + (Ez[k]-Ez[k-1]) on a non-periodic array
    (first row/column left zero)."""
    d = np.zeros_like(Ex)
    d[1:, 1:, 1:] = ((Ex[1:, 1:, 1:] - Ex[:-1, 1:, 1:]) + (Ey[1:, 1:, 1:] - Ey[1:, :-1, 1:]) +
                     (Ez[1:, 1:, 1:] - Ez[1:, 1:, :-1]))
    return d


def thermal_particles(rng, n, box_lo, box_hi, umax=0.6):
    """Random positions inside [box_lo, box_hi) and momenta with |u_i| < umax."""
    lo = np.asarray(box_lo, dtype=np.float64)
    hi = np.asarray(box_hi, dtype=np.float64)
    pos = lo + rng.random((n, 3)) * (hi - lo)
    u = (rng.random((n, 3)) * 2 - 1) * umax
    return np.concatenate([pos, u], axis=1)
