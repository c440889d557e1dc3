"""A small pure-PyTorch PIC step on one periodic box (test infrastructure).

It restates the discrete step of one lap of projects/pic-turbulence/pic.py
(the oracle's lap without the sort) with torch operations only, so that
torch autograd differentiates it: the reference against which the HIP
adjoint kernels are checked.  It is itself checked against the oracle
(tests/test_torch_pic.py).  Vectorised over particles; fields [3, n0, n1, n2]
on the tile lattice's Yee stagger, particles [3, Np].

Citations (hel-astro-lab/runko v5.0.0):
  fdtd2            src/runko/emf/yee_lattice_fdtd2.c++:19-111
  interpolation    src/runko/emf/yee_lattice_interpolate_linear_1st.h:29-138
  Boris            src/runko/pic/particle_boris.h:15-62
  zigzag deposit   src/runko/pic/particle_current_zigzag_1st.c++:225-339
  binomial2        src/runko/emf/yee_lattice_current_filter_binomial2.c++:23-76
  E -= J           src/runko/emf/yee_lattice.c++:170-179
"""

import torch


def _sh(F, axis, off):
    """F[..., i + off, ...] along axis (periodic)."""
    return torch.roll(F, shifts=-off, dims=axis)


def curl_e(E):
    Ex, Ey, Ez = E[0], E[1], E[2]
    cx = (_sh(Ey, 2, 1) - Ey) - (_sh(Ez, 1, 1) - Ez)
    cy = (_sh(Ez, 0, 1) - Ez) - (_sh(Ex, 2, 1) - Ex)
    cz = (_sh(Ex, 1, 1) - Ex) - (_sh(Ey, 0, 1) - Ey)
    return torch.stack([cx, cy, cz])


def curl_b(B):
    Bx, By, Bz = B[0], B[1], B[2]
    cx = (_sh(By, 2, -1) - By) - (_sh(Bz, 1, -1) - Bz)
    cy = (_sh(Bz, 0, -1) - Bz) - (_sh(Bx, 2, -1) - Bx)
    cz = (_sh(Bx, 1, -1) - Bx) - (_sh(By, 0, -1) - By)
    return torch.stack([cx, cy, cz])


def binomial2(J):
    out = torch.zeros_like(J)
    w1 = (1.0, 2.0, 1.0)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                w = w1[a] * w1[b] * w1[c] / 64.0
                out = out + w * torch.roll(J, shifts=(-(a - 1), -(b - 1), -(c - 1)), dims=(1, 2, 3))
    return out


def node_means(E, B):
    """Field values at the cell nodes (interpolation_EB_linear_1st means)."""
    mE = torch.stack([(_sh(E[0], 0, -1) + E[0]) / 2, (_sh(E[1], 1, -1) + E[1]) / 2, (_sh(E[2], 2, -1) + E[2]) / 2])
    bx = (B[0] + (_sh(B[0], 1, -1) + (_sh(B[0], 2, -1) + _sh(_sh(B[0], 1, -1), 2, -1)))) / 4
    by = (B[1] + (_sh(B[1], 0, -1) + (_sh(B[1], 2, -1) + _sh(_sh(B[1], 0, -1), 2, -1)))) / 4
    bz = (B[2] + (_sh(B[2], 0, -1) + (_sh(B[2], 1, -1) + _sh(_sh(B[2], 0, -1), 1, -1)))) / 4
    return mE, torch.stack([bx, by, bz])


def interpolate(E, B, x):
    n = E.shape[1:]
    p = x + 3.0
    i = torch.floor(p)
    d = p - i
    base = i.to(torch.int64) - 3
    mE, mB = node_means(E, B)
    flatE, flatB = mE.reshape(3, -1), mB.reshape(3, -1)

    def corner(ic, jc, kc):
        ii = (base[0] + ic) % n[0]
        jj = (base[1] + jc) % n[1]
        kk = (base[2] + kc) % n[2]
        q = (ii * n[1] + jj) * n[2] + kk
        return flatE[:, q], flatB[:, q]

    def lerp(t, A, Bv):
        return (1 - t) * A + t * Bv

    c = {(ic, jc, kc): corner(ic, jc, kc) for ic in (0, 1) for jc in (0, 1) for kc in (0, 1)}
    out = []
    for which in (0, 1):
        m = {key: v[which] for key, v in c.items()}
        c00 = lerp(d[0], m[(0, 0, 0)], m[(1, 0, 0)])
        c01 = lerp(d[0], m[(0, 0, 1)], m[(1, 0, 1)])
        c10 = lerp(d[0], m[(0, 1, 0)], m[(1, 1, 0)])
        c11 = lerp(d[0], m[(0, 1, 1)], m[(1, 1, 1)])
        c0 = lerp(d[1], c00, c10)
        c1 = lerp(d[1], c01, c11)
        out.append(lerp(d[2], c0, c1))
    return out[0], out[1]


def boris(u, EBE, EBB, qm, cfl, x):
    hq = 0.5 * qm
    v0 = u * cfl
    E0 = EBE * hq
    u0 = v0 + E0
    g = cfl / torch.sqrt(cfl * cfl + (u0 * u0).sum(0))
    bfac = hq * g
    B0 = EBB * bfac / cfl
    f = 2.0 / (1.0 + (B0 * B0).sum(0))
    u1 = (u0 + torch.linalg.cross(u0, B0, dim=0)) * f
    u2 = (u0 + torch.linalg.cross(u1, B0, dim=0)) + E0
    un = u2 / cfl
    ginv2 = cfl / torch.sqrt(cfl * cfl + (u2 * u2).sum(0))
    return un, x + (un * ginv2) * cfl


def wrap(x, n):
    L = torch.tensor(n, dtype=x.dtype, device=x.device).view(3, 1)
    return torch.where(x < 0, L + x, torch.where(x >= L, x - L, x))


def deposit(x, u, q, cfl, n):
    invgam = 1.0 / torch.sqrt(1.0 + (u * u).sum(0))
    s = cfl * invgam
    x2 = x + 3.0
    x1 = x2 - u * s
    fi1, fi2 = torch.floor(x1), torch.floor(x2)
    xr = torch.minimum(torch.minimum(fi1, fi2) + 1.0, torch.maximum(torch.maximum(fi1, fi2), 0.5 * (x1 + x2)))
    F1, F2 = (xr - x1) * q, (x2 - xr) * q
    W1, W2 = (x1 + xr) * 0.5 - fi1, (x2 + xr) * 0.5 - fi2
    N = n[0] * n[1] * n[2]
    J = [torch.zeros(N, dtype=x.dtype, device=x.device) for _ in range(3)]

    def store(fi, di, dj, dk, vals):
        a = (fi[0].to(torch.int64) - 3 + di) % n[0]
        b = (fi[1].to(torch.int64) - 3 + dj) % n[1]
        c = (fi[2].to(torch.int64) - 3 + dk) % n[2]
        idx = (a * n[1] + b) * n[2] + c
        for comp, v in enumerate(vals):
            if v is not None:
                J[comp] = J[comp].index_add(0, idx, v)

    for fi, F, W in ((fi1, F1, W1), (fi2, F2, W2)):
        Wx, Wy, Wz = W[0], W[1], W[2]
        Fx, Fy, Fz = F[0], F[1], F[2]
        store(fi, 0, 0, 0, (Fx * (1 - Wy) * (1 - Wz), Fy * (1 - Wx) * (1 - Wz), Fz * (1 - Wx) * (1 - Wy)))
        store(fi, 1, 0, 0, (None, Fy * Wx * (1 - Wz), Fz * Wx * (1 - Wy)))
        store(fi, 0, 1, 0, (Fx * Wy * (1 - Wz), None, Fz * (1 - Wx) * Wy))
        store(fi, 0, 0, 1, (Fx * (1 - Wy) * Wz, Fy * (1 - Wx) * Wz, None))
        store(fi, 0, 1, 1, (Fx * Wy * Wz, None, None))
        store(fi, 1, 0, 1, (None, Fy * Wx * Wz, None))
        store(fi, 1, 1, 0, (None, None, Fz * Wx * Wy))
    return torch.stack([j.reshape(n) for j in J])


def step(E, B, x, u, q, qm, cfl, filter_passes=3):
    """One lap.  q and qm: per-particle charge and sign(q)/m tensors [Np]."""
    n = tuple(E.shape[1:])
    half = cfl / 2
    B1 = B + half * curl_e(E)
    EBE, EBB = interpolate(E, B1, x)
    un, xn = boris(u, EBE, EBB, qm, cfl, x)
    xw = wrap(xn, n)
    J = deposit(xw, un, q, cfl, n)
    for _ in range(filter_passes):
        J = binomial2(J)
    B2 = B1 + half * curl_e(E)
    E2 = (E + cfl * curl_b(B2)) - J
    return E2, B2, xw, un


def per_particle(values, counts, dtype, device):
    return torch.cat([torch.full((c,), float(v), dtype=dtype, device=device) for v, c in zip(values, counts)])
