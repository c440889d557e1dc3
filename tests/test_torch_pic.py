"""The pure-PyTorch PIC step (tests/torch_pic.py) against the oracle, on the
CPU: it is the reference of the adjoint checks, so it is pinned first.  One
periodic tile of the oracle is the same periodic box."""

import numpy as np
import pytest
import torch

import torch_pic as tp

CFL = 0.45
Q = [-0.02, 0.03]
M = [1.0, 2.0]


def random_box(n, ppc, seed, umax=0.9, famp=0.05):
    rng = np.random.default_rng(seed)
    F = np.zeros((9,) + n)
    F[0:6] = famp * rng.standard_normal((6,) + n)
    cells = np.stack(np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij"), -1).reshape(-1, 3)
    parts = []
    for s in range(2):
        pos = np.repeat(cells, ppc, 0) + rng.random((len(cells) * ppc, 3))
        u = (rng.random((len(pos), 3)) * 2 - 1) * umax
        parts.append(np.concatenate([pos, u], 1))
    return F, parts


def torch_state(F, parts, dtype=torch.float64):
    E = torch.tensor(F[0:3], dtype=dtype)
    B = torch.tensor(F[3:6], dtype=dtype)
    st = np.concatenate(parts, 0)
    x = torch.tensor(st[:, 0:3].T.copy(), dtype=dtype)
    u = torch.tensor(st[:, 3:6].T.copy(), dtype=dtype)
    counts = [len(p) for p in parts]
    q = tp.per_particle(Q, counts, dtype, "cpu")
    qm = tp.per_particle([np.sign(a) / b for a, b in zip(Q, M)], counts, dtype, "cpu")
    return E, B, x, u, q, qm


@pytest.mark.parametrize("prec", [64, 32])
def test_torch_step_matches_oracle(orc, prec):
    n = (6, 5, 7)
    F, parts = random_box(n, 2, seed=41)
    dtype = torch.float64 if prec == 64 else torch.float32
    parts_c = [p.astype(np.float32).astype(np.float64) for p in parts] if prec == 32 else parts
    F_c = F.astype(np.float32).astype(np.float64) if prec == 32 else F
    w = orc.OracleWorld((1, 1, 1), n, CFL, Q, M, prec=prec)
    w.set_fields(0, F_c)
    for s, st in enumerate(parts_c):
        w.add_particles(0, s, st)
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    E, B, x, u, q, qm = torch_state(F_c, parts_c, dtype)
    tol = 1e-12 if prec == 64 else 1e-5
    for lap in range(1, 4):
        w.lap(lap)
        E, B, x, u = tp.step(E, B, x, u, q, qm, CFL)
        o = w.global_fields()
        g = np.concatenate([E.double().numpy(), B.double().numpy()])
        assert np.linalg.norm(g - o[0:6]) <= tol * np.linalg.norm(o[0:6]), lap
        off = 0
        for s in range(2):
            _, so, io = w.all_particles(s)
            so = so[np.argsort(io, kind="stable")]
            k = len(so)
            st = torch.cat([x[:, off:off + k], u[:, off:off + k]]).t().double().numpy()
            off += k
            assert np.abs(st - so).max() <= tol * 10 * max(1.0, np.abs(so).max()), (lap, s)


def test_torch_step_is_differentiable():
    n = (5, 4, 6)
    F, parts = random_box(n, 1, seed=42)
    E, B, x, u, q, qm = torch_state(F, parts)
    for t in (E, B, x, u):
        t.requires_grad_(True)
    out = tp.step(E, B, x, u, q, qm, CFL)
    loss = sum((o * o).sum() for o in out)
    loss.backward()
    for t in (E, B, x, u):
        assert t.grad is not None and torch.isfinite(t.grad).all()
