"""Differentiable periodic box on the GPU: the forward step against the
oracle and the PyTorch restatement, the hand-written adjoint against torch
autograd of the FP64 PyTorch PIC, and gradients of the differentiation
workload's loss against FP64 central finite differences.

Bars: one forward step bit-exact in particles against the FP32 oracle
(fields within 1e-6, the current's float sums differ in order); FP64
forward within 1e-12 of the PyTorch PIC; FP64 adjoint within 1e-9 of
autograd; FP32 gradients within 1e-3 (relative, normwise) of FP64 finite
differences."""

import numpy as np
import pytest
import torch

import torch_pic as tp
from runko_amd.diff import DiffBox, PairPlasmaProblem, band_magnetic_energy, rollout
from test_torch_pic import CFL, M, Q, random_box

pytestmark = pytest.mark.gpu

DEV = "cuda"


def gpu_state(F, parts, dtype):
    E = torch.tensor(F[0:3], dtype=dtype, device=DEV).contiguous()
    B = torch.tensor(F[3:6], dtype=dtype, device=DEV).contiguous()
    st = np.concatenate(parts, 0)
    x = torch.tensor(st[:, 0:3].T.copy(), dtype=dtype, device=DEV).contiguous()
    u = torch.tensor(st[:, 3:6].T.copy(), dtype=dtype, device=DEV).contiguous()
    return E, B, x, u


def make_box(n, parts, dtype, filter_passes=3):
    return DiffBox(n, Q, M, [len(p) for p in parts], cfl=CFL, filter_passes=filter_passes, dtype=dtype)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_forward_matches_oracle(orc):
    n = (8, 6, 7)
    F, parts = random_box(n, 3, seed=51)
    F = F.astype(np.float32).astype(np.float64)
    parts = [p.astype(np.float32).astype(np.float64) for p in parts]
    w = orc.OracleWorld((1, 1, 1), n, CFL, Q, M, prec=32)
    w.set_fields(0, F)
    for s, st in enumerate(parts):
        w.add_particles(0, s, st)
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    box = make_box(n, parts, torch.float32)
    s = gpu_state(F, parts, torch.float32)
    for lap in range(1, 6):
        w.lap(lap)
        s = box.step(s)
        o = w.global_fields()
        g = torch.cat([s[0], s[1]]).double().cpu().numpy()
        assert rel(g, o[0:6]) < (1e-6 if lap == 1 else 1e-5), lap
        off = 0
        for sp in range(2):
            _, so, io = w.all_particles(sp)
            so = so[np.argsort(io, kind="stable")]
            k = len(so)
            st = torch.cat([s[2][:, off:off + k], s[3][:, off:off + k]]).t().double().cpu().numpy()
            off += k
            if lap == 1:
                assert np.array_equal(st, so), sp
            else:
                assert np.abs(st - so).max() < 1e-4, (lap, sp)
    assert box.error_count() == 0


def test_forward_fp64_matches_torch_pic():
    n = (6, 7, 5)
    F, parts = random_box(n, 2, seed=52)
    box = make_box(n, parts, torch.float64)
    s = gpu_state(F, parts, torch.float64)
    counts = [len(p) for p in parts]
    q = tp.per_particle(Q, counts, torch.float64, DEV)
    qm = tp.per_particle([np.sign(a) / b for a, b in zip(Q, M)], counts, torch.float64, DEV)
    r = s
    for _ in range(4):
        s = box.step(s)
        r = tp.step(*r, q, qm, CFL)
        for a, b in zip(s, r):
            assert rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("passes", [3, 0])
def test_adjoint_step_matches_autograd(passes):
    n = (6, 5, 7)
    F, parts = random_box(n, 2, seed=53, umax=1.2, famp=0.08)
    box = make_box(n, parts, torch.float64, filter_passes=passes)
    s = gpu_state(F, parts, torch.float64)
    counts = [len(p) for p in parts]
    q = tp.per_particle(Q, counts, torch.float64, DEV)
    qm = tp.per_particle([np.sign(a) / b for a, b in zip(Q, M)], counts, torch.float64, DEV)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(5)
    cot = tuple(torch.randn(t.shape, generator=gen, device=DEV, dtype=torch.float64) for t in s)
    adj = box.step_adjoint(s, cot)
    leaves = tuple(t.clone().requires_grad_(True) for t in s)
    out = tp.step(*leaves, q, qm, CFL, filter_passes=passes)
    ref = torch.autograd.grad(out, leaves, grad_outputs=cot)
    for name, a, b in zip("EBxu", adj, ref):
        assert rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-9, name


def test_rollout_gradient_matches_autograd():
    # 7 steps, checkpoints every 3: segments of 3, 3 and 1 steps
    n = (6, 6, 5)
    F, parts = random_box(n, 2, seed=54, famp=0.06)
    box = make_box(n, parts, torch.float64)
    counts = [len(p) for p in parts]
    q = tp.per_particle(Q, counts, torch.float64, DEV)
    qm = tp.per_particle([np.sign(a) / b for a, b in zip(Q, M)], counts, torch.float64, DEV)
    s0 = gpu_state(F, parts, torch.float64)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(6)
    w = tuple(torch.randn(t.shape, generator=gen, device=DEV, dtype=torch.float64) for t in s0)
    a = tuple(t.clone().requires_grad_(True) for t in s0)
    out = rollout(box, *a, steps=7, checkpoint_every=3)
    sum((o * c).sum() for o, c in zip(out, w)).backward()
    b = tuple(t.clone().requires_grad_(True) for t in s0)
    r = b
    for _ in range(7):
        r = tp.step(*r, q, qm, CFL)
    sum((o * c).sum() for o, c in zip(r, w)).backward()
    for name, ga, gb in zip("EBxu", a, b):
        assert rel(ga.grad.cpu().numpy(), gb.grad.cpu().numpy()) < 1e-9, name


def _workload_grads(steps, every, kw, h):
    """(FP32 adjoint gradient, FP64 adjoint gradient, FP64 central differences)
    of the band-energy loss w.r.t. the 12 mode amplitudes and the drift."""
    p32 = PairPlasmaProblem(dtype=torch.float32, **kw)
    p64 = PairPlasmaProblem(dtype=torch.float64, **kw)
    amp0 = torch.linspace(-1.0, 1.0, 12, dtype=torch.float64, device=DEV)
    drift0 = torch.tensor(0.1, dtype=torch.float64, device=DEV)

    def grad(p):
        a = amp0.clone().requires_grad_(True)
        d = drift0.clone().requires_grad_(True)
        p.loss(a, d, steps, every).backward()
        return torch.cat([a.grad, d.grad.view(1)]).cpu().numpy()

    def f(a, d):
        with torch.no_grad():
            return float(p64.loss(a, d, steps, every))

    fd = np.zeros(13)
    for i in range(13):
        e = torch.zeros(13, dtype=torch.float64, device=DEV)
        e[i] = h
        fd[i] = (f(amp0 + e[:12], drift0 + e[12]) - f(amp0 - e[:12], drift0 - e[12])) / (2 * h)
    assert p32.box.error_count() == 0 and p64.box.error_count() == 0
    return grad(p32), grad(p64), fd


# The zigzag deposit is discontinuous where a particle's end point crosses a
# cell face (the relay point jumps from the midpoint to the face), so the
# loss is piecewise smooth: central differences use a step small enough that
# no particle of the rollout changes cell (h = 1e-8; with h = 1e-6 some do).


def test_workload_gradient_fp32_vs_fp64_finite_differences():
    # 16^3 box, 2 ppc, 16 steps, checkpoints every 8
    g32, g64, fd = _workload_grads(16, 8, dict(n=(16, 16, 16), ppc=2, theta=0.3, seed=12), h=1e-8)
    print(f"[gradient] FP32 adjoint vs FP64 FD {rel(g32, fd):.2e}, FP64 adjoint vs FP64 FD {rel(g64, fd):.2e}")
    assert np.linalg.norm(fd) > 0
    assert rel(g32, fd) < 1e-3
    assert rel(g64, fd) < 1e-4


def test_workload_gradient_fp64_longer_rollout():
    # 24 steps in three checkpoint segments: the FP64 adjoint against FP64 FD.
    # (The FP32 trajectory of this rollout takes the other side of a cell face
    # than the FP64 one for some particle, which moves its gradient by ~2e-3.)
    g32, g64, fd = _workload_grads(24, 8, dict(n=(16, 16, 16), ppc=2, theta=0.3, seed=12), h=1e-8)
    print(f"[gradient] 24 steps: FP64 adjoint vs FP64 FD {rel(g64, fd):.2e}, FP32 {rel(g32, fd):.2e}")
    assert rel(g64, fd) < 1e-4
    assert rel(g32, fd) < 1e-2


def test_band_energy_is_parseval_share():
    B = torch.randn((3, 12, 10, 8), dtype=torch.float64, device=DEV)
    total = band_magnetic_energy(B, 0.0, 1e9)
    assert float(total) == pytest.approx(0.5 * float((B * B).sum()), rel=1e-12)
