"""Ensembles of differentiable periodic boxes (rk_diff_create_ensemble): one
launch per kernel over every box must give, box by box, what the box gives
alone (particles bit-exact after a step from identical inputs, fields and
adjoints to the order of the current sums), and the ensemble workload's
gradient must match central differences."""

import math

import numpy as np
import pytest
import torch

from runko_amd.diff import DiffBox, DiffEnsemble, EnsembleProblem, rollout
from helpers import rel_err

pytestmark = pytest.mark.gpu

BOXES = [
    dict(n=(8, 8, 8), charges=[-0.02, 0.02], masses=[1.0, 1.0], counts=[1024, 1024]),
    dict(n=(12, 8, 10), charges=[-0.01, 0.01, 0.005], masses=[1.0, 1.0, 4.0], counts=[900, 700, 300]),
    dict(n=(6, 6, 6), charges=[-0.03], masses=[1.0], counts=[500]),
    dict(n=(10, 10, 10), charges=[-0.02, 0.02], masses=[1.0, 1.0], counts=[0, 0]),
]


def random_state(box, n, npart, gen, dtype, umax=0.8, famp=0.05):
    dev = "cuda"
    E = (famp * torch.randn((3,) + n, generator=gen, device=dev, dtype=torch.float64)).to(dtype)
    B = (famp * torch.randn((3,) + n, generator=gen, device=dev, dtype=torch.float64)).to(dtype)
    x = torch.rand((3, npart), generator=gen, device=dev, dtype=torch.float64)
    x = (x * torch.tensor(n, device=dev, dtype=torch.float64).view(3, 1)).to(torch.float32).to(dtype)
    u = ((torch.rand((3, npart), generator=gen, device=dev, dtype=torch.float64) * 2 - 1) * umax).to(dtype)
    return E, B, x, u


def build(dtype, seed=3):
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    ens = DiffEnsemble(BOXES, cfl=0.45, filter_passes=3, dtype=dtype)
    singles, states = [], []
    for bx in BOXES:
        b = DiffBox(bx["n"], bx["charges"], bx["masses"], bx["counts"], cfl=0.45, filter_passes=3, dtype=dtype)
        singles.append(b)
        states.append(random_state(b, tuple(bx["n"]), b.np, gen, dtype))
    cat = (torch.cat([s[0].reshape(3, -1) for s in states], 1).contiguous(),
           torch.cat([s[1].reshape(3, -1) for s in states], 1).contiguous(),
           torch.cat([s[2] for s in states], 1).contiguous(),
           torch.cat([s[3] for s in states], 1).contiguous())
    return ens, singles, states, cat


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_ensemble_step_matches_single_boxes(dtype):
    ens, singles, states, cat = build(dtype)
    assert ens.n_boxes == len(BOXES) and ens.np == sum(b.np for b in singles)
    out = ens.step(cat)
    ftol = 1e-6 if dtype == torch.float32 else 1e-12
    for i, (b, st) in enumerate(zip(singles, states)):
        ref = b.step(st)
        assert torch.equal(ens.particle_view(out[2], i), ref[2]), i
        assert torch.equal(ens.particle_view(out[3], i), ref[3]), i
        assert rel_err(ens.field_view(out[0], i).cpu().numpy(), ref[0].cpu().numpy()) < ftol, i
        assert rel_err(ens.field_view(out[1], i).cpu().numpy(), ref[1].cpu().numpy()) < ftol, i
    assert ens.error_count() == 0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_ensemble_adjoint_matches_single_boxes(dtype):
    ens, singles, states, cat = build(dtype, seed=4)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    lam = tuple(torch.randn(t.shape, generator=gen, device="cuda", dtype=torch.float64).to(dtype) for t in cat)
    got = ens.step_adjoint(cat, lam)
    tol = 1e-5 if dtype == torch.float32 else 1e-11
    for i, (b, st) in enumerate(zip(singles, states)):
        lam_i = (lam[0][:, ens.c0[i]:ens.c0[i] + ens.cells[i]].reshape((3,) + ens.ns[i]).contiguous(),
                 lam[1][:, ens.c0[i]:ens.c0[i] + ens.cells[i]].reshape((3,) + ens.ns[i]).contiguous(),
                 ens.particle_view(lam[2], i).contiguous(), ens.particle_view(lam[3], i).contiguous())
        ref = b.step_adjoint(st, lam_i)
        assert rel_err(ens.field_view(got[0], i).cpu().numpy(), ref[0].cpu().numpy()) < tol, i
        assert rel_err(ens.field_view(got[1], i).cpu().numpy(), ref[1].cpu().numpy()) < tol, i
        if b.np:
            assert rel_err(ens.particle_view(got[2], i).cpu().numpy(), ref[2].cpu().numpy()) < tol, i
            assert rel_err(ens.particle_view(got[3], i).cpu().numpy(), ref[3].cpu().numpy()) < tol, i


def test_ensemble_workload_gradient_matches_central_differences():
    prob = EnsembleProblem(n_boxes=4, sides=(8, 12), ppcs=(2, 4), n_modes=3, seed=2, dtype=torch.float64)
    steps, every = 8, 3
    amp = torch.linspace(-1.0, 1.0, 4 * 3, dtype=torch.float64, device="cuda").view(4, 3).requires_grad_(True)
    loss = prob.loss(amp, steps, every)
    loss.backward()
    g = amp.grad.clone()
    assert torch.isfinite(g).all()
    h = 1e-8  # the zigzag current jumps where a perturbed particle changes cell
    for (i, m) in [(0, 0), (1, 2), (3, 1)]:
        a = amp.detach().clone()
        a[i, m] += h
        lp = prob.loss(a, steps, every).item()
        a[i, m] -= 2 * h
        lm = prob.loss(a, steps, every).item()
        fd = (lp - lm) / (2 * h)
        assert g[i, m].item() == pytest.approx(fd, rel=1e-4, abs=1e-12), (i, m)
    # boxes are independent: a box's amplitudes move only its own energy
    with torch.no_grad():
        e0 = prob.box_magnetic_energies(rollout(prob.box, *prob.initial_state(amp.detach()), steps, every)[1])
        a = amp.detach().clone()
        a[2] += 0.5
        e1 = prob.box_magnetic_energies(rollout(prob.box, *prob.initial_state(a), steps, every)[1])
    rel = ((e1 - e0).abs() / e0.abs()).cpu().numpy()
    assert rel[2] > 1e-3 and rel[[0, 1, 3]].max() < 1e-12
