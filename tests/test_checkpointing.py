"""Checkpointed rollouts (runko_amd.diff.rollout) on a toy differentiable
step, on the CPU: flat and nested checkpoint spacings give the gradient of
plain autograd through the unrolled steps, and the backward pass replays the
number of forward steps the spacing implies."""

import pytest
import torch

from runko_amd.diff import checkpoint_levels, checkpoint_states, rollout


def f(s, A):
    E, B, x, u = s
    return (E + 0.1 * torch.sin(B), B + A @ u, x + 0.2 * torch.tanh(u), u - 0.1 * x * E)


class ToyBox:
    """rollout's box interface: a nonlinear map on (E, B, x, u) whose adjoint
    is the transposed Jacobian (from autograd of one step)."""

    def __init__(self):
        g = torch.Generator().manual_seed(3)
        self.A = torch.randn(4, 4, generator=g, dtype=torch.float64) * 0.3
        self.steps = 0

    def empty_state(self):
        return tuple(torch.empty(4, dtype=torch.float64) for _ in range(4))

    def step(self, state, out=None):
        self.steps += 1
        new = f(state, self.A)
        if out is None:
            out = self.empty_state()
        for o, n in zip(out, new):
            o.copy_(n)
        return out

    def step_adjoint(self, state, adj_out, adj_in=None):
        with torch.enable_grad():
            s = tuple(t.detach().clone().requires_grad_(True) for t in state)
            g = torch.autograd.grad(f(s, self.A), s, [a.clone() for a in adj_out])
        target = adj_out if adj_in is None else adj_in
        for t, v in zip(target, g):
            t.copy_(v)
        return target


# (steps, spacing, forward steps replayed by the backward pass)
CASES = [
    (17, 5, 17 - 4),                   # segments 5,5,5,2: each replays len-1
    (30, (10, 3), 3 * (9 + 6)),        # per 10-segment: 3 checkpoints (9 steps), then sub-segments 3,3,3,1 replay 6
    (8, 20, 7),                        # one segment
    (25, (25, 5), 20 + 5 * 4),         # checkpoints at 5,10,15,20 (20 steps), five 5-segments replay 4 each
    (37, (12, 4, 2), None),
]


@pytest.mark.parametrize("steps,every,replayed", CASES)
def test_rollout_gradient_matches_autograd(steps, every, replayed):
    box = ToyBox()
    g = torch.Generator().manual_seed(7)
    x0 = [torch.randn(4, generator=g, dtype=torch.float64).requires_grad_(True) for _ in range(4)]
    w = [torch.randn(4, generator=g, dtype=torch.float64) for _ in range(4)]
    s = tuple(x0)
    for _ in range(steps):
        s = f(s, box.A)
    gref = torch.autograd.grad(sum((r * c).sum() for r, c in zip(s, w)), x0)
    out = rollout(box, *x0, steps, every)
    assert box.steps == steps
    box.steps = 0
    got = torch.autograd.grad(sum((o * c).sum() for o, c in zip(out, w)), x0)
    for a, b in zip(got, gref):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-14)
    if replayed is not None:
        assert box.steps == replayed


def test_checkpoint_states_counts():
    assert checkpoint_states(300, 20) == 15 + 20
    assert checkpoint_states(300, (60, 8)) == 5 + 8 + 8
    assert checkpoint_states(200, 20) == 10 + 20
    assert checkpoint_states(7, 20) == 1 + 7
    with pytest.raises(ValueError):
        checkpoint_levels((4, 8))
    with pytest.raises(ValueError):
        checkpoint_levels(0)
