"""Mailbox spills of the fused particle step.

Every brick writes its leavers into its own mailbox; past the mailbox's
capacity they spill into the shared record sets 0 and 1, which
k_append_records appends after the mailboxes (and then clears its record
counters itself).  Two copies of a hot 64^3 turbulence world, one with the
default mailboxes and one with 64 records per brick (about a tenth of a
brick's leavers per lap here), must stay identical lap for lap: the same
particles (compared by id) bit for bit, and the same fields bit for bit,
since the current windows accumulate in order-free fixed point."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _particles(w, s):
    pos, ids = [], []
    for t in range(w.ntiles):
        p, v, x = w.tile_particles(t, s)
        pos.append(torch.cat([p, v]).cpu().numpy())
        ids.append(x.cpu().numpy())
    pos, ids = np.concatenate(pos, axis=1), np.concatenate(ids)
    o = np.argsort(ids)
    return pos[:, o], ids[o]


def test_spilled_leavers_match_mailboxed_leavers():
    from runko_amd.init import turbulence_world
    from runko_amd.stepper import BatchedLap
    kw = dict(N=64, tile=16, ppc=8, theta=1.0, sigma=1.0, seed=5, mig_fraction=0.5)
    wa, _ = turbulence_world(mailbox_fraction=0.12, **kw)
    wb, _ = turbulence_world(mailbox_fraction=1e-6, **kw)
    assert wb._mbox_per_brick == 64 and wa._mbox_per_brick > 4 * 64
    la, lb = BatchedLap(wa), BatchedLap(wb)
    br = np.array(wb.brick)
    spilled = 0
    for lap in range(3):
        before, ids0 = _particles(wb, 0)
        la()
        la.settle()
        lb()
        lb.settle()
        for s in range(2):
            pa, ia = _particles(wa, s)
            pb, ib = _particles(wb, s)
            assert np.array_equal(ia, ib), f"lap {lap} species {s}: ids"
            assert np.array_equal(pa, pb), f"lap {lap} species {s}: positions and momenta"
        assert torch.equal(wa.F, wb.F), f"lap {lap}: fields"
        assert list(wa.error_counters()) == [0, 0] and list(wb.error_counters()) == [0, 0]
        # leavers per source brick of species 0 (world coordinates: cell = floor(x))
        after, ids1 = _particles(wb, 0)
        assert np.array_equal(ids0, ids1)
        b0 = np.floor(before[:3]).astype(np.int64) // br[:, None]
        b1 = np.floor(after[:3]).astype(np.int64) // br[:, None]
        moved = np.any(b0 != b1, axis=0)
        key = (b0[0] * 64 + b0[1]) * 64 + b0[2]
        per_brick = np.bincount(key[moved])
        spilled += int(np.maximum(per_brick - 64, 0).sum())
    print(f"species-0 leavers past the 64-record mailboxes over 3 laps: {spilled}")
    assert spilled > 0
