"""Particle pools under overflow: bins that fill move into the pool's spare
arena (rk_move_bins) and, once it is used up, the pool is relaid out in place
(rk_relayout).  Neither replaces a pool buffer, and neither loses, duplicates
or alters a particle: worlds laid out tightly (every bin exactly full, so
the first laps overflow) must end where a world with ample bins ends."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from runko_amd.init import turbulence_world  # noqa: E402
from runko_amd.stepper import BatchedLap, GraphLap  # noqa: E402

SETUP = dict(N=32, tile=16, ppc=8, seed=7, brick=(8, 8, 8))


def _addresses(w):
    return [(p.pos.data_ptr(), p.vel.data_ptr(), p.id.data_ptr(), p.start.data_ptr(), p.cap.data_ptr())
            for p in w.pools]


def _state(w):
    """Per species: ids sorted, with positions and velocities in that order."""
    out = []
    for s in range(w.nsp):
        xs, vs, ids = [], [], []
        for t in range(w.ntiles):
            p, v, i = w.tile_particles(t, s)
            xs.append(p.cpu().numpy())
            vs.append(v.cpu().numpy())
            ids.append(i.cpu().numpy())
        x, v, i = np.concatenate(xs, 1), np.concatenate(vs, 1), np.concatenate(ids)
        o = np.argsort(i)
        out.append((i[o], x[:, o], v[:, o]))
    return out


def _check_bins(w):
    """Every bin lies inside its pool, bins are disjoint and hold at most
    their capacity."""
    for s, p in enumerate(w.pools):
        st = p.start.cpu().numpy()
        cap = p.cap.cpu().numpy().view(np.uint32).astype(np.int64)
        cnt = w.count[s].cpu().numpy().view(np.uint32).astype(np.int64)
        assert np.array_equal(st, p.start_h) and np.array_equal(cap, p.cap_h)
        assert np.all(cnt <= cap)
        assert np.all(st + cap <= p.capacity)
        o = np.argsort(st, kind="stable")
        nz = cap[o] > 0
        s0, e0 = st[o][nz], (st + cap)[o][nz]
        assert np.all(s0[1:] >= e0[:-1])


def _run(w, laps, graph=False):
    lap = GraphLap(w) if graph else BatchedLap(w)
    for i in range(laps):
        lap()
        lap.settle()
        print(f"lap {i}: repacks {w.repacks} moves {w.bin_moves} relayouts {w.relayouts}", flush=True)
    torch.cuda.synchronize()
    return lap


def _compare(wa, wb):
    sa, sb = _state(wa), _state(wb)
    for s in range(wa.nsp):
        ia, xa, va = sa[s]
        ib, xb, vb = sb[s]
        assert np.array_equal(ia, ib), f"species {s}: id sets differ"
        dx, dv = np.max(np.abs(xa - xb)), np.max(np.abs(va - vb))
        print(f"species {s}: max |dx| {dx:.3e}, max |du| {dv:.3e}")
        assert dx < 1e-4 and dv < 1e-5
    fa, fb = wa.F[:2].cpu().numpy(), wb.F[:2].cpu().numpy()
    err = np.max(np.abs(fa - fb)) / max(np.max(np.abs(fa)), 1e-30)
    print(f"E, B rel diff {err:.3e}")
    assert err < 1e-5


@pytest.fixture(scope="module")
def reference():
    """The same world with ample bins (no overflow in these laps)."""
    w, _ = turbulence_world(**SETUP)
    n0 = w.total_particles()
    _run(w, 4)
    assert w.repacks == 0
    assert w.total_particles() == n0
    return w


@pytest.mark.parametrize("arena,stage", [(0, None), (0, "small"), ("large", None), ("small", "small")])
def test_overflow_paths_keep_buffers_and_particles(reference, arena, stage):
    per_bin = 8 ** 3 * SETUP["ppc"]
    nb = (SETUP["N"] // 8) ** 3
    slots = {0: 0, "small": nb * per_bin // 20, "large": 4 * nb * per_bin}[arena]
    w, _ = turbulence_world(**SETUP, layout=dict(slack=1.0, pad=0, arena=slots))
    w.grow_pools = False  # the in-place paths: no pool may be replaced
    if stage == "small":
        # tables (28 B per bin) plus room for about three bins: many batches per pass
        w.stage_bytes = 28 * w.nbricks + 256 + 3 * 32 * (per_bin + 512)
    n0 = w.total_particles()
    addr = _addresses(w)
    _run(w, 4)
    assert _addresses(w) == addr, "a pool buffer was replaced"
    assert w.new_pools == 0
    assert w.total_particles() == n0
    assert w.repacks > 0
    if arena == 0:
        assert w.relayouts > 0 and w.bin_moves == 0
    elif arena == "large":
        assert w.bin_moves > 0 and w.relayouts == 0
    else:
        assert w.bin_moves > 0 and w.relayouts > 0
    _check_bins(w)
    assert list(w.error_counters()) == [0, 0]
    _compare(reference, w)


def test_overflow_under_graph_replay(reference):
    per_bin = 8 ** 3 * SETUP["ppc"]
    nb = (SETUP["N"] // 8) ** 3
    w, _ = turbulence_world(**SETUP, layout=dict(slack=1.0, pad=0, arena=nb * per_bin // 20))
    w.grow_pools = False  # the in-place paths: no pool may be replaced
    n0 = w.total_particles()
    addr = _addresses(w)
    lap = _run(w, 4, graph=True)
    assert _addresses(w) == addr
    assert w.bin_moves > 0 and w.relayouts > 0 and w.new_pools == 0
    # neither a move nor a relayout replaces a buffer: the graphs stay valid
    assert lap.recaptures == 0
    assert w.total_particles() == n0
    _check_bins(w)
    _compare(reference, w)


def test_tight_relayout_grows_the_pool(reference):
    # no arena: the first relayout would leave under 5 % of the pool spare,
    # so with room on the device the species gets a larger pool instead
    w, _ = turbulence_world(**SETUP, layout=dict(slack=1.0, pad=0, arena=0))
    n0 = w.total_particles()
    _run(w, 4)
    print(f"new pools {w.new_pools}, relayouts {w.relayouts}, moves {w.bin_moves}")
    assert w.new_pools > 0
    assert w.total_particles() == n0
    _compare(reference, w)
