"""Host-side planning of the particle pools: bin capacities, spare arenas
sized from free device memory, and the capacities of an in-place relayout
(no GPU needed)."""

import numpy as np
import pytest

from runko_amd.world import PAD, SLACK, SLOT_BYTES, bin_caps, plan_arena, plan_relayout


def test_bin_caps_slack_pad_and_rounding():
    c = bin_caps([0, 1, 100, 8192])
    assert c[0] == 0
    assert all(v % 32 == 0 for v in c)
    assert c[1] == 96  # ceil(1.25) + 64 = 66 -> 96
    assert c[3] == int(np.ceil(8192 * SLACK)) + PAD  # 10304 is already a multiple of 32
    assert list(bin_caps([4096], 1.0, 0)) == [4096]


def test_plan_arena_targets_half_the_bins_when_memory_allows():
    used = [1000, 3000]
    a = plan_arena(used, free_bytes=10**9, reserve_bytes=0)
    assert list(a) == [500, 1500]


def test_plan_arena_shares_what_fits_in_proportion():
    used = [1000, 3000]
    free = (4000 + 400) * SLOT_BYTES + 31
    a = plan_arena(used, free_bytes=free, reserve_bytes=0)
    assert list(a) == [100, 300]
    assert (sum(used) + a.sum()) * SLOT_BYTES <= free


def test_plan_arena_respects_the_reserve_and_raises_when_bins_do_not_fit():
    used = [1000]
    a = plan_arena(used, free_bytes=1200 * SLOT_BYTES, reserve_bytes=100 * SLOT_BYTES)
    assert list(a) == [100]
    with pytest.raises(RuntimeError, match="particle bins need"):
        plan_arena(used, free_bytes=1000 * SLOT_BYTES, reserve_bytes=1)


def test_plan_relayout_gives_every_bin_its_demand_plus_slack():
    # with room to spare: demand * 1.5 + 64 per bin (multiples of 32), the
    # full bin at least its grown capacity; the bin that emptied shrinks
    caps = np.array([64, 64, 64])
    demand = np.array([10, 80, 64])
    grown = np.array([0, 192, 0])
    c = plan_relayout(demand, caps, grown, capacity=1000)
    assert list(c) == [96, 192, 160]
    assert list(c) == list(np.maximum(bin_caps(demand, 1.5), np.where(demand > caps, grown, 0)))
    caps = np.array([512, 64])
    demand = np.array([8, 100])
    c = plan_relayout(demand, caps, np.array([0, 256]), capacity=1000)
    assert list(c) == [96, 256]
    # the looser slack only while it leaves 15 % of the pool spare: 1.5 gives
    # 96 + 384 = 480 > 0.85 * 420, the set-up's 1.25 gives 96 + 320 = 416,
    # which fits the pool
    c = plan_relayout(np.array([8, 200]), np.array([512, 64]), np.array([0, 256]), capacity=420)
    assert list(c) == [96, 320]
    assert list(c) == list(np.maximum(bin_caps([8, 200]), [0, 256]))
    demand = np.array([1000, 1000])
    assert list(plan_relayout(demand, np.array([1600, 1600]), np.array([0, 0]), capacity=10**5)) == \
        list(bin_caps(demand, 1.5))
    assert list(plan_relayout(demand, np.array([1600, 1600]), np.array([0, 0]), capacity=3000)) == \
        list(bin_caps(demand))


def test_plan_relayout_keeps_capacities_when_the_fresh_layout_does_not_fit():
    # demand + slack for all: 96 + 192 + 160 = 448 > 330; full bin grown and
    # the others kept: 64 + 192 + 64 = 320 fits
    caps = np.array([64, 64, 64])
    demand = np.array([10, 80, 64])
    grown = np.array([0, 192, 0])
    assert list(plan_relayout(demand, caps, grown, capacity=330)) == [64, 192, 64]


def test_plan_relayout_falls_back_to_tighter_capacities():
    caps = np.array([64, 64, 64])
    demand = np.array([10, 80, 64])
    grown = np.array([0, 192, 0])
    # grown (320) does not fit 250; demand * slack + pad = 96 + 160 + 160 = 416 neither;
    # demand + pad rounded = 96 + 160 + 128 = 384 neither; the exact demand (154) does
    c = plan_relayout(demand, caps, grown, capacity=250)
    assert list(c) == [10, 80, 64]
    # never below what a bin stores (min(demand, cap)), and None when the demand exceeds the pool
    assert plan_relayout(demand, caps, grown, capacity=153) is None


def test_plan_relayout_middle_fallback():
    caps = np.array([32, 32])
    demand = np.array([40, 8])
    grown = np.array([128, 0])
    c = plan_relayout(demand, caps, grown, capacity=200)
    # grown + kept = 160 fits
    assert list(c) == [128, 32]
    c = plan_relayout(demand, caps, grown, capacity=159)
    # demand*1.25+64 -> 114 -> 128, 74 -> 96: 224 > 159; demand + 64 -> 104 -> 128, 72 -> 96: 224; exact 48
    assert list(c) == [40, 8]
