"""Host logic of the brick-shape autotuner (no GPU)."""

import numpy as np

from runko_amd.autotune import LDS_LIMIT, candidate_bricks, fused_lds_bytes


def test_lds_footprint_formula():
    # node means (float, 6 per node) over (b+2)^3 plus the current window
    # (double, 3 per node) over (b+3)^3
    assert fused_lds_bytes((8, 8, 8)) == 4 * 6 * 10**3 + 8 * 3 * 11**3
    assert fused_lds_bytes((4, 8, 16)) == 4 * 6 * 6 * 10 * 18 + 8 * 3 * 7 * 11 * 19


def test_candidates_divide_the_tile_and_fit_lds():
    for n in [(32, 32, 32), (16, 16, 16), (8, 8, 8), (32, 16, 8), (12, 12, 12)]:
        c = candidate_bricks(n)
        assert len(set(c)) == len(c)
        for b in c:
            assert all(v in (4, 8, 16) and nn % v == 0 for v, nn in zip(b, n))
            assert fused_lds_bytes(b) <= LDS_LIMIT
        # the cubes lead, the larger first
        cubes = [b for b in c if len(set(b)) == 1]
        assert c[:len(cubes)] == sorted(cubes, reverse=True)
    assert candidate_bricks((32, 32, 32))[0] == (8, 8, 8)
    assert (16, 16, 16) not in candidate_bricks((32, 32, 32))  # 173 KiB of LDS
    assert candidate_bricks((12, 12, 12)) == [(4, 4, 4)]
    assert candidate_bricks((6, 6, 6)) == []
    assert np.all([fused_lds_bytes(b) > LDS_LIMIT for b in [(16, 16, 8), (16, 8, 8)]])
