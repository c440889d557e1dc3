"""rk_energies_device, the allocation-free form of rk_energies: called through
the C ABI, its values must equal rk_energies' bit for bit, on two calls in a
row (the second on a smaller region, so stale partial sums of the first call
sit in the static array), for the whole world and for single tiles."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from runko_amd import _lib as L  # noqa: E402
from runko_amd.init import turbulence_world  # noqa: E402
from runko_amd.stepper import BatchedLap  # noqa: E402
from runko_amd.world import _stream  # noqa: E402


def _host(w, tile):
    out = np.zeros(2 + w.nsp, dtype=np.float64)
    L.check(L.lib().rk_energies(w.h, tile, out.ctypes.data, _stream()))
    return out


def _device(w, tile):
    out = torch.full((2 + w.nsp,), float("nan"), dtype=torch.float64, device="cuda")
    L.check(L.lib().rk_energies_device(w.h, tile, ctypes.c_void_p(out.data_ptr()), _stream()))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.fixture(scope="module")
def worlds():
    big, _ = turbulence_world(N=64, tile=16, ppc=4, seed=3)
    small, _ = turbulence_world(N=32, tile=16, ppc=2, seed=5)
    for w in (big, small):
        lap = BatchedLap(w)
        lap()
        lap.settle()
    torch.cuda.synchronize()
    return big, small


def test_two_calls_match_rk_energies(worlds):
    big, small = worlds
    for w, tile in ((big, -1), (small, -1), (big, 5), (small, 0)):
        h, d = _host(w, tile), _device(w, tile)
        print(f"tile {tile}: {d}")
        assert np.all(np.isfinite(d)) and np.all(d[:2] > 0)
        assert np.array_equal(h.view(np.uint64), d.view(np.uint64)), (tile, h, d)


def test_bad_arguments_and_region_limit_are_errors():
    from runko_amd.world import World

    w, _ = turbulence_world(N=32, tile=16, ppc=1, seed=1)
    out = torch.zeros(2 + w.nsp, dtype=torch.float64, device="cuda")
    with pytest.raises(L.RkError):
        L.check(L.lib().rk_energies_device(w.h, 99, ctypes.c_void_p(out.data_ptr()), _stream()))
    with pytest.raises(L.RkError):
        L.check(L.lib().rk_energies_device(w.h, -1, None, _stream()))
    # a world of 8196 x planes: beyond the static array's 8192
    long = World((8196, 4, 4), (1, 1, 1), (0, 0, 0), (1, 1, 1), True, 0.45, [-1e-3, 1e-3], [1.0, 1.0])
    with pytest.raises(L.RkError, match="8192"):
        L.check(L.lib().rk_energies_device(long.h, -1, ctypes.c_void_p(out.data_ptr()), _stream()))
    torch.cuda.synchronize()
