"""Three binomial2 passes in one launch (rk_filter_current_passes) against the
passes run one by one in the driver's order (filter, comm J, filter, filter;
projects/pic-turbulence/pic.py:208-212) and against the passes without the
ghost sync between them: the world interior must agree bit for bit, on
worlds whose sizes are and are not multiples of the launch's column shape,
with x segments of 4 planes (small worlds) up to the full 128."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from runko_amd.tools import comm_mode  # noqa: E402
from runko_amd.world import HALO, World  # noqa: E402


def _world(n, nt, J):
    w = World(n, nt, (0, 0, 0), nt, True, 0.45, [-1e-3, 1e-3], [1.0, 1.0])
    w.J.copy_(torch.from_numpy(J).to("cuda"))
    w.comm_local(int(comm_mode.emf_J))
    return w


def _interior(w):
    N = w.N
    return w.J[:, HALO:HALO + N[0], HALO:HALO + N[1], HALO:HALO + N[2]].cpu().numpy()


@pytest.mark.parametrize("n,nt", [((16, 16, 16), (2, 2, 2)), ((8, 40, 12), (3, 1, 6)), ((130, 9, 33), (1, 1, 1)),
                                  ((32, 64, 64), (4, 8, 4))])
def test_fused_passes_match_pass_by_pass(n, nt):
    rng = np.random.default_rng(sum(n) + sum(nt))
    L = tuple(a * b + 2 * HALO for a, b in zip(n, nt))
    J = rng.standard_normal((3,) + L).astype(np.float32)
    J[:, :, :, ::7] *= 1e-30  # tiny values: products near the subnormal range
    ref = _world(n, nt, J)
    ref.call("rk_filter_current", -1)
    ref.comm_local(int(comm_mode.emf_J))
    ref.call("rk_filter_current", -1)
    ref.call("rk_filter_current", -1)
    nocomm = _world(n, nt, J)
    for _ in range(3):
        nocomm.call("rk_filter_current", -1)
    fused = _world(n, nt, J)
    fused.filter_current_passes(3)
    torch.cuda.synchronize()
    a, b, c = _interior(ref), _interior(nocomm), _interior(fused)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    assert np.array_equal(b.view(np.uint32), c.view(np.uint32))
    # the ghost layer afterwards holds the periodic images of the result
    F = fused.J.cpu().numpy()
    N = fused.N
    assert np.array_equal(F[:, :HALO, HALO:HALO + N[1], HALO:HALO + N[2]],
                          F[:, N[0]:N[0] + HALO, HALO:HALO + N[1], HALO:HALO + N[2]])
    # other pass counts run pass by pass with the ghost sync between passes
    for p in (1, 2):
        one = _world(n, nt, J)
        one.filter_current_passes(p)
        two = _world(n, nt, J)
        for q in range(p):
            if q:
                two.comm_local(int(comm_mode.emf_J))
            two.call("rk_filter_current", -1)
        assert np.array_equal(_interior(one), _interior(two))
