"""Alternative pushers and the extended-stencil B push on the GPU against
the oracle: one push bit-exact, one stencil half-B push bit-exact, one full
lap in the plain and batched forms at the one-lap bars."""

import numpy as np
import pytest

import runko_amd as runko
from gpu_common import (build_gpu, build_oracle, canonical_lap, gpu_fields, gpu_particles, oracle_particles,
                        random_inputs)
from helpers import rel_err

pytestmark = pytest.mark.gpu

Q = [-0.02, 0.02]
M = [1.0, 1.0]
STENCIL = dict(stencil_delta=-0.06, stencil_gamma=0.004, stencil_beta_p1=0.011, stencil_beta_p2=-0.007,
               stencil_zeta_p1=0.003, stencil_x_delta=-0.05, stencil_y_beta2_p2=0.002, stencil_z_zeta3_p1=0.0004)


@pytest.mark.parametrize("pusher", ["higuera_cary", "faraday"])
def test_push_bitexact(orc, pusher):
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=81, ppc=3, field_amp=0.2)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, pusher=pusher)
    w = build_oracle(orc, nt, n, Q, M, inputs, pusher=pusher)
    for tile in grid.tiles():
        tile.push_particles()
    w.push_particles()
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(sg, so), s


def test_pushers_are_distinct_on_gpu():
    nt, n = (1, 1, 1), (8, 8, 8)
    inputs = random_inputs(nt, n, 1, seed=82, ppc=2, field_amp=0.3)
    out = []
    for pusher in ("boris", "higuera_cary", "faraday"):
        conf, grid, sim = build_gpu(nt, n, Q[:1], M[:1], inputs, pusher=pusher)
        grid.tiles()[0].push_particles()
        out.append(gpu_particles(grid, 0)[1])
    assert not np.array_equal(out[0], out[1]) and not np.array_equal(out[0], out[2])


def test_stencil_half_b_bitexact(orc):
    nt, n = (2, 1, 2), (8, 9, 7)
    inputs = random_inputs(nt, n, 0, seed=83)
    conf, grid, sim = build_gpu(nt, n, [], [], inputs, stencil=STENCIL)
    w = build_oracle(orc, nt, n, [], [], inputs, stencil=STENCIL)
    cm = runko.tools.comm_mode

    def lap(x):
        x.grid_push_half_b()
        x.comm_local(cm.emf_B)
        x.grid_push_e()
        x.comm_local(cm.emf_E)
        x.grid_push_half_b()

    for _ in range(3):
        sim.for_one_lap(lap)
        w.push_half_b(); w.comm_local("emf_B"); w.push_e(); w.comm_local("emf_E"); w.push_half_b()
    assert np.array_equal(gpu_fields(grid), w.global_fields())


@pytest.mark.parametrize("execution", ["plain", "batched"])
def test_stencil_half_b_bitexact_long_regions(orc, execution):
    # tiles of 37 x planes (per tile: a 32-plane march and a partial one) and
    # a 74-plane world (batched: two full marches and a partial one); 10 y
    # rows end in a partial 4-row block
    nt, n = (2, 1, 1), (37, 10, 7)
    inputs = random_inputs(nt, n, 0, seed=86)
    conf, grid, sim = build_gpu(nt, n, [], [], inputs, execution=execution, stencil=STENCIL)
    w = build_oracle(orc, nt, n, [], [], inputs, stencil=STENCIL)
    cm = runko.tools.comm_mode

    def lap(x):
        x.grid_push_half_b()
        x.comm_local(cm.emf_B)
        x.grid_push_e()
        x.comm_local(cm.emf_E)
        x.grid_push_half_b()

    for _ in range(2):
        sim.for_one_lap(lap)
        w.push_half_b(); w.comm_local("emf_B"); w.push_e(); w.comm_local("emf_E"); w.push_half_b()
    assert np.array_equal(gpu_fields(grid), w.global_fields())


@pytest.mark.parametrize("execution", ["plain", "batched"])
def test_unrolled_filter_bitexact_long_regions(orc, execution):
    # the x-column sweep of the unrolled filter over 37-plane tiles / a
    # 74-plane world (columns of 8 outputs, the last one partial)
    nt, n = (2, 1, 1), (37, 10, 7)
    inputs = random_inputs(nt, n, 0, seed=87)
    rng = np.random.default_rng(88)
    for F, _ in inputs.values():
        F[6:9] = rng.standard_normal((3,) + n)
    conf, grid, sim = build_gpu(nt, n, [], [], inputs, execution=execution, filt="binomial2_unrolled")
    w = build_oracle(orc, nt, n, [], [], inputs, filt="binomial2_unrolled")
    cm = runko.tools.comm_mode

    def lap(x):
        x.comm_local(cm.emf_J)
        x.grid_filter_current()
        x.comm_local(cm.emf_J)
        x.grid_filter_current()

    sim.for_one_lap(lap)
    for _ in range(2):
        w.comm_local("emf_J")
        w.filter_current()
    out = gpu_fields(grid)[6:9]
    assert np.array_equal(out, w.global_fields()[6:9])
    assert 0.0 < out.std() < 0.8  # unit-variance noise, smoothed by two passes


def test_zero_stencil_is_fdtd2():
    nt, n = (1, 2, 1), (8, 8, 8)
    inputs = random_inputs(nt, n, 0, seed=84)
    _, ga, sa = build_gpu(nt, n, [], [], inputs, stencil={"stencil_delta": 0.0})
    _, gb, sb = build_gpu(nt, n, [], [], inputs)
    for g in (ga, gb):
        for tile in g.tiles():
            tile.push_half_b()
    assert np.array_equal(gpu_fields(ga), gpu_fields(gb))


@pytest.mark.parametrize("execution", ["plain", "batched"])
@pytest.mark.parametrize("pusher", ["higuera_cary", "faraday"])
def test_one_lap_parity(orc, execution, pusher):
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=85, ppc=4, umax=1.5)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution=execution, pusher=pusher, stencil=STENCIL)
    w = build_oracle(orc, nt, n, Q, M, inputs, pusher=pusher, stencil=STENCIL)
    sim.for_one_lap(canonical_lap(sim))
    w.lap(0)
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(tg, to)
        assert np.array_equal(sg, so)
    g, o = gpu_fields(grid), w.global_fields()
    assert np.array_equal(g[3:6], o[3:6])
    assert rel_err(g[0:3], o[0:3]) < 1e-6
    assert rel_err(g[6:9], o[6:9]) < 1e-6
