"""GPU parity: the HIP path (through the runko-compatible API and the C ABI)
against the oracle on identical seeded inputs.

Bars: particle positions, momenta, ids, per-tile counts, migration lists and
cell indices bit-exact after one step; fields bit-exact where no atomic sum
is involved (fdtd2, E -= J, filters, interpolation + Boris); J and everything
downstream of it within 1e-6 normwise relative error (float32 sums in
another order), and checked against the FP64 oracle."""

import numpy as np
import pytest
import torch

import runko_amd as runko
from gpu_common import (build_gpu, build_oracle, canonical_lap, gpu_fields, gpu_particles, make_config,
                        oracle_particles, random_inputs)
from helpers import rel_err

pytestmark = pytest.mark.gpu

Q = [-0.02, 0.02]
M = [1.0, 1.0]


def _standalone_pair(orc, n, filt="binomial2", seed=0):
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((9,) + n) * 0.1
    conf = make_config((3, 3, 3), n, [], [], filt=filt)
    tile = runko.emf.threeD.Tile((1, 2, 0), conf)
    tile.batch_set_EBJ(*[lambda x, y, z, a=F[c]: a for c in range(9)])
    w = orc.OracleWorld((3, 3, 3), n, 0.45, [], [], prec=32, current_filter=filt)
    t = w.tile_index(1, 2, 0)
    w.set_fields(t, F)
    return tile, w, t


@pytest.mark.parametrize("filt", ["binomial2", "binomial2_unrolled"])
def test_standalone_tile_fields_bitexact(orc, filt):
    n = (10, 11, 13)
    tile, w, t = _standalone_pair(orc, n, filt)
    for step in ("push_half_b", "push_e", "add_current", "filter_current", "filter_current", "push_half_b"):
        getattr(tile, step)()
        getattr(w, {"filter_current": "filter_current"}.get(step, step))(t)
        E, B, J = tile.get_EBJ_with_halo()
        got = np.stack(list(E) + list(B) + list(J))
        ref = w.get_fields_halo(t)
        inner = (slice(None),) + (slice(1, -1),) * 3  # outer shell: filter scratch, not compared
        assert np.array_equal(got[inner], ref[inner]), step


def test_grid_fields_periodic_bitexact(orc):
    nt, n = (2, 3, 2), (6, 5, 7)
    inputs = random_inputs(nt, n, 0, seed=3)
    for F, _ in inputs.values():
        F[6:9] = np.random.default_rng(4).standard_normal((3,) + n) * 0.01
    conf, grid, sim = build_gpu(nt, n, [], [], inputs)
    w = build_oracle(orc, nt, n, [], [], inputs)
    cm = runko.tools.comm_mode

    def lap(x):
        x.grid_push_half_b()
        x.comm_local(cm.emf_B)
        x.comm_local(cm.emf_J)
        x.grid_filter_current()
        x.comm_local(cm.emf_J)
        x.grid_filter_current()
        x.grid_filter_current()
        x.grid_push_half_b()
        x.comm_local(cm.emf_B)
        x.grid_push_e()
        x.grid_add_current()
        x.comm_local(cm.emf_E)

    for _ in range(3):
        sim.for_one_lap(lap)
        w.push_half_b(); w.comm_local("emf_B"); w.comm_local("emf_J"); w.filter_current()
        w.comm_local("emf_J"); w.filter_current(); w.filter_current(); w.push_half_b()
        w.comm_local("emf_B"); w.push_e(); w.add_current(); w.comm_local("emf_E")
    assert np.array_equal(gpu_fields(grid), w.global_fields())


def test_push_bitexact(orc):
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=5, ppc=3)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    for tile in grid.tiles():
        tile.push_particles()
    w.push_particles()
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(tg, to)
        assert np.array_equal(sg, so)


def test_migration_bitexact(orc):
    nt, n = (2, 2, 3), (6, 6, 4)
    inputs = random_inputs(nt, n, 2, seed=6, ppc=4, umax=4.0)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    tiles = grid.tiles()
    for tile in tiles:
        tile.push_particles()
    for tile in tiles:
        tile.pack_outgoing_particles()
    w.push_particles()
    w.pack_outgoing()
    # outgoing lists: (source tile, direction, species, id) sets
    pos, vel, ids, src, dirs, sp = grid.world.outgoing()
    got = set(zip(src.tolist(), dirs.tolist(), sp.tolist(), ids.tolist()))
    ref = set()
    for t in range(w.ntiles):
        _, oi, od, osp = w.get_outgoing(t)
        ref |= set(zip([t] * len(oi), od.tolist(), osp.tolist(), oi.tolist()))
    assert len(ref) > 100
    assert got == ref
    runko_world = grid.world
    runko_world.comm_local(int(runko.tools.comm_mode.pic_particle))
    w.comm_local("pic_particle")
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(tg, to)  # per-tile membership
        assert np.array_equal(sg, so)  # wrapped positions bit-exact
        for t in range(w.ntiles):
            assert tiles[t].number_of_particles(s) == w.num_alive(t, s)


def test_deposit_matches_oracle(orc):
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=7, ppc=4, umax=2.0)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w32 = build_oracle(orc, nt, n, Q, M, inputs)
    w64 = build_oracle(orc, nt, n, Q, M, inputs, prec=64)
    cm = runko.tools.comm_mode

    def lap(x):
        x.prtcl_deposit_current()
        x.comm_local(cm.emf_J_exchange)
        x.comm_local(cm.emf_J)

    sim.for_one_lap(lap)
    for w in (w32, w64):
        w.deposit_current()
        w.comm_local("emf_J_exchange")
        w.comm_local("emf_J")
    g = gpu_fields(grid)[6:9]
    assert rel_err(g, w32.global_fields()[6:9]) < 1e-6
    assert rel_err(g, w64.global_fields()[6:9]) < 1e-5


@pytest.mark.parametrize("execution", ["plain", "batched"])
def test_one_lap_parity(orc, execution):
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=8, ppc=4, umax=1.5)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution=execution)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    w64 = build_oracle(orc, nt, n, Q, M, inputs, prec=64)
    sim.for_one_lap(canonical_lap(sim))
    w.lap(0)
    w64.lap(0)
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(tg, to)
        assert np.array_equal(sg, so)
        cells_g = np.floor(sg[:, :3]).astype(np.int64)
        cells_o = np.floor(so[:, :3]).astype(np.int64)
        assert np.array_equal(cells_g, cells_o)
    g, o, o64 = gpu_fields(grid), w.global_fields(), w64.global_fields()
    assert np.array_equal(g[3:6], o[3:6])  # B does not see this lap's current
    assert rel_err(g[0:3], o[0:3]) < 1e-6
    assert rel_err(g[6:9], o[6:9]) < 1e-6
    assert rel_err(g[0:3], o64[0:3]) < 1e-5


@pytest.mark.parametrize("execution", ["plain", "batched"])
def test_multi_lap_energy_and_conservation(orc, execution):
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=9, ppc=4, umax=1.0)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution=execution)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    ids0 = [np.sort(gpu_particles(grid, s)[2]) for s in range(2)]
    step = canonical_lap(sim)
    for lap in range(12):
        sim.for_one_lap(step)
        w.lap(lap)
        eg = grid.world.energies()
        eo = w.energies()
        assert rel_err(eg, eo) < 1e-5
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert np.array_equal(np.sort(ig), ids0[s])
        assert np.max(np.abs(sg - so)) < 1e-3
    assert np.all(grid.world.error_counters() == 0)


def test_sort_keeps_particles(orc):
    nt, n = (2, 1, 1), (16, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=10, ppc=2)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    before = [gpu_particles(grid, s) for s in range(2)]
    for tile in grid.tiles():
        tile.sort_particles()
    for s in range(2):
        after = gpu_particles(grid, s)
        for a, b in zip(before[s], after):
            assert np.array_equal(a, b)


def test_sort_key_order_matches_oracle(orc):
    # pic/tile.c++:417-438 -> particle.h:574-703: after sort_particles the
    # storage order follows the haloed layout_right cell key.  The GPU sorts
    # inside each brick bin; restricted to a bin, the oracle's sorted order
    # must visit the same cells in the same sequence (ties inside a cell are
    # unordered in both).  The injection order is a random permutation, so
    # a sort that did nothing fails.
    nt, n = (2, 1, 2), (12, 8, 10)
    inputs = random_inputs(nt, n, 2, seed=21, ppc=4)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    for tile in grid.tiles():
        tile.sort_particles()
    w.sort_particles(-1)
    world = grid.world
    checked = 0
    for s in range(2):
        counts = world.count[s].cpu().numpy()
        for t in range(world.ntiles):
            ost, oid = w.get_particles(t, s)
            rank = {int(i): r for r, i in enumerate(oid)}
            pos, vel, ids = world.tile_particles(t, s)
            pos = pos.cpu().numpy()
            ids = ids.cpu().numpy().view(np.uint64)
            cells = np.floor(pos).astype(np.int64).T
            assert sorted(ids.tolist()) == sorted(int(i) for i in oid)
            o = 0
            for b in range(world.bpt):
                c = int(counts[t * world.bpt + b])
                g_cells = [tuple(v) for v in cells[o:o + c]]
                r = np.array([rank[int(i)] for i in ids[o:o + c]])
                o_cells = [g_cells[k] for k in np.argsort(r, kind="stable")]
                assert g_cells == o_cells, f"species {s} tile {t} bin {b}"
                assert len(set(g_cells)) > 1
                checked += c
                o += c
    assert checked == sum(len(p) for _, ps in inputs.values() for p in ps)


def test_bins_grow_on_demand():
    # every particle of a tile injected into one brick: the bins repack and nothing is lost
    conf = make_config((1, 1, 1), (16, 16, 16), Q, M)
    tile = runko.pic.threeD.Tile((0, 0, 0), conf)
    rng = np.random.default_rng(0)
    total = 0
    for k in range(5):
        x = 1.0 + rng.random(3000) * 2
        st = runko.pic.threeD.ParticleStateBatch(pos=(x, x, x), vel=(x * 0, x * 0, x * 0))
        tile.batch_inject_to_cells(0, lambda a, b, c, st=st: st)
        total += 3000
        assert tile.number_of_particles(0) == total
    ids = tile.get_ids(0)
    assert len(np.unique(ids)) == total


def test_energy_diagnostics(orc):
    nt, n = (2, 1, 2), (8, 6, 8)
    inputs = random_inputs(nt, n, 2, seed=11, ppc=3)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w64 = build_oracle(orc, nt, n, Q, M, inputs, prec=64)
    e_g = grid.world.energies()
    e_o = w64.energies()
    assert rel_err(e_g, e_o) < 1e-6


def test_comm_type_error():
    nt, n = (1, 1, 1), (8, 8, 8)
    inputs = random_inputs(nt, n, 1, seed=1, ppc=1)
    conf, grid, sim = build_gpu(nt, n, Q[:1], M[:1], inputs)
    with pytest.raises(TypeError):
        sim.for_one_lap(lambda x: x.comm_local(3))
    with pytest.raises(RuntimeError):
        sim.for_one_lap(lambda x: x.not_a_method())


def _gauss_numpy(grid, passes, charges):
    """div E - F^passes(rho_CIC) in FP64 from the readback (same stencil and
    summation order as rk_gauss_residual); returns (max, rms rho, rms div E)."""
    world = grid.world
    N, n = world.N, world.n
    rho = np.zeros(N)
    for s, q in enumerate(charges):
        for t, tile in enumerate(grid.tiles()):
            pos = np.stack(tile.get_positions(s), 1).astype(np.float32)
            if len(pos) == 0:
                continue
            origo = np.array([(c * m) for c, m in zip(tile.index, n)], dtype=np.float32) - np.float32(3.0)
            p = pos - origo  # float32, as the kernel computes it
            f = np.floor(p)
            wgt = p.astype(np.float64) - f.astype(np.float64)
            base = np.array(tile.index) * np.array(n) + f.astype(np.int64) - 3
            for a in (0, 1):
                for b in (0, 1):
                    for c in (0, 1):
                        wa = (wgt[:, 0] if a else 1.0 - wgt[:, 0]) * (wgt[:, 1] if b else 1.0 - wgt[:, 1]) * \
                             (wgt[:, 2] if c else 1.0 - wgt[:, 2])
                        idx = tuple((base[:, d] + (a, b, c)[d]) % N[d] for d in range(3))
                        np.add.at(rho, idx, q * wa)
    w1 = (1.0, 2.0, 1.0)
    for _ in range(passes):
        acc = np.zeros_like(rho)
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    acc += w1[a] * w1[b] * w1[c] / 64.0 * np.roll(rho, (1 - a, 1 - b, 1 - c), (0, 1, 2))
        rho = acc
    E = world.F[0].double().cpu().numpy()
    inner = tuple(slice(3, 3 + v) for v in N)
    div = ((E[0][inner] - E[0][tuple(slice(2, 2 + v) if d == 0 else slice(3, 3 + v) for d, v in enumerate(N))]) +
           (E[1][inner] - E[1][tuple(slice(2, 2 + v) if d == 1 else slice(3, 3 + v) for d, v in enumerate(N))]) +
           (E[2][inner] - E[2][tuple(slice(2, 2 + v) if d == 2 else slice(3, 3 + v) for d, v in enumerate(N))]))
    return np.max(np.abs(div - rho)), np.sqrt(np.mean(rho**2)), np.sqrt(np.mean(div**2))


@pytest.mark.parametrize("execution", ["plain", "batched"])
def test_gauss_residual(execution):
    # neutral start (species 1 on top of species 0, E = 0): div E - F^3 rho
    # starts at 0; after laps it stays at the size of the 3D zigzag defect.
    # The device diagnostic must equal the FP64 numpy restatement.
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=12, ppc=4, umax=1.0)
    for F, parts in inputs.values():
        F[0:3] = 0.0
        parts[1][:, :3] = parts[0][:, :3]
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution=execution)
    g0 = grid.world.gauss_residual(3)
    assert g0[0] < 1e-12 and g0[1] < 1e-12
    step = canonical_lap(sim)
    for _ in range(6):
        sim.for_one_lap(step)
    g = grid.world.gauss_residual(3)
    h = _gauss_numpy(grid, 3, Q)
    print(f"gauss residual {execution}: device {g}, numpy {h}")
    assert g[1] > 0 and g[2] > 0
    assert abs(g[0] - h[0]) <= 1e-9 * h[1]
    assert g[1] == pytest.approx(h[1], rel=1e-9)
    assert g[2] == pytest.approx(h[2], rel=1e-9)
    assert g[0] < 0.05 * g[1]


@pytest.mark.parametrize("execution", ["plain", "batched"])
def test_gauss_residual_matches_fp64_oracle(orc, execution):
    # The same neutral start and laps on the GPU (FP32) and in the FP64 oracle:
    # each world's residual div E - F^3 rho on its own state.  The residual is
    # the 3D zigzag defect, a property of the scheme (not of rounding), so the
    # FP32 GPU run must reproduce the FP64 oracle's residual; its size
    # relative to rms rho is printed (the reference's scheme keeps no Gauss
    # law, so no bound like 1e-5 of rms rho can hold).
    nt, n = (2, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=12, ppc=4, umax=1.0)
    for F, parts in inputs.values():
        F[0:3] = 0.0
        parts[1][:, :3] = parts[0][:, :3]
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution=execution)
    w64 = build_oracle(orc, nt, n, Q, M, inputs, prec=64)
    w32 = build_oracle(orc, nt, n, Q, M, inputs)
    assert w64.gauss_residual(3)[0] < 1e-15
    step = canonical_lap(sim)
    for lap in range(6):
        sim.for_one_lap(step)
        w64.lap(lap)
        w32.lap(lap)
    g = grid.world.gauss_residual(3)
    o64, o32 = w64.gauss_residual(3), w32.gauss_residual(3)
    print(f"gauss residual {execution}: GPU {g}, FP64 oracle {o64}, FP32 oracle {o32}; "
          f"max residual / rms rho: GPU {g[0] / g[1]:.3e}, FP64 oracle {o64[0] / o64[1]:.3e}")
    # rms rho and rms div E agree within the FP32 bar; the max residual sits
    # at one node where the defect peaks, and there the J rounding differs
    # (the GPU rounds each contribution to 2^-30 of max|q| and sums exactly,
    # the oracle sums in FP32 or FP64): measured 0.56 % from the FP64 oracle
    # and 0.66 % from the FP32 one, while the two oracles differ by 0.1 %
    assert g[0] == pytest.approx(o64[0], rel=2e-2)
    assert g[0] == pytest.approx(o32[0], rel=2e-2)
    assert g[1] == pytest.approx(o64[1], rel=1e-6)
    assert g[2] == pytest.approx(o64[2], rel=1e-6)


def test_graph_lap_matches_eager():
    # two identical worlds: one advanced eagerly, one by graph replays with
    # eager work (energies, syncs) between the replays; every lap compared
    from runko_amd.init import turbulence_world
    from runko_amd.stepper import BatchedLap, GraphLap
    wa, _ = turbulence_world(N=64, tile=16, ppc=4, seed=3)
    wb, _ = turbulence_world(N=64, tile=16, ppc=4, seed=3)
    ea = BatchedLap(wa)
    gb = GraphLap(wb)
    for lap in range(5):
        ea()
        ea.settle()
        gb()
        gb.settle()
        e1, e2 = wa.energies(), wb.energies()
        assert rel_err(e2, e1) < 1e-6, f"lap {lap}"
    assert wa.total_particles() == wb.total_particles()
    assert rel_err(wb.F.cpu().numpy(), wa.F.cpu().numpy()) < 1e-5
    for s in range(2):
        ia, ib = [], []
        for t in range(wa.ntiles):
            pa, va, xa = wa.tile_particles(t, s)
            pb, vb, xb = wb.tile_particles(t, s)
            oa, ob = torch.argsort(xa), torch.argsort(xb)
            assert torch.equal(xa[oa], xb[ob])
            assert torch.max(torch.abs(pa[:, oa] - pb[:, ob])).item() < 1e-4


@pytest.mark.parametrize("n", [(10, 6, 9), (4, 4, 4), (6, 11, 5)])
def test_one_lap_parity_uneven_bricks(orc, n):
    # non-cubic and tiny bricks (generic-shape fused kernel, windows reached
    # by bricks two away when br <= 2) against the oracle, batched form
    nt = (2, 2, 2)
    inputs = random_inputs(nt, n, 2, seed=13, ppc=3, umax=1.2)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution="batched")
    w = build_oracle(orc, nt, n, Q, M, inputs)
    sim.for_one_lap(canonical_lap(sim))
    w.lap(0)
    for s in range(2):
        tg, sg, ig = gpu_particles(grid, s)
        to, so, io = oracle_particles(w, s)
        assert np.array_equal(ig, io)
        assert np.array_equal(tg, to)
        assert np.array_equal(sg, so)
    g, o = gpu_fields(grid), w.global_fields()
    assert rel_err(g[0:3], o[0:3]) < 1e-6
    assert rel_err(g[6:9], o[6:9]) < 1e-6
