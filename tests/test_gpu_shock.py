"""The reflecting-wall shock as a workload: the reflect inside the fused
particle step against the oracle (particles bit-exact, including the parked
ones it removes), the world-level shock lap (runko_amd.stepper.ShockLap,
the lap of projects/pic-shock/pic.py:250-299 issued as one launch per
method over every tile, fused or not, eager or replayed from HIP graphs)
against the Simulation's shock lap, whose tile-by-tile form the wall tests
compare with the oracle; and the on-GPU shock set-up (cold flow along -x,
density ramp at the wall)."""

import ctypes
import math

import numpy as np
import pytest
import torch

import runko_amd as runko
from runko_amd import _lib as L
from runko_amd import stepper
from runko_amd.init import shock_parameters, shock_world
from gpu_common import build_gpu, build_oracle, gpu_fields, gpu_particles
from helpers import rel_err
from test_gpu_wall import M, Q, assert_particles_equal, shock_lap, wall_inputs
from test_gpu_wall import shock_world as wall_test_world

pytestmark = pytest.mark.gpu


def world_particles(w, s):
    """(states [k, 6] float64, ids) of species s over every tile, sorted by id."""
    sts, ids = [], []
    for t in range(w.ntiles):
        pos, vel, i = w.tile_particles(t, s)
        sts.append(torch.cat([pos, vel]).t().double().cpu().numpy())
        ids.append(i.cpu().numpy())
    sts, ids = np.concatenate(sts), np.concatenate(ids)
    o = np.argsort(ids, kind="stable")
    return sts[o], ids[o]


@pytest.mark.parametrize("beta,pusher", [(0.0, "boris"), (0.25, "boris"), (0.25, "higuera_cary"),
                                         (0.0, "faraday")])
def test_fused_reflect_parity_with_oracle(orc, beta, pusher):
    # particles start on both sides of the wall: in the fused step some
    # reflect, some are parked and removed, the rest are untouched
    nt, n = (2, 2, 1), (8, 8, 8)
    walloc = 5.25
    inputs = wall_inputs(nt, n, seed=21)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs, execution="batched", pusher=pusher)
    w = build_oracle(orc, nt, n, Q, M, inputs, pusher=pusher)
    gam = 1.0 / np.sqrt(1.0 - beta**2)
    wall = runko.pic.threeD.reflector_wall(walloc=walloc, betawall=beta, gammawall=gam)
    for tile in grid.tiles():
        tile.register_reflector_wall(wall)
    cm = runko.tools.comm_mode

    def lap(x):
        x.prtcl_push()
        x.prtcl_reflect_particles()
        x.prtcl_pack_outgoing()
        x.comm_local(cm.pic_particle)
        x.prtcl_deposit_current()
        x.comm_local(cm.emf_J_exchange)
        x.comm_local(cm.emf_J)

    L.check(L.lib().rk_profile_enable(grid.world.h, 1))
    sim.for_one_lap(lap)
    ms = np.zeros(9)
    cnt = np.zeros(9, dtype=np.int64)
    L.check(L.lib().rk_profile_read(grid.world.h, ms.ctypes.data, cnt.ctypes.data))
    L.check(L.lib().rk_profile_enable(grid.world.h, 0))
    assert cnt[0] == 1 and cnt[8] == 1  # one fused launch (and its J fill); no separate reflect
    w.push_particles()
    w.reflect_particles(-1, [(wall.walloc, wall.betawall, wall.gammawall)])
    w.pack_outgoing()
    w.comm_local("pic_particle")
    w.deposit_current()
    w.comm_local("emf_J_exchange")
    w.comm_local("emf_J")
    n_before = sum(len(st) for _, parts in inputs.values() for st in parts)
    assert grid.world.total_particles() < n_before  # some were parked
    assert_particles_equal(grid, w)
    assert rel_err(gpu_fields(grid)[6:9], w.global_fields()[6:9]) < 1e-6


@pytest.mark.parametrize("beta,fused", [(0.0, True), (0.2, True), (0.2, False)])
def test_stepper_shock_lap_matches_simulation(beta, fused):
    nt, n = (3, 2, 1), (8, 8, 8)
    walloc = 4.0
    res = []
    for kind in ("simulation", "stepper"):
        inputs, grid, sim = wall_test_world("batched", nt, n, walloc, beta, seed=31)
        tile = grid.tiles()[0]
        if kind == "simulation":
            for _ in range(5):
                sim.for_one_lap(shock_lap())
            wall_now = tile._walls[0].walloc
        else:
            walls = [L.ReflectorWall(wl.walloc, wl.betawall, wl.gammawall) for wl in tile._walls]
            lap = stepper.ShockLap(grid.world, walls, [bc._struct() for bc in tile._edge_bcs], fused=fused)
            assert lap.fused == fused
            for _ in range(5):
                lap()
                lap.settle()
            torch.cuda.synchronize()
            wall_now = lap.walls[0].walloc
        res.append(([gpu_particles(grid, s) for s in range(2)], gpu_fields(grid), wall_now))
    (ps, fs, ws), (pt, ft, wt) = res
    assert ws == pytest.approx(wt, abs=0.0)
    for (tsim, ssim, isim), (tst, sst, ist) in zip(ps, pt):
        assert np.array_equal(isim, ist) and np.array_equal(tsim, tst)
        assert np.abs(ssim - sst).max() < 1e-4
    assert rel_err(ft, fs) < 1e-5


def small_shock(seed=5):
    return shock_world(nx=128, ny=32, nz=32, tile=16, ppc=2, theta=1e-4, gamma_flow=10.0, sigma=0.1,
                       walloc=15.0, ramp_len=40.0, ramp_factor=4.0, seed=seed)


def test_shock_world_initial_state():
    w, walls, bcs, par = small_shock()
    gam = 10.0
    beta = math.sqrt(1 - 1 / gam**2)
    q0, q1, B0, Ey = shock_parameters(0.45, 10.0, 2, gam, 0.1)
    assert (par["q0"], par["B0"]) == (q0, B0) and Ey == pytest.approx(-beta * B0)
    # column counts: none behind the wall, 4x ppc at the wall, ppc far upstream
    xc = np.arange(128)
    factor = 1.0 + 3.0 * np.clip(1.0 - (xc - 15.0) / 40.0, 0.0, 1.0)
    want = np.where(xc + 1 <= 15, 0, np.rint(2 * factor)).astype(np.int64) * 32 * 32
    for s in range(2):
        st, ids = world_particles(w, s)
        assert len(np.unique(ids)) == len(ids)
        assert st[:, 0].min() >= 15.0
        got = np.bincount(np.floor(st[:, 0]).astype(np.int64), minlength=128)
        assert np.array_equal(got, want)
        # cold flow along -x.  The low-theta sampler (the reference's
        # Box-Muller magnitude, Rayleigh with sigma = sqrt(2 theta), pointed
        # isotropically) has <u^2> = 4 theta, so each rest-frame component
        # has spread sqrt(4 theta / 3); the boost stretches u_x by gamma
        sig = math.sqrt(4e-4 / 3)
        assert np.all(st[:, 3] < 0.0)
        assert st[:, 3].mean() == pytest.approx(-gam * beta, abs=5e-3)
        assert st[:, 3].std() == pytest.approx(gam * sig, rel=0.05)
        assert st[:, 4:6].std(0) == pytest.approx([sig] * 2, rel=0.05)
    # species 1 sits on species 0 (same ids, same positions)
    s0, i0 = world_particles(w, 0)
    s1, i1 = world_particles(w, 1)
    assert np.array_equal(i0, i1) and np.array_equal(s0[:, :3], s1[:, :3])
    # upstream E_y, B_z in every interior cell
    E, B = w.E, w.B
    inner = tuple(slice(3, 3 + v) for v in w.N)
    assert torch.all(E[1][inner] == np.float32(Ey)) and torch.all(B[2][inner] == np.float32(B0))


@pytest.mark.parametrize("form", ["fused", "graph", "graph_queue"])
def test_shock_world_batched_matches_plain_and_keeps_particles(form):
    runs = []
    for kind in (form, "plain"):
        w, walls, bcs, par = small_shock(seed=7)
        n0 = w.total_particles()
        lap = stepper.PlainShockLap(w, walls, bcs) if kind == "plain" else stepper.ShockLap(w, walls, bcs)
        if kind == "graph_queue":
            w.set_fused_launch("queue")
        lap()
        lap.settle()
        if kind.startswith("graph"):
            lap = stepper.GraphLap(w, lap=lap)
        for _ in range(4):
            lap()
            lap.settle()
        torch.cuda.synchronize()
        assert w.total_particles() == n0
        assert all(v == 0 for v in w.error_counters())
        parts = [world_particles(w, s) for s in range(2)]
        for st, _ in parts:
            assert st[:, 0].min() >= 15.0
        runs.append((parts, torch.stack([w.E, w.B]).cpu().numpy()))
        del lap, w
    (pb, fb), (pp, fp) = runs
    for (sb, ib), (sp, ip) in zip(pb, pp):
        assert np.array_equal(ib, ip)
        assert np.abs(sb - sp).max() < 1e-4
    assert rel_err(fb, fp) < 1e-5


@pytest.mark.parametrize("direction,side,position", [(0, 0, 11.5), (0, 1, 11.5), (1, 0, 3.0), (2, 1, 12.0),
                                                     (1, 1, 0.5), (2, 0, 24.0)])
def test_world_edge_bc_matches_per_tile(direction, side, position):
    # one launch per slab of tiles (tile -1) sets exactly the cells that the
    # per-tile calls set, world halo included
    from gpu_common import random_inputs
    nt, n = (3, 2, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 1, seed=41, ppc=1)
    conf, grid, sim = build_gpu(nt, n, Q[:1], M[:1], inputs)
    w = grid.world
    bc = L.EdgeBC(direction, side, position, (1.0, 2.0, 3.0), (4.0, 5.0, 6.0), (7.0, 8.0, 9.0), 0b101, 0b011, 0b111)
    for mode in (0, 1, 2):  # emf_J, emf_E, emf_B
        w.F[2].normal_()
        w.F[3].copy_(w.F[2])
        start = w.F.clone()
        L.check(L.lib().rk_apply_edge_bc(w.h, -1, ctypes.byref(bc), mode, stepper._stream()))
        torch.cuda.synchronize()
        got = w.F.clone()
        w.F.copy_(start)
        for t in range(w.ntiles):
            L.check(L.lib().rk_apply_edge_bc(w.h, t, ctypes.byref(bc), mode, stepper._stream()))
        torch.cuda.synchronize()
        assert torch.equal(got, w.F), mode
        assert not torch.equal(got, start), mode


def test_moving_wall_device_walls_and_graph_match_eager():
    # a wall moving at 0.2 c: the walls held on the device (advanced by a
    # kernel) and the same lap replayed from graphs give the eager lap's
    # particles (within the float-atomic order of the correction currents)
    # and the same float32 wall position
    gam = 1.0 / math.sqrt(1.0 - 0.2 ** 2)
    runs = []
    for kind in ("eager", "device", "graph"):
        w, walls, bcs, par = small_shock(seed=11)
        wall = [L.ReflectorWall(15.0, 0.2, gam)]
        lap = stepper.ShockLap(w, wall, bcs[1:], device_walls=(kind != "eager"))
        lap()
        lap.settle()
        run = lap
        if kind == "graph":
            run = stepper.GraphLap(w, lap=lap)
        for _ in range(4):
            run()
            run.settle()
        torch.cuda.synchronize()
        runs.append(([world_particles(w, s) for s in range(2)], lap.walls[0].walloc, w.total_particles()))
        del run, lap, w
    (pe, we, ne) = runs[0]
    expect = np.float32(15.0)
    for _ in range(5):
        expect = np.float32(expect + np.float32(0.2) * np.float32(0.45))
    assert we == pytest.approx(float(expect), abs=0.0)
    for (pk, wk, nk) in runs[1:]:
        assert wk == we and nk == ne
        for (se, ie), (sk, ik) in zip(pe, pk):
            assert np.array_equal(ie, ik)
            assert np.abs(se - sk).max() < 1e-4
