"""The reflecting-wall shock as a workload: the world-level shock lap
(runko_amd.stepper.ShockLap, the lap of projects/pic-shock/pic.py:250-299
issued as one launch per method over every tile) against the Simulation's
shock lap, whose tile-by-tile form the wall tests compare with the oracle;
and the on-GPU shock set-up (cold flow along -x, density ramp at the wall)."""

import math

import numpy as np
import pytest
import torch

import runko_amd as runko
from runko_amd import _lib as L
from runko_amd import stepper
from runko_amd.init import shock_parameters, shock_world
from gpu_common import gpu_fields, gpu_particles
from helpers import rel_err
from test_gpu_wall import shock_lap
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


@pytest.mark.parametrize("beta", [0.0, 0.2])
def test_stepper_shock_lap_matches_simulation(beta):
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
            lap = stepper.ShockLap(grid.world, walls, [bc._struct() for bc in tile._edge_bcs])
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


def test_shock_world_batched_matches_plain_and_keeps_particles():
    runs = []
    for cls in (stepper.ShockLap, stepper.PlainShockLap):
        w, walls, bcs, par = small_shock(seed=7)
        n0 = w.total_particles()
        lap = cls(w, walls, bcs)
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
