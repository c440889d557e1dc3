"""Snapshot writers on the GPU: packed arrays bit-exact against the oracle's
restatement (oracle/snapshots.py) on random fields and particles, E/B at
positions bit-exact against the oracle's interpolation, and the reference's
writer behaviours (tests/py/test_mpiio_fields.py, test_mpiio_particles.py,
test_mpiio_spectra.py) through the files and readers."""

import os

import numpy as np
import pytest

import runko_amd as runko
import snapshots as osnap
from gpu_common import build_gpu, build_oracle, gpu_fields, random_inputs
from runko_amd import snapshots as S

pytestmark = pytest.mark.gpu

F32 = np.float32
Q = [-0.02, 0.02]
M = [1.0, 1.0]


def tile_parts(grid, nsp):
    pos, vel = [], []
    for tile in grid.tiles():
        pos.append([np.stack(tile.get_positions(s)).astype(F32) for s in range(nsp)])
        vel.append([np.stack(tile.get_velocities(s)).astype(F32) for s in range(nsp)])
    return pos, vel


def random_grid(nt, n, seed, nsp=2, umax=3.0):
    inputs = random_inputs(nt, n, nsp, seed=seed, ppc=3, umax=umax)
    rng = np.random.default_rng(seed + 100)
    for F, _ in inputs.values():
        F[6:9] = rng.standard_normal((3,) + n)
    return inputs, build_gpu(nt, n, Q[:nsp], M[:nsp], inputs)


@pytest.mark.parametrize("stride", [1, 2, 3, 8])
def test_fields_pack_matches_oracle(stride):
    nt, n = (2, 2, 3), (8, 8, 8)
    inputs, (conf, grid, sim) = random_grid(nt, n, seed=61)
    for nspecies in (2, 1, 5):
        wr = S.MpiioFieldsWriter("unused", *[v for a in zip(nt, n) for v in a], stride, nspecies)
        got = wr.pack(grid).cpu().numpy()
        F = gpu_fields(grid).astype(F32)
        pos, _ = tile_parts(grid, 2)
        ref = osnap.fields_pack([F[0], F[1], F[2]], [F[3], F[4], F[5]], [F[6], F[7], F[8]], pos, nt, n, stride,
                                min(nspecies, 5))
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), nspecies


@pytest.mark.parametrize("stride", [1, 2, 4])
def test_spectra_pack_matches_oracle(stride):
    nt, n = (2, 3, 1), (8, 8, 8)
    inputs, (conf, grid, sim) = random_grid(nt, n, seed=62, umax=5.0)
    for nspecies in (2, 1):
        wr = S.MpiioSpectraWriter("unused", *[v for a in zip(nt, n) for v in a], stride, 60, 1e-2, 1e2, nspecies)
        got = wr.pack(grid).cpu().numpy()
        pos, vel = tile_parts(grid, 2)
        ref = osnap.spectra_pack([[p[0] for p in tp] for tp in pos], vel, nt, n, stride, 60, 1e-2, 1e2, nspecies)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), nspecies
        assert got.sum() == 4 * sum(len(p[s][0]) for p in pos for s in range(nspecies))


def test_interpolate_at_matches_oracle(orc):
    nt, n = (2, 2, 1), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=63, ppc=2)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    import torch
    from runko_amd import _lib as L
    world = grid.world
    for t, tile in enumerate(grid.tiles()):
        for s in range(2):
            p = torch.tensor(np.stack(tile.get_positions(s)), device=world.dev).contiguous()
            k = p.shape[1]
            E = torch.empty((3, k), device=world.dev)
            B = torch.empty((3, k), device=world.dev)
            L.check(L.lib().rk_interpolate_at(world.h, t, k, L.p3(p), L.p3(E), L.p3(B), S._stream()))
            got = torch.cat([E, B]).t().double().cpu().numpy()
            ref = w.interpolate(t, s)
            assert np.array_equal(got, ref), (t, s)


def write_fields(grid, conf, outdir, stride=1, nspecies=2, lap=0):
    n, nt = conf.n_cells_per_tile, conf.n_tiles
    wr = runko.emf.threeD.MpiioFieldsWriter(str(outdir), nt[0], n[0], nt[1], n[1], nt[2], n[2], stride, nspecies)
    assert wr.write(grid._corgi_grid, lap)
    path = os.path.join(outdir, f"flds_{lap}.bin")
    return S.read_header(path), S.read_field_snapshot(path)


def uniform_grid(nt, n, ppc=1, nsp=2, vel=(0.0, 0.0, 0.0), E=(1, 2, 3), B=(4, 5, 6), J=(7, 8, 9)):
    conf = runko.Configuration(None)
    conf.n_tiles, conf.n_cells_per_tile, conf.cfl = list(nt), list(n), 0.45
    conf.field_propagator, conf.particle_pusher = "fdtd2", "boris"
    conf.field_interpolator, conf.current_depositer = "linear_1st", "zigzag_1st"
    conf.tile_partitioning, conf.n_laps = "hilbert_curve", 1
    for s in range(nsp):
        setattr(conf, f"q{s}", -1.0 if s == 0 else 1.0)
        setattr(conf, f"m{s}", 1.0)
    P = runko.pic.threeD.ParticleState
    grid = runko.TileGrid(conf)
    for idx in grid.local_tile_indices():
        tile = runko.pic.threeD.Tile(idx, conf)
        tile.set_EBJ(lambda x, y, z: E, lambda x, y, z: B, lambda x, y, z: J)
        for s in range(nsp):
            tile.inject_to_each_cell(s, lambda x, y, z: [P(pos=(x + 0.5, y + 0.5, z + 0.5), vel=vel)] * ppc)
        grid.add_tile(tile, idx)
    grid.configure_simulation(conf)
    return conf, grid


def test_fields_file_behaviours(tmp_path):
    # single tile values, header completeness, stride, densities, species counts
    conf, grid = uniform_grid((2, 2, 1), (8, 8, 8), ppc=4)
    h, f = write_fields(grid, conf, tmp_path, stride=2, nspecies=2, lap=3)
    assert (h["magic"], h["version"], h["header_size"], h["dtype_size"], h["lap"]) == (S.MAGIC, 3, 512, 4, 3)
    assert (h["nx"], h["ny"], h["nz"], h["stride"]) == (8, 8, 4, 2)
    assert (h["Nx"], h["Ny"], h["Nz"], h["NxMesh"], h["NyMesh"], h["NzMesh"]) == (2, 2, 1, 8, 8, 8)
    assert h["field_names"] == ["ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "n0", "n1"]
    for name, v in zip(("ex", "ey", "ez", "bx", "by", "bz"), (1, 2, 3, 4, 5, 6)):
        assert f[name].shape == (4, 8, 8) and np.all(f[name] == v)
    for name, v in zip(("jx", "jy", "jz"), (7, 8, 9)):
        assert np.all(f[name] == v * 8)
    assert np.all(f["n0"] == 32) and np.all(f["n1"] == 32)
    h, f = write_fields(grid, conf, tmp_path, stride=1, nspecies=5, lap=0)
    assert h["num_fields"] == 14 and np.all(f["n2"] == 0) and np.all(f["n4"] == 0) and np.all(f["n0"] == 4)
    assert S.read_field(os.path.join(tmp_path, "flds_0.bin"), "jy").shape == (8, 16, 16)
    h, f = write_fields(grid, conf, tmp_path, stride=8, nspecies=1, lap=1)
    assert (h["nx"], h["ny"], h["nz"]) == (2, 2, 1) and np.all(f["jx"] == 7 * 512) and np.all(f["n0"] == 4 * 512)


def test_particle_file_behaviours(tmp_path):
    conf, grid = uniform_grid((2, 2, 1), (8, 8, 8), ppc=1, vel=(1.0, 2.0, 3.0))
    total = 4 * 8**3
    # all particles when n_prtcls exceeds the total; velocities, global positions
    wr = runko.emf.threeD.MpiioParticlesWriter(str(tmp_path), total * 10, 0)
    assert wr.write(grid._corgi_grid, 0)
    h = S.read_prtcl_header(os.path.join(tmp_path, "prtcls_0_0.bin"))
    f = S.read_prtcl_snapshot(os.path.join(tmp_path, "prtcls_0_0.bin"))
    assert (h["magic"], h["version"], h["num_fields"], h["species"], h["n_prtcls"]) == (S.MAGIC, 3, 12, 0, total)
    assert h["field_names"] == list(S.PRTCL_FIELD_NAMES)
    assert np.all(f["ux"] == 1.0) and np.all(f["uy"] == 2.0) and np.all(f["uz"] == 3.0)
    assert f["x"].min() == 0.5 and f["x"].max() == 15.5 and f["y"].max() == 15.5 and f["z"].max() == 7.5
    # uniform fields interpolate to themselves (halos synced first: the
    # stencil of particles next to the world edge reads them)
    grid.world.comm_local(int(runko.tools.comm_mode.emf_E))
    grid.world.comm_local(int(runko.tools.comm_mode.emf_B))
    assert wr.write(grid._corgi_grid, 0)
    f = S.read_prtcl_snapshot(os.path.join(tmp_path, "prtcls_0_0.bin"))
    for name, v in zip(("ex", "ey", "ez", "bx", "by", "bz"), (1, 2, 3, 4, 5, 6)):
        assert np.allclose(f[name], v, atol=1e-6)
    # sampling reduces the count by the per-tile quota rule
    wr = runko.emf.threeD.MpiioParticlesWriter(str(tmp_path), 300, 1)
    assert wr.write(grid._corgi_grid, 2)
    h = S.read_prtcl_header(os.path.join(tmp_path, "prtcls_1_2.bin"))
    assert h["species"] == 1 and h["lap"] == 2
    assert h["n_prtcls"] == sum(k for _, k in osnap.sample_plan([512] * 4, 300))


def test_particle_sample_values_match_tiles(tmp_path, orc):
    # every sampled row is a particle of the grid, with E, B interpolated at it
    nt, n = (2, 1, 2), (8, 8, 8)
    inputs = random_inputs(nt, n, 2, seed=64, ppc=2)
    conf, grid, sim = build_gpu(nt, n, Q, M, inputs)
    w = build_oracle(orc, nt, n, Q, M, inputs)
    wr = runko.emf.threeD.MpiioParticlesWriter(str(tmp_path), 700, 1)
    assert wr.write(grid._corgi_grid, 0)
    f = S.read_prtcl_snapshot(os.path.join(tmp_path, "prtcls_1_0.bin"))
    rows = np.stack([f[k] for k in S.PRTCL_FIELD_NAMES], 1).astype(np.float64)
    known = {}
    for t in range(w.ntiles):
        st, _ = w.get_particles(t, 1)
        eb = w.interpolate(t, 1)
        for r, e in zip(st, eb):
            known[tuple(r.astype(F32))] = e
    counts = [len(tile.get_ids(1)) for tile in grid.tiles()]
    assert len(rows) == sum(k for _, k in osnap.sample_plan(counts, 700))
    for r in rows:
        e = known[tuple(r[:6].astype(F32))]
        assert np.array_equal(r[6:], e)


def test_empty_species_file(tmp_path):
    conf, grid = uniform_grid((1, 1, 1), (8, 8, 8), ppc=1, nsp=1)
    wr = runko.emf.threeD.MpiioParticlesWriter(str(tmp_path), 1000, 1)
    assert wr.write(grid._corgi_grid, 0)
    h = S.read_prtcl_header(os.path.join(tmp_path, "prtcls_1_0.bin"))
    assert h["n_prtcls"] == 0 and h["species"] == 1


def test_spectra_file_behaviours(tmp_path):
    conf, grid = uniform_grid((2, 1, 1), (8, 8, 8), ppc=1, vel=(3.0, 4.0, 0.0))
    wr = runko.emf.threeD.MpiioSpectraWriter(str(tmp_path), 2, 8, 1, 8, 1, 8, 2, 100, 1e-2, 1e4, 2)
    assert wr.write(grid._corgi_grid, 5)
    path = os.path.join(tmp_path, "pspectra_5.bin")
    h = S.read_spectra_header(path)
    f = S.read_spectra_snapshot(path)
    assert (h["nx"], h["ny"], h["nz"], h["nbins"], h["lap"]) == (8, 1, 1, 100, 5)
    assert h["umin"] == np.float32(1e-2) and h["umax"] == np.float32(1e4)
    assert h["field_names"] == ["s0_u", "s0_bx", "s0_by", "s0_bz", "s1_u", "s1_bx", "s1_by", "s1_bz"]
    tot = f["s0_u"].sum(axis=(0, 1, 2))
    assert tot[np.searchsorted(S.u_bin_edges(h), 5.0) - 1] == 1024 and tot.sum() == 1024
    assert np.all(f["s1_u"].sum(axis=(0, 1, 3)) == 128)  # 2 x 8 x 8 particles per x bin of stride 2
    # the layout rows of tile (i, j, k): with Ny > 1 each tile row has its own block
    conf, grid = uniform_grid((1, 2, 1), (8, 8, 8), ppc=1, vel=(0.5, 0.0, 0.0))
    wr = runko.emf.threeD.MpiioSpectraWriter(str(tmp_path), 1, 8, 2, 8, 1, 8, 1, 20, 1e-2, 1e2, 1)
    assert wr.write(grid._corgi_grid, 6)
    f = S.read_spectra_snapshot(os.path.join(tmp_path, "pspectra_6.bin"))
    assert f["s0_u"].shape == (1, 2, 8, 20)
    assert np.all(f["s0_u"].sum(axis=3) == 64)


def test_simulation_io_snapshots(tmp_path):
    conf, grid = uniform_grid((2, 1, 1), (8, 8, 8), ppc=1, vel=(0.1, 0.0, 0.0))
    conf.io_outdir = str(tmp_path)
    conf.io_grid_stride = 2
    conf.n_sampled_prtcls = 100
    sim = grid.configure_simulation(conf)

    def lap(x):
        x.io_emf_snapshot()
        x.io_prtcl_snapshot()
        x.io_spectra_snapshot()

    sim.for_one_lap(lap)
    names = sorted(os.listdir(tmp_path))
    assert "flds_0.bin" in names and "prtcls_0_0.bin" in names and "prtcls_1_0.bin" in names
    assert "pspectra_0.bin" in names
    h = S.read_spectra_header(os.path.join(tmp_path, "pspectra_0.bin"))
    assert (h["nbins"], h["stride"]) == (200, 2)
