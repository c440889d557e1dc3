"""Builders of identical inputs for the GPU path (runko_amd) and the oracle."""

import numpy as np

import runko_amd as runko  # noqa: F401
from runko_amd import workloads

CFL = 0.45


make_config = workloads.make_configuration


def random_inputs(nt, n, nsp, seed, ppc=4, field_amp=0.05, umax=1.5):
    """Per tile: fields [9, n] (J zero) and per species particle states [k, 6]."""
    rng = np.random.default_rng(seed)
    tiles = {}
    for i in range(nt[0]):
        for j in range(nt[1]):
            for k in range(nt[2]):
                F = np.zeros((9,) + tuple(n))
                F[0:6] = field_amp * rng.standard_normal((6,) + tuple(n))
                parts = []
                cells = np.prod(n)
                lo = np.array([i * n[0], j * n[1], k * n[2]], dtype=np.float64)
                base = np.stack(np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij"), -1)
                base = base.reshape(-1, 3)
                for s in range(nsp):
                    pos = np.repeat(base, ppc, 0) + rng.random((cells * ppc, 3)) + lo
                    u = (rng.random((cells * ppc, 3)) * 2 - 1) * umax
                    parts.append(np.concatenate([pos, u], 1))
                tiles[(i, j, k)] = (F, parts)
    return tiles


build_gpu = workloads.build_simulation


def build_oracle(orc, nt, n, charges, masses, inputs, prec=32, filt="binomial2", comm_prelude=True, pusher="boris",
                 stencil=None):
    """stencil: the configuration keys (stencil_<name>, stencil_<axis>_<name>)."""
    w = orc.OracleWorld(nt, n, CFL, charges, masses, prec=prec, current_filter=filt)
    w.set_pusher(pusher)
    if stencil is not None:
        per_axis = []
        for ax in "xyz":
            d = {}
            for key, v in stencil.items():
                name = key[len("stencil_"):]
                if name.startswith(ax + "_"):
                    d[name[2:]] = v
                elif name[:2] not in ("x_", "y_", "z_"):
                    d.setdefault(name, v)
            per_axis.append(d)
        w.set_stencil(orc.stencil_matrices(per_axis))
    for idx, (F, parts) in inputs.items():
        t = w.tile_index(*idx)
        w.set_fields(t, F)
        for s, st in enumerate(parts):
            w.add_particles(t, s, st)
    if comm_prelude:
        w.comm_local("emf_E")
        w.comm_local("emf_B")
    return w


def gpu_particles(grid, s):
    """(tile index per particle, states [k,6] float64 of the float32 values, ids) sorted by id."""
    ts, sts, ids = [], [], []
    for t, tile in enumerate(grid.tiles()):
        x, y, z = tile.get_positions(s)
        u, v, w = tile.get_velocities(s)
        i = tile.get_ids(s)
        ts.append(np.full(len(i), t))
        sts.append(np.stack([x, y, z, u, v, w], 1).astype(np.float64))
        ids.append(i)
    ts, sts, ids = np.concatenate(ts), np.concatenate(sts), np.concatenate(ids)
    o = np.argsort(ids, kind="stable")
    return ts[o], sts[o], ids[o]


def oracle_particles(w, s):
    ts, sts, ids = w.all_particles(s)
    o = np.argsort(ids, kind="stable")
    return ts[o], sts[o], ids[o]


def gpu_fields(grid):
    """[9, Nx, Ny, Nz] interior fields assembled from the tiles."""
    tiles = grid.tiles()
    n = tiles[0].n
    nt = grid.n_tiles
    out = np.zeros((9,) + tuple(a * b for a, b in zip(nt, n)))
    for t, tile in enumerate(tiles):
        i, j, k = tile.index
        (E, B, J) = tile.get_EBJ()
        arr = np.stack(list(E) + list(B) + list(J))
        out[:, i * n[0]:(i + 1) * n[0], j * n[1]:(j + 1) * n[1], k * n[2]:(k + 1) * n[2]] = arr
    return out


def canonical_lap(sim, io=False, sort_every=5):
    """projects/pic-turbulence/pic.py:187-221 without snapshot IO."""
    return workloads.turbulence_lap(sim, io=io, sort_every=sort_every)
