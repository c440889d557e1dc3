"""TileGrid: the grid of tiles on one GPU (runko/tile_grid.py:16-227).

Tiles are built and initialised stand-alone, registered with add_tile, and
moved into one grid world by configure_simulation: their fields are copied
into the tile windows of the grid lattice and their particles into the grid's
brick bins (sized once from the per-brick counts), after which every tile
object is a handle into the grid world.
"""

import itertools
import logging

import numpy as np
import torch

from .simulation import Simulation
from .world import HALO, World

_logger = logging.getLogger("runko_amd.TileGrid")


class TileGrid:
    def __init__(self, conf):
        for var in ("n_tiles", "n_cells_per_tile", "tile_partitioning"):
            if getattr(conf, var) is None:
                raise RuntimeError(f"Can not construct TileGrid without: {var}")
        self._Nx, self._Ny, self._Nz = (int(v) for v in conf.n_tiles)
        self._NxMesh, self._NyMesh, self._NzMesh = (int(v) for v in conf.n_cells_per_tile)
        if conf.tile_partitioning not in ("hilbert_curve", "catepillar_track"):
            raise RuntimeError(f"invalid `tile_partitioning`: {conf.tile_partitioning}")
        if conf.tile_partitioning == "catepillar_track" and conf.catepillar_track_length is None:
            raise RuntimeError("Catepillar track requested, but no `catepillar_track_length` given!")
        self._tiles = {}
        self.world = None

    @property
    def n_tiles(self):
        return (self._Nx, self._Ny, self._Nz)

    def add_tile(self, tile, tile_grid_idx):
        idx = tuple(int(v) for v in tile_grid_idx)
        if any(i < 0 or i >= n for i, n in zip(idx, self.n_tiles)):
            raise RuntimeError(f"Trying to add tile outside of the grid: {idx}")
        self._tiles[idx] = tile

    def initialized_from_restart_file(self):
        return False

    def local_tile_indices(self):
        yield from itertools.product(range(self._Nx), range(self._Ny), range(self._Nz))

    def tiles(self):
        """Tiles in layout_right order of their grid index."""
        return [self._tiles[i] for i in self.local_tile_indices()]

    def build_world(self, brick=None):
        """Create the grid world and move every registered tile into it."""
        if self.world is not None:
            return self.world
        missing = [i for i in self.local_tile_indices() if i not in self._tiles]
        if missing:
            raise RuntimeError(f"tile grid is missing tiles, e.g. {missing[0]}")
        tiles = self.tiles()
        t0 = tiles[0]
        workgroup = None
        if brick == "auto":
            brick, workgroup = self._tuned_shape(tiles)
        world = World(t0.n, self.n_tiles, (0, 0, 0), self.n_tiles, periodic=True, cfl=t0.cfl,
                      charges=t0._charges, masses=t0._masses, current_filter=t0.current_filter or "binomial2",
                      brick=brick)
        world.set_schemes(getattr(t0, "particle_pusher", "boris"), t0._stencil)
        if workgroup:
            world.set_fused_workgroup(workgroup)
        torch.cuda.synchronize()
        # fields
        for t, tile in enumerate(tiles):
            src = tile._world
            dst = world.tile_interior(t)
            inner = (slice(HALO, -HALO),) * 3
            world.F[0][(slice(None),) + dst] = src.F[0][(slice(None),) + inner]
            world.F[1][(slice(None),) + dst] = src.F[1][(slice(None),) + inner]
            world.J[(slice(None),) + dst] = src.J[(slice(None),) + inner]
        # particles: size the bins once from the per-brick counts, then append
        for s in range(world.nsp):
            parts = []
            demand = np.zeros(world.nbricks, dtype=np.int64)
            for t, tile in enumerate(tiles):
                pos, vel, ids = tile._world.tile_particles(0, s)
                if ids.numel():
                    b = world.bricks_of(t, pos)
                    demand += np.bincount(b.cpu().numpy(), minlength=world.nbricks)
                parts.append((pos.clone(), vel.clone(), ids.clone()))
            world.presize(s, demand)
            world.ensure_migration(max(int(max((p[2].numel() for p in parts), default=0)), 1))
            for t, (pos, vel, ids) in enumerate(parts):
                # storage order inside a bin is free (the reference's is unpinned
                # too); a random order keeps the lanes of a wave on different
                # cells, so their LDS current atomics rarely collide
                perm = torch.randperm(ids.numel(), device=ids.device)
                world.inject(t, s, pos[:, perm], vel[:, perm], ids[perm])
        torch.cuda.synchronize()
        for t, tile in enumerate(tiles):
            tile._rebind(world, t)
        total = world.total_particles()
        world.ensure_migration(max(1 << 16, int(0.12 * total)))
        self.world = world
        return world

    @staticmethod
    def _tuned_shape(tiles):
        """The autotuner's brick and workgroup size for cubic tiles at the
        grid's mean ppc (the rule of choose_brick otherwise)."""
        n = tuple(tiles[0].n)
        if len(set(n)) != 1:
            return None, None
        from .autotune import tune_shape
        nsp = max(1, len(tiles[0]._charges))
        total = sum(tile._world.tile_count(0, s) for tile in tiles for s in range(len(tiles[0]._charges)))
        ppc = max(1, int(round(total / (np.prod(n) * len(tiles) * nsp))))
        brick, wg, _ = tune_shape(n[0], ppc=ppc)
        return brick, wg

    @property
    def _corgi_grid(self):
        """The grid the snapshot writers take (the reference passes its corgi grid)."""
        return self

    def configure_simulation(self, config):
        if getattr(config, "n_laps") is None:
            raise RuntimeError("Can not configure simulation without: n_laps")
        self.build_world(brick=getattr(config, "brick_shape"))
        nspecies = 0
        for i in range(6):
            if getattr(config, f"q{i}") is not None and getattr(config, f"m{i}") is not None:
                nspecies += 1
            else:
                break
        outdir = getattr(config, "io_outdir") or "runko-output"
        # runko/tile_grid.py:200-209
        stride = getattr(config, "io_grid_stride") or 1
        io_config = dict(outdir=outdir, nspecies=nspecies, stride=stride,
                         n_prtcls=getattr(config, "n_sampled_prtcls") or 0,
                         laps_in_timer_statistics=getattr(config, "io_n_laps_in_timer_stats"),
                         spectra_nbins=getattr(config, "io_n_spectra_bins") or 200,
                         spectra_umin=getattr(config, "io_spectra_umin") or 1e-4,
                         spectra_umax=getattr(config, "io_spectra_umax") or 1e3,
                         spectra_stride=getattr(config, "io_spectra_stride") or stride)
        mode = getattr(config, "execution") or "plain"
        return Simulation(self, Simulation._im_not_user, Nt=config.n_laps, io_config=io_config,
                          verbose=config.verbose, mode=mode)
