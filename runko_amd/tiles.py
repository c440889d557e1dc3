"""emf / pic tile classes with runko's Python-visible method API.

A Tile is a handle (world, tile index).  Until the tile grid is configured a
tile owns a private stand-alone world (one tile, no periodic neighbours, halo
initially zero -- like a reference tile that was never communicated); after
TileGrid.configure_simulation every tile points into the grid world, where
its haloed lattice is a window of the one grid lattice.

Bindings mirrored: src/runko/bindings/pypic.c++:47-140 (pic.threeD),
src/runko/bindings/pyemf.c++:225-260 (emf.threeD).  Constructor validation and
error texts follow src/runko/emf/tile.c++:81-180 and src/runko/pic/tile.c++:118-147.
"""

import ctypes
import itertools

import numpy as np
import torch

from . import _lib
from . import _lib as L
from .world import HALO, World, _stream

SUPPORTED_PUSHERS = ("boris", "higuera_cary", "faraday")
SUPPORTED_PROPAGATORS = ("fdtd2", "stencil")

# StencilAxisCoeffs entries read from the configuration (emf/tile.c++:118-135):
# (row, column) -> key suffix; M[0][0] follows from the normalisation
_STENCIL_KEYS = (((1, 0), "delta"), ((2, 0), "gamma"), ((0, 1), "beta_p1"), ((0, 2), "beta_p2"),
                 ((1, 1), "beta2_p1"), ((1, 2), "beta2_p2"), ((2, 1), "beta3_p1"), ((2, 2), "beta3_p2"),
                 ((0, 3), "zeta_p1"), ((0, 4), "zeta_p2"), ((1, 3), "zeta2_p1"), ((1, 4), "zeta2_p2"),
                 ((2, 3), "zeta3_p1"), ((2, 4), "zeta3_p2"))


def stencil_from_config(conf):
    """[3][3][5] float32 matrices of the stencil B push: each coefficient from
    stencil_<axis>_<name>, else stencil_<name>, else 0 (emf/tile.c++:103-141);
    M[0][0] = StencilAxisCoeffs::alpha() (stencil_coefficients.h:42-60), in
    float arithmetic."""
    f = np.float32
    out = np.zeros((3, 3, 5), dtype=f)
    for a, ax in enumerate("xyz"):
        M = out[a]
        for (r, c), name in _STENCIL_KEYS:
            v = conf.get(f"stencil_{ax}_{name}")
            if v is None:
                v = conf.get(f"stencil_{name}")
            M[r, c] = f(float(v)) if v is not None else f(0)
        s = f(0)
        s = f(s + f(3) * f(M[1, 0] + f(2) * f(f(f(M[1, 1] + M[1, 2]) + M[1, 3]) + M[1, 4])))
        s = f(s + f(5) * f(M[2, 0] + f(2) * f(f(f(M[2, 1] + M[2, 2]) + M[2, 3]) + M[2, 4])))
        s = f(s + f(1) * f(2) * f(f(f(M[0, 1] + M[0, 2]) + M[0, 3]) + M[0, 4]))
        M[0, 0] = f(f(1) - s)
    return out
SUPPORTED_INTERPOLATORS = ("linear_1st", "linear_1st_unrolled")
SUPPORTED_DEPOSITERS = ("zigzag", "zigzag_1st", "zigzag_1st_atomic")
SUPPORTED_FILTERS = ("binomial2", "binomial2_unrolled")


class _Conf:
    """Read-only view of a configuration object's __dict__ (ConfigParser)."""

    def __init__(self, config):
        self.d = dict(getattr(config, "__dict__", {}))

    def get(self, key):
        v = self.d.get(key)
        return v

    def require(self, key):
        v = self.d.get(key)
        if v is None:
            raise RuntimeError(f"Required configuration value missing: {key}")
        return v


def _species_from_config(conf):
    """q{i}/m{i} pairs read until the first missing pair (tile.c++:28-71)."""
    qs, ms = [], []
    for i in itertools.count():
        q, m = conf.get(f"q{i}"), conf.get(f"m{i}")
        if q is None or m is None:
            break
        qs.append(float(q))
        ms.append(float(m))
    return qs, ms


class ParticleState:
    """One particle (pos, vel) in double precision (particles_common.h:25-34)."""

    def __init__(self, pos, vel):
        self.pos = tuple(float(v) for v in pos)
        self.vel = tuple(float(v) for v in vel)


class ParticleStateBatch:
    """A batch of particles: pos = (x, y, z), vel = (ux, uy, uz) 1D arrays."""

    def __init__(self, pos, vel):
        self.pos = tuple(np.asarray(v, dtype=np.float64) for v in pos)
        self.vel = tuple(np.asarray(v, dtype=np.float64) for v in vel)


class reflector_wall:
    """pic::reflector_wall (src/runko/pic/reflector_wall.h:13-25, bound at
    src/runko/bindings/pypic.c++:72-85): a conducting piston normal to x at
    walloc moving with betawall (Lorentz factor gammawall).  float32 fields."""

    def __init__(self, walloc, betawall=0.0, gammawall=1.0):
        self.walloc = float(np.float32(walloc))
        self.betawall = float(np.float32(betawall))
        self.gammawall = float(np.float32(gammawall))

    def _copy(self):
        return reflector_wall(self.walloc, self.betawall, self.gammawall)


class edge_bc:
    """emf::edge_bc (src/runko/emf/edge_bc.h:22-38, bound at
    src/runko/bindings/pyemf.c++:172-205): set E, B or J components in the
    region left (side 0) or right (side 1) of a global position along one axis."""

    def __init__(self, *, direction=0, side=0, position=0.0, Ex=0.0, Ey=0.0, Ez=0.0, Bx=0.0, By=0.0, Bz=0.0,
                 Jx=0.0, Jy=0.0, Jz=0.0, E_components=0b111, B_components=0b111, J_components=0b111):
        self.direction = int(direction)
        self.side = int(side)
        self.position = float(np.float32(position))
        self.Ex, self.Ey, self.Ez = (float(np.float32(v)) for v in (Ex, Ey, Ez))
        self.Bx, self.By, self.Bz = (float(np.float32(v)) for v in (Bx, By, Bz))
        self.Jx, self.Jy, self.Jz = (float(np.float32(v)) for v in (Jx, Jy, Jz))
        self.E_components = int(E_components)
        self.B_components = int(B_components)
        self.J_components = int(J_components)

    def _copy(self):
        c = edge_bc()
        c.__dict__.update(self.__dict__)
        return c

    def _struct(self):
        r = L.EdgeBC()
        r.direction, r.side, r.position = self.direction, self.side, self.position
        r.E[:] = [self.Ex, self.Ey, self.Ez]
        r.B[:] = [self.Bx, self.By, self.Bz]
        r.J[:] = [self.Jx, self.Jy, self.Jz]
        r.E_components, r.B_components, r.J_components = self.E_components, self.B_components, self.J_components
        return r


def _walls_array(walls):
    arr = (L.ReflectorWall * max(len(walls), 1))()
    for i, w in enumerate(walls):
        arr[i].walloc, arr[i].betawall, arr[i].gammawall = w.walloc, w.betawall, w.gammawall
    return arr


class EmfTile:
    """emf.threeD.Tile"""

    _has_particles = False

    def __init__(self, tile_grid_idx, config):
        conf = _Conf(config)
        n = [int(v) for v in conf.require("n_cells_per_tile")]
        if any(v < HALO for v in n):
            raise RuntimeError(f"Yee Lattice extents ({n[0]}, {n[1]}, {n[2]}) are assumed to be at least halo size: 3")
        self.cfl = float(conf.require("cfl"))
        fp = conf.require("field_propagator")
        if fp not in SUPPORTED_PROPAGATORS:
            raise RuntimeError(f"{fp} is not supported field propagator.")
        self.field_propagator = fp
        self._stencil = stencil_from_config(conf) if fp == "stencil" else None
        cf = conf.get("current_filter")
        if cf is not None and cf not in SUPPORTED_FILTERS:
            raise RuntimeError(f"{cf} is not supported current filter.")
        self.current_filter = cf
        tiles = [int(v) for v in conf.require("n_tiles")]
        idx = tuple(int(v) for v in tile_grid_idx)
        if any(i >= t or i < 0 for i, t in zip(idx, tiles)):
            raise RuntimeError("Trying to create tile outside of configured grid.")
        self.index = idx
        self.n = tuple(n)
        self.n_tiles = tuple(tiles)
        self.mins = [float(idx[a] * n[a]) for a in range(3)]
        self.maxs = [float((idx[a] + 1) * n[a]) for a in range(3)]
        self._init_species(conf)
        self._world = World(self.n, (1, 1, 1), idx, tiles, periodic=False, cfl=self.cfl,
                            charges=self._charges, masses=self._masses,
                            current_filter=cf or "binomial2")
        self._world.set_schemes(getattr(self, "particle_pusher", "boris"), self._stencil)
        self._t = 0
        self._edge_bcs = []
        self._walls = []

    def _init_species(self, conf):
        self._charges, self._masses = [], []

    # -------------------------------------------------------------- metadata
    @staticmethod
    def canonical_type():
        return EmfTile

    @staticmethod
    def virtual_tile_specialization():
        raise TypeError("no virtual tiles on a single process")

    def extents_wout_halo(self):
        return self.n

    def global_coordinate_map(self):
        mins = self.mins
        return lambda idx: (mins[0] + idx[0], mins[1] + idx[1], mins[2] + idx[2])

    def _rebind(self, world, t):
        self._world = world
        self._t = t

    # -------------------------------------------------------------- fields
    def _cell_coords(self):
        n = self.n
        i, j, k = np.meshgrid(np.arange(n[0], dtype=np.float64), np.arange(n[1], dtype=np.float64),
                              np.arange(n[2], dtype=np.float64), indexing="ij")
        return self.mins[0] + i, self.mins[1] + j, self.mins[2] + k

    def _write_fields(self, arrays):
        """arrays: 9 float64 [n0,n1,n2] arrays Ex..Jz (cast to float32)."""
        w = self._world
        sl = w.tile_interior(self._t)
        data = torch.from_numpy(np.stack([np.asarray(a, dtype=np.float64) for a in arrays]).astype(np.float32))
        data = data.to(w.dev)
        w.F[0][(slice(None),) + sl] = data[0:3]
        w.F[1][(slice(None),) + sl] = data[3:6]
        w.J[(slice(None),) + sl] = data[6:9]

    def set_EBJ(self, E, B, J):
        """E, B, J: callables (x, y, z) -> 3-tuple, sampled on the staggered
        Yee positions (emf/tile.c++:183-221)."""
        x, y, z = self._cell_coords()
        out = np.zeros((9,) + self.n)
        for i, j, k in itertools.product(range(self.n[0]), range(self.n[1]), range(self.n[2])):
            X, Y, Z = x[i, j, k], y[i, j, k], z[i, j, k]
            out[0, i, j, k] = E(X + 0.5, Y, Z)[0]
            out[1, i, j, k] = E(X, Y + 0.5, Z)[1]
            out[2, i, j, k] = E(X, Y, Z + 0.5)[2]
            out[3, i, j, k] = B(X, Y + 0.5, Z + 0.5)[0]
            out[4, i, j, k] = B(X + 0.5, Y, Z + 0.5)[1]
            out[5, i, j, k] = B(X + 0.5, Y + 0.5, Z)[2]
            out[6, i, j, k] = J(X + 0.5, Y, Z)[0]
            out[7, i, j, k] = J(X, Y + 0.5, Z)[1]
            out[8, i, j, k] = J(X, Y, Z + 0.5)[2]
        self._write_fields(out)

    def batch_set_EBJ(self, Ex, Ey, Ez, Bx, By, Bz, Jx, Jy, Jz):
        """Nine vectorised callables (x, y, z arrays) -> array of the tile
        shape (emf/tile.c++:224-335)."""
        x, y, z = self._cell_coords()
        xp, yp, zp = x + 0.5, y + 0.5, z + 0.5
        vals = [Ex(xp, y, z), Ey(x, yp, z), Ez(x, y, zp), Bx(x, yp, zp), By(xp, y, zp), Bz(xp, yp, z),
                Jx(xp, y, z), Jy(x, yp, z), Jz(x, y, zp)]
        for v in vals:
            if np.shape(v) != self.n:
                raise RuntimeError("Batch field setter returned array with incorrect shape!")
        self._write_fields(vals)

    def _fields_host(self, window):
        w = self._world
        torch.cuda.synchronize()
        sl = w.tile_window(self._t) if window else w.tile_interior(self._t)
        out = []
        for F in (w.F[0], w.F[1], w.J):
            a = F[(slice(None),) + sl].double().cpu().numpy()
            out.append((a[0], a[1], a[2]))
        return tuple(out)

    def get_EBJ(self):
        """((Ex,Ey,Ez),(Bx,By,Bz),(Jx,Jy,Jz)) float64 arrays of the interior."""
        return self._fields_host(False)

    def get_EBJ_with_halo(self):
        return self._fields_host(True)

    def push_half_b(self):
        self._world.call("rk_push_half_b", self._t)

    def push_e(self):
        self._world.call("rk_push_e", self._t)

    def add_current(self):
        self._world.call("rk_add_current", self._t)

    def filter_current(self):
        if self.current_filter is None:
            raise RuntimeError("Trying to filter current without specifying `current_filter`!")
        self._world.call("rk_filter_current", self._t)

    def total_energy_E(self):
        return float(self._world.energies(self._t)[0])

    def total_energy_B(self):
        return float(self._world.energies(self._t)[1])

    def register_antenna(self, *a, **k):
        raise RuntimeError("antenna currents are not part of runko_amd")

    def register_edge_bc(self, bc):
        """Tile::register_edge_bc (emf/tile.c++:831-835): stores a copy."""
        self._edge_bcs.append(bc._copy())

    def apply_edge_bc(self, bc, mode):
        """Tile::apply_edge_bc for comm_mode emf_E, emf_B or emf_J (emf/tile.c++:837-841)."""
        L.check(L.lib().rk_apply_edge_bc(self._world.h, self._t, ctypes.byref(bc._struct()), int(mode),
                                         _stream()))

    def apply_edge_bcs(self, mode):
        """Every registered edge bc in registration order (emf/tile.c++:843-847)."""
        for bc in self._edge_bcs:
            self.apply_edge_bc(bc, mode)


class PicTile(EmfTile):
    """pic.threeD.Tile"""

    _has_particles = True

    def __init__(self, tile_grid_idx, config):
        conf = _Conf(config)
        pp = conf.require("particle_pusher")
        if pp not in SUPPORTED_PUSHERS:
            raise RuntimeError(f"{pp} is not supported particle pusher.")
        fi = conf.require("field_interpolator")
        if fi not in SUPPORTED_INTERPOLATORS:
            raise RuntimeError(f"{fi} is not supported field_interpolator.")
        cd = conf.require("current_depositer")
        if cd not in SUPPORTED_DEPOSITERS:
            raise RuntimeError(f"{cd} is not supported current depositer.")
        self.particle_pusher, self.field_interpolator, self.current_depositer = pp, fi, cd
        super().__init__(tile_grid_idx, config)
        nt = self.n_tiles
        if nt[0] * nt[1] * nt[2] >= (1 << 24):
            raise RuntimeError("PIC tile does not support this many tiles.")
        i, j, k = self.index
        self.tile_tag = (i * nt[1] + j) * nt[2] + k
        self._next_ordinal = [0] * len(self._charges)

    def _init_species(self, conf):
        self._charges, self._masses = _species_from_config(conf)

    @staticmethod
    def canonical_type():
        return PicTile

    # -------------------------------------------------------------- ids
    def _consume_ids(self, s, n):
        """(tile_tag << 40) | ordinal per species (pic/tile.c++:468-481)."""
        o = self._next_ordinal[s]
        if o + n > (2 << 40):
            raise RuntimeError("PIC tile ran out of particle ids (2^40 per species)!")
        self._next_ordinal[s] = o + n
        return (np.uint64(self.tile_tag) << np.uint64(40)) | np.arange(o, o + n, dtype=np.uint64)

    def _inject_arrays(self, s, pos, vel):
        if s < 0 or s >= len(self._charges):
            raise IndexError("particle type out of range")
        n = len(pos[0])
        if n == 0:
            return
        ids = self._consume_ids(s, n)
        w = self._world
        p = torch.from_numpy(np.stack([np.asarray(v, dtype=np.float64) for v in pos]).astype(np.float32)).to(w.dev)
        u = torch.from_numpy(np.stack([np.asarray(v, dtype=np.float64) for v in vel]).astype(np.float32)).to(w.dev)
        i = torch.from_numpy(ids.view(np.int64)).to(w.dev)
        w.inject(self._t, s, p, u, i)

    def inject_to_each_cell(self, particle_type, pgen):
        x, y, z = self._cell_coords()
        pos, vel = ([], [], []), ([], [], [])
        for i, j, k in itertools.product(range(self.n[0]), range(self.n[1]), range(self.n[2])):
            for p in pgen(x[i, j, k], y[i, j, k], z[i, j, k]):
                for a in range(3):
                    pos[a].append(p.pos[a])
                    vel[a].append(p.vel[a])
        self._inject_arrays(particle_type, pos, vel)

    def inject(self, particle_type, new_particles):
        pos, vel = ([], [], []), ([], [], [])
        for p in new_particles:
            for a in range(3):
                pos[a].append(p.pos[a])
                vel[a].append(p.vel[a])
        self._inject_arrays(particle_type, pos, vel)

    def batch_inject_to_cells(self, particle_type, pgen):
        self.batch_inject_in_x_stripe(particle_type, pgen, self.mins[0], self.maxs[0])

    def batch_inject_in_x_stripe(self, particle_type, pgen, x_left, x_right):
        """pic/tile.c++:207-322: cells whose x range meets [x_left, x_right)."""
        xmin, xmax = self.mins[0], self.maxs[0]
        if x_right <= xmin or x_left >= xmax:
            return
        dl, dr = x_left - xmin, x_right - xmin
        i0 = 0 if dl <= 0.0 else min(self.n[0], int(np.floor(dl)))
        i1 = self.n[0] if dr >= float(self.n[0]) else int(np.ceil(dr))
        if i0 >= i1:
            return
        i, j, k = np.meshgrid(np.arange(i0, i1, dtype=np.float64), np.arange(self.n[1], dtype=np.float64),
                              np.arange(self.n[2], dtype=np.float64), indexing="ij")
        x = (self.mins[0] + i).ravel()
        y = (self.mins[1] + j).ravel()
        z = (self.mins[2] + k).ravel()
        b = pgen(x, y, z)
        size = len(b.pos[0])
        for a in list(b.pos) + list(b.vel):
            if np.ndim(a) != 1:
                raise RuntimeError("pic::Tile::batch_inject_in_x_stripe: given batch must be one dimensional.")
            if len(a) != size:
                raise RuntimeError("pic::Tile::batch_inject_in_x_stripe: batches must have same length.")
        self._inject_arrays(particle_type, b.pos, b.vel)

    # -------------------------------------------------------------- readback
    def _particles(self, s):
        if s < 0 or s >= len(self._charges):
            raise IndexError("particle type out of range")
        torch.cuda.synchronize()
        p, v, i = self._world.tile_particles(self._t, s)
        # id order: the order of injection, independent of the bricks' layout
        o = torch.sort(i, stable=True).indices
        return p[:, o], v[:, o], i[o]

    def get_positions(self, s):
        p, _, _ = self._particles(s)
        p = p.cpu().numpy()
        return (p[0].copy(), p[1].copy(), p[2].copy())

    def get_velocities(self, s):
        _, v, _ = self._particles(s)
        v = v.cpu().numpy()
        return (v[0].copy(), v[1].copy(), v[2].copy())

    def get_ids(self, s):
        _, _, i = self._particles(s)
        return i.cpu().numpy().view(np.uint64).copy()

    def number_of_species(self):
        return len(self._charges)

    def number_of_particles(self, s):
        """Alive particles of the species in this tile (bins never hold dead slots)."""
        return self._world.tile_count(self._t, s)

    def total_kinetic_energy(self, s):
        return float(self._world.energies(self._t)[2 + s])

    # -------------------------------------------------------------- hot methods
    def push_particles(self):
        self._world.call("rk_push_particles", self._t)

    def deposit_current(self):
        self._world.call("rk_deposit_current", self._t)

    def sort_particles(self):
        self._world.call("rk_sort_particles", self._t)

    def pack_outgoing_particles(self):
        self._world.call("rk_pack_outgoing_particles", self._t)

    def register_reflector_wall(self, wall):
        """Tile::register_reflector_wall (pic/reflector_wall.c++:240-245): stores a copy."""
        self._walls.append(wall._copy())

    def reflect_particles(self):
        """Tile::reflect_particles (pic/reflector_wall.c++:253-289)."""
        if not self._walls:
            return
        arr = _walls_array(self._walls)
        L.check(L.lib().rk_reflect_particles(self._world.h, self._t, arr, len(self._walls), _stream()))

    def advance_reflector_walls(self):
        """walloc += betawall * cfl in float32 (pic/reflector_wall.c++:293-300)."""
        cfl = np.float32(self.cfl)
        for w in self._walls:
            w.walloc = float(np.float32(w.walloc) + np.float32(w.betawall) * cfl)


class _Namespace:
    pass


def make_modules():
    pic3 = _Namespace()
    pic3.Tile = PicTile
    pic3.ParticleState = ParticleState
    pic3.ParticleStateBatch = ParticleStateBatch
    pic3.reflector_wall = reflector_wall
    emf3 = _Namespace()
    emf3.Tile = EmfTile
    emf3.edge_bc = edge_bc
    from .snapshots import MpiioFieldsWriter, MpiioParticlesWriter, MpiioSpectraWriter
    emf3.MpiioFieldsWriter = MpiioFieldsWriter
    emf3.MpiioParticlesWriter = MpiioParticlesWriter
    emf3.MpiioSpectraWriter = MpiioSpectraWriter
    return pic3, emf3
