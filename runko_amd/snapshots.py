"""Snapshot files in runko's "RNKO" v3 binary format: writers and readers.

Writers (the reference's MPI-IO writers, src/runko/io/snapshots/, bound at
src/runko/bindings/pyemf.c++:271-296; one process writes the whole file here):

* ``MpiioFieldsWriter(prefix, Nx, NxMesh, Ny, NyMesh, Nz, NzMesh, stride, nspecies=2)``
  -> ``<prefix>/flds_<lap>.bin``: ex ey ez bx by bz jx jy jz n0 .. n{k-1}, each a
  (nz, ny, nx) float32 array, nx = Nx * max(1, NxMesh // stride).  E and B are
  sampled at the stride points of every tile, J summed over stride^3 cells,
  n_s counts species s particles per coarse cell (mpiio_fields.c++:218-407).
* ``MpiioParticlesWriter(prefix, n_prtcls, species=0)``
  -> ``<prefix>/prtcls_<species>_<lap>.bin``: x y z ux uy uz ex ey ez bx by bz
  of a sample of about n_prtcls particles: per tile, quota = round(count * f)
  with f = min(1, n_prtcls / total), every (count // quota)-th particle, E and
  B interpolated at its position (mpiio_particles.c++:117-305).
* ``MpiioSpectraWriter(prefix, Nx, NxMesh, Ny, NyMesh, Nz, NzMesh, stride, nbins,
  umin, umax, nspecies=2)`` -> ``<prefix>/pspectra_<lap>.bin``: per species
  histograms of log10|u| on [umin, umax] and of beta_x, beta_y, beta_z on
  [-1, 1], per coarse x bin of every tile, each field (Nz, Ny, nx, nbins)
  (mpiio_spectra.c++:97-289).

``write(grid, lap)`` takes the TileGrid (or its ``_corgi_grid``), returns True.
The packing runs on the GPU (rk_snapshot_fields, rk_snapshot_spectra,
rk_interpolate_at); only the finished arrays cross to the host.

Two places where the reference's code and its own comments and reader
disagree follow the comments and the reader:
* particle positions are written as stored (global coordinates, as
  mpiio_particles.h:22-27 and the reader say); the reference adds the tile
  mins once more (mpiio_particles.c++:238-240), which only tile (0,0,0) of its
  tests can see;
* spectra rows of tile (i, j, k) start at ((k * Ny + j) * nx + i * nxt) * nbins
  floats (the (Nz, Ny, nx, nbins) layout of mpiio_spectra.h:24-27 and the
  reader); the reference's byte offset omits the nbins * 4 factor on the
  (k * Ny + j) * nx part (mpiio_spectra.c++:266-268), which is zero only on
  single-row grids.

Readers (runko/mpiio_reader.py, mpiio_prtcl_reader.py, mpiio_spectra_reader.py
counterparts): ``read_header``, ``read_field_snapshot``, ``read_field``,
``read_prtcl_header``, ``read_prtcl_snapshot``, ``read_spectra_header``,
``read_spectra_snapshot``, ``u_bin_edges``, ``u_bin_centers``,
``beta_bin_edges``, ``beta_bin_centers``.
"""

import ctypes
import math
import os
import struct

import numpy as np
import torch

from . import _lib as L

MAGIC = 0x524E4B4F  # "RNKO"
VERSION = 3
HEADER_SIZE = 512
MAX_SPECIES = 5
MAX_SPECTRA_SPECIES = 3
EMF_FIELD_NAMES = ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz")
PRTCL_FIELD_NAMES = ("x", "y", "z", "ux", "uy", "uz", "ex", "ey", "ez", "bx", "by", "bz")
SPECTRA_SUFFIXES = ("u", "bx", "by", "bz")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _names_block(names):
    b = bytearray(HEADER_SIZE - 64)
    for f, name in enumerate(names):
        raw = name.encode()[:15]
        b[f * 16:f * 16 + len(raw)] = raw
    return bytes(b)


def pack_header(nx, ny, nz, stride, Nx, Ny, Nz, NxMesh, NyMesh, NzMesh, lap, names):
    """The 512-byte header of field and spectra files (mpiio_header.h:55-79)."""
    head = struct.pack("<IIII" + "i" * 11 + "I", MAGIC, VERSION, HEADER_SIZE, len(names), nx, ny, nz, stride, Nx,
                       Ny, Nz, NxMesh, NyMesh, NzMesh, lap, 4)
    return head + _names_block(names)


def pack_prtcl_header(n_prtcls, species, lap):
    """Particle header: n_prtcls as int64 at 16, species at 28 (mpiio_particles.c++:23-75)."""
    head = struct.pack("<IIIIqiiiiiiiiiI", MAGIC, VERSION, HEADER_SIZE, len(PRTCL_FIELD_NAMES), int(n_prtcls), 0,
                       int(species), 0, 0, 0, 0, 0, 0, int(lap), 4)
    return head + _names_block(PRTCL_FIELD_NAMES)


def pack_spectra_header(nx, ny, nz, stride, Nx, Ny, Nz, NxMesh, NyMesh, NzMesh, lap, names, nbins, umin, umax):
    """Spectra header: field header + nbins, umin, umax at 256 (mpiio_spectra.c++:19-79)."""
    h = bytearray(pack_header(nx, ny, nz, stride, Nx, Ny, Nz, NxMesh, NyMesh, NzMesh, lap, names))
    h[256:268] = struct.pack("<iff", int(nbins), float(umin), float(umax))
    return bytes(h)


def _grid_of(grid):
    g = getattr(grid, "_tile_grid", grid)
    if getattr(g, "world", None) is None:
        raise RuntimeError("snapshot writers need a configured TileGrid (configure_simulation first)")
    return g


def _write(path, header, arrays):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "wb") as f:
        f.write(header)
        for a in arrays:
            f.write(np.ascontiguousarray(a, dtype="<f4").tobytes())


class MpiioFieldsWriter:
    def __init__(self, prefix, Nx, NxMesh, Ny, NyMesh, Nz, NzMesh, stride, nspecies=2):
        self.prefix = str(prefix)
        self.N = (int(Nx), int(Ny), int(Nz))
        self.mesh = (int(NxMesh), int(NyMesh), int(NzMesh))
        self.stride = int(stride)
        self.nspecies = min(int(nspecies), MAX_SPECIES)
        self.nt = tuple(max(1, m // self.stride) for m in self.mesh)
        self.no = tuple(a * b for a, b in zip(self.N, self.nt))

    def pack(self, grid):
        """(num_fields, nz, ny, nx) float32 device tensor of the file's arrays."""
        w = _grid_of(grid).world
        if tuple(w.nt) != self.N or tuple(w.n) != self.mesh:
            raise RuntimeError("MpiioFieldsWriter: tile grid does not match the writer's dimensions")
        out = torch.empty((9 + self.nspecies, self.no[2], self.no[1], self.no[0]), dtype=torch.float32,
                          device=w.dev)
        L.check(L.lib().rk_snapshot_fields(w.h, self.stride, self.nspecies, ctypes.c_void_p(out.data_ptr()),
                                           _stream()))
        return out

    def write(self, grid, lap):
        arr = self.pack(grid).cpu().numpy()
        names = list(EMF_FIELD_NAMES) + [f"n{s}" for s in range(self.nspecies)]
        hdr = pack_header(self.no[0], self.no[1], self.no[2], self.stride, *self.N, *self.mesh, int(lap), names)
        _write(os.path.join(self.prefix, f"flds_{int(lap)}.bin"), hdr, arr)
        return True

    write_collective = write


def sample_plan(counts, n_prtcls):
    """Per-tile (stride, sampled) of mpiio_particles.c++:126-153."""
    total = int(sum(counts))
    frac = min(1.0, float(n_prtcls) / float(total)) if total > 0 else 0.0
    plan = []
    for c in counts:
        c = int(c)
        if c == 0:
            plan.append((1, 0))
            continue
        quota = int(math.floor(c * frac + 0.5))  # std::round, half away from zero (positive here)
        if quota <= 0:
            plan.append((1, 0))
            continue
        quota = min(quota, c)
        stride = c // quota
        plan.append((stride, c // stride))
    return plan


class MpiioParticlesWriter:
    def __init__(self, prefix, n_prtcls, species=0):
        self.prefix = str(prefix)
        self.n_prtcls = int(n_prtcls)
        self.species = int(species)

    def pack(self, grid):
        """[12, n] float32 host array of the sampled particles."""
        g = _grid_of(grid)
        w = g.world
        s = self.species
        if s >= w.nsp:
            return np.zeros((12, 0), dtype=np.float32)
        counts = w.count[s].view(w.ntiles, -1).sum(1).cpu().tolist()
        plan = sample_plan(counts, self.n_prtcls)
        parts = []
        for t, (stride, n) in enumerate(plan):
            if n == 0:
                continue
            pos, vel, _ = w.tile_particles(t, s)
            pos = pos[:, ::stride][:, :n].contiguous()
            vel = vel[:, ::stride][:, :n].contiguous()
            E = torch.empty((3, n), dtype=torch.float32, device=w.dev)
            B = torch.empty((3, n), dtype=torch.float32, device=w.dev)
            L.check(L.lib().rk_interpolate_at(w.h, t, n, L.p3(pos), L.p3(E), L.p3(B), _stream()))
            parts.append(torch.cat([pos, vel, E, B]))
        if not parts:
            return np.zeros((12, 0), dtype=np.float32)
        return torch.cat(parts, 1).cpu().numpy()

    def write(self, grid, lap):
        arr = self.pack(grid)
        hdr = pack_prtcl_header(arr.shape[1], self.species, int(lap))
        _write(os.path.join(self.prefix, f"prtcls_{self.species}_{int(lap)}.bin"), hdr, arr)
        return True


class MpiioSpectraWriter:
    def __init__(self, prefix, Nx, NxMesh, Ny, NyMesh, Nz, NzMesh, stride, nbins, umin, umax, nspecies=2):
        self.prefix = str(prefix)
        self.N = (int(Nx), int(Ny), int(Nz))
        self.mesh = (int(NxMesh), int(NyMesh), int(NzMesh))
        self.stride = int(stride)
        self.nbins = int(nbins)
        self.umin = float(np.float32(umin))
        self.umax = float(np.float32(umax))
        self.nspecies = min(int(nspecies), MAX_SPECTRA_SPECIES)
        self.nxt = max(1, self.mesh[0] // self.stride)
        self.nxg = self.N[0] * self.nxt

    def pack(self, grid):
        """(4 * nspecies, Nz, Ny, nx, nbins) float32 device tensor."""
        w = _grid_of(grid).world
        if tuple(w.nt) != self.N or tuple(w.n) != self.mesh:
            raise RuntimeError("MpiioSpectraWriter: tile grid does not match the writer's dimensions")
        out = torch.empty((4 * self.nspecies, self.N[2], self.N[1], self.nxg, self.nbins), dtype=torch.float32,
                          device=w.dev)
        L.check(L.lib().rk_snapshot_spectra(w.h, self.stride, self.nbins, ctypes.c_float(self.umin),
                                            ctypes.c_float(self.umax), self.nspecies,
                                            ctypes.c_void_p(out.data_ptr()), _stream()))
        return out

    def write(self, grid, lap):
        arr = self.pack(grid).cpu().numpy()
        names = [f"s{s}_{q}" for s in range(self.nspecies) for q in SPECTRA_SUFFIXES]
        hdr = pack_spectra_header(self.nxg, self.N[1], self.N[2], self.stride, *self.N, *self.mesh, int(lap), names,
                                  self.nbins, self.umin, self.umax)
        _write(os.path.join(self.prefix, f"pspectra_{int(lap)}.bin"), hdr, arr)
        return True


# ---------------------------------------------------------------------- readers
def _check_magic(magic, version):
    if magic != MAGIC:
        raise ValueError(f"Bad magic number: expected 0x{MAGIC:08X}, got 0x{magic:08X}")
    if version != VERSION:
        raise ValueError(f"Unsupported format version: expected one of ({VERSION},), got {version}")


def _names(raw, n):
    return [raw[64 + 16 * f:64 + 16 * (f + 1)].split(b"\0", 1)[0].decode() for f in range(n)]


def read_header(path):
    with open(path, "rb") as f:
        raw = f.read(HEADER_SIZE)
    v = struct.unpack_from("<IIII" + "i" * 11 + "I", raw, 0)
    _check_magic(v[0], v[1])
    keys = ("magic", "version", "header_size", "num_fields", "nx", "ny", "nz", "stride", "Nx", "Ny", "Nz", "NxMesh",
            "NyMesh", "NzMesh", "lap", "dtype_size")
    h = dict(zip(keys, v))
    h["field_names"] = _names(raw, h["num_fields"])
    return h


def read_field_snapshot(path):
    h = read_header(path)
    n = h["nx"] * h["ny"] * h["nz"]
    data = np.fromfile(path, dtype="<f4", offset=h["header_size"], count=n * h["num_fields"])
    return {name: data[i * n:(i + 1) * n].reshape(h["nz"], h["ny"], h["nx"]).copy()
            for i, name in enumerate(h["field_names"])}


def read_field(path, field_name):
    h = read_header(path)
    if field_name not in h["field_names"]:
        raise KeyError(field_name)
    i = h["field_names"].index(field_name)
    n = h["nx"] * h["ny"] * h["nz"]
    data = np.fromfile(path, dtype="<f4", offset=h["header_size"] + 4 * i * n, count=n)
    return data.reshape(h["nz"], h["ny"], h["nx"]).copy()


def read_prtcl_header(path):
    with open(path, "rb") as f:
        raw = f.read(HEADER_SIZE)
    v = struct.unpack_from("<IIIIqiiiiiiiiiI", raw, 0)
    _check_magic(v[0], v[1])
    h = dict(magic=v[0], version=v[1], header_size=v[2], num_fields=v[3], n_prtcls=v[4], species=v[6], lap=v[13],
             dtype_size=v[14])
    h["field_names"] = _names(raw, h["num_fields"])
    return h


def read_prtcl_snapshot(path):
    h = read_prtcl_header(path)
    n = h["n_prtcls"]
    data = np.fromfile(path, dtype="<f4", offset=h["header_size"], count=n * h["num_fields"])
    return {name: data[i * n:(i + 1) * n].copy() for i, name in enumerate(h["field_names"])}


def read_spectra_header(path):
    h = read_header(path)
    with open(path, "rb") as f:
        raw = f.read(HEADER_SIZE)
    h["nbins"], h["umin"], h["umax"] = struct.unpack_from("<iff", raw, 256)
    return h


def read_spectra_snapshot(path):
    h = read_spectra_header(path)
    shape = (h["nz"], h["ny"], h["nx"], h["nbins"])
    n = int(np.prod(shape))
    data = np.fromfile(path, dtype="<f4", offset=h["header_size"], count=n * h["num_fields"])
    return {name: data[i * n:(i + 1) * n].reshape(shape).copy() for i, name in enumerate(h["field_names"])}


def u_bin_edges(hdr):
    return np.logspace(np.log10(hdr["umin"]), np.log10(hdr["umax"]), hdr["nbins"] + 1)


def u_bin_centers(hdr):
    e = u_bin_edges(hdr)
    return np.sqrt(e[:-1] * e[1:])


def beta_bin_edges(hdr):
    return np.linspace(-1.0, 1.0, hdr["nbins"] + 1)


def beta_bin_centers(hdr):
    e = beta_bin_edges(hdr)
    return 0.5 * (e[:-1] + e[1:])
