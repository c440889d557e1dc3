"""Oracle (test infrastructure only): numpy restatement of the snapshot
packers of runko's MPI-IO writers, for comparison with the GPU packing.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may use this
module; the product never imports it.

Inputs are plain arrays: interior fields of the whole grid as float32
[3, X, Y, Z] (x, y, z order, as the tile readback assembles them) and, per
tile and species, float32 particle arrays.  Float arithmetic is float32 in
the reference's order.

Citations (hel-astro-lab/runko v5.0.0):
  fields_pack     src/runko/io/snapshots/mpiio_fields.c++:218-321 (pack_tile),
                  :360-405 (placement of tile blocks in the file arrays)
  sample_plan     src/runko/io/snapshots/mpiio_particles.c++:126-153
  spectra_pack    src/runko/io/snapshots/mpiio_spectra.c++:97-209
"""

import math

import numpy as np

F32 = np.float32


def fields_pack(E, B, J, parts, n_tiles, n_cells, stride, nspecies):
    """(9 + nspecies, nz, ny, nx) float32.  parts[t][s] = [3, k] float32 global
    positions of species s in tile t (layout_right tile index)."""
    nt = [max(1, n // stride) for n in n_cells]
    no = [a * b for a, b in zip(n_tiles, nt)]
    out = np.zeros((9 + nspecies, no[2], no[1], no[0]), dtype=F32)
    # global sample coordinates: tile * n + i * stride for output index tile * nt + i
    idx = []
    for a in range(3):
        o = np.arange(no[a])
        idx.append((o // nt[a]) * n_cells[a] + (o % nt[a]) * stride)
    X, Y, Z = np.meshgrid(idx[0], idx[1], idx[2], indexing="ij")  # [ox, oy, oz]
    for c in range(3):
        out[c] = E[c][X, Y, Z].transpose(2, 1, 0)
        out[3 + c] = B[c][X, Y, Z].transpose(2, 1, 0)
    for c in range(3):
        acc = np.zeros(X.shape, dtype=F32)
        for kk in range(stride):
            for jj in range(stride):
                for ii in range(stride):
                    acc = (acc + J[c][X + ii, Y + jj, Z + kk]).astype(F32)
        out[6 + c] = acc.transpose(2, 1, 0)
    inv_stride = F32(1.0) / F32(stride)
    ntiles = n_tiles[0] * n_tiles[1] * n_tiles[2]
    for t in range(ntiles):
        tc = (t // (n_tiles[1] * n_tiles[2]), (t // n_tiles[2]) % n_tiles[1], t % n_tiles[2])
        mins = [F32(tc[a] * n_cells[a]) for a in range(3)]
        for s in range(min(nspecies, len(parts[t]))):
            p = parts[t][s]
            if p.shape[1] == 0:
                continue
            ci = [np.floor(((p[a] - mins[a]).astype(F32) * inv_stride).astype(F32)) for a in range(3)]
            ok = np.ones(p.shape[1], dtype=bool)
            for a in range(3):
                ok &= (ci[a] >= 0) & (ci[a] < nt[a])
            ox = tc[0] * nt[0] + ci[0][ok].astype(np.int64)
            oy = tc[1] * nt[1] + ci[1][ok].astype(np.int64)
            oz = tc[2] * nt[2] + ci[2][ok].astype(np.int64)
            np.add.at(out[9 + s], (oz, oy, ox), F32(1.0))
    return out


def sample_plan(counts, n_prtcls):
    """Per tile (stride, sampled): quota = round(count * min(1, n / total))."""
    total = int(sum(int(c) for c in counts))
    frac = min(1.0, float(n_prtcls) / float(total)) if total > 0 else 0.0
    plan = []
    for c in counts:
        c = int(c)
        q = math.floor(c * frac + 0.5) if c > 0 else 0
        if q <= 0:
            plan.append((1, 0))
            continue
        q = min(q, c)
        plan.append((c // q, c // (c // q)))
    return plan


def spectra_pack(parts_pos_x, parts_vel, n_tiles, n_cells, stride, nbins, umin, umax, nspecies):
    """(4 * nspecies, Nz, Ny, nx, nbins) float32.  parts_pos_x[t][s]: float32
    global x positions; parts_vel[t][s]: [3, k] float32 momenta."""
    nxt = max(1, n_cells[0] // stride)
    nxg = n_tiles[0] * nxt
    out = np.zeros((4 * nspecies, n_tiles[2], n_tiles[1], nxg, nbins), dtype=F32)
    inv_stride = F32(1.0) / F32(stride)
    # logarithms correctly rounded to float (through double), the rest float32
    log_umin = F32(np.log10(np.float64(F32(umin))))
    log_umax = F32(np.log10(np.float64(F32(umax))))
    inv_dlog = F32(nbins) / F32(log_umax - log_umin)
    inv_dbeta = F32(nbins) / F32(2.0)
    last = F32(nbins - 1)
    ntiles = n_tiles[0] * n_tiles[1] * n_tiles[2]
    for t in range(ntiles):
        tc = (t // (n_tiles[1] * n_tiles[2]), (t // n_tiles[2]) % n_tiles[1], t % n_tiles[2])
        mx = F32(tc[0] * n_cells[0])
        for s in range(min(nspecies, len(parts_vel[t]))):
            x = parts_pos_x[t][s]
            u = parts_vel[t][s]
            if x.size == 0:
                continue
            fx = np.floor(((x - mx).astype(F32) * inv_stride).astype(F32))
            ok = (fx >= 0) & (fx < nxt)
            fx, ux, uy, uz = fx[ok], u[0][ok], u[1][ok], u[2][ok]
            u2 = ((ux * ux).astype(F32) + (uy * uy).astype(F32)).astype(F32)
            u2 = (u2 + (uz * uz).astype(F32)).astype(F32)
            umag = np.sqrt(u2).astype(F32)
            gam = np.sqrt((F32(1.0) + u2).astype(F32)).astype(F32)
            ig = (F32(1.0) / gam).astype(F32)
            with np.errstate(divide="ignore"):
                lu = np.log10(umag.astype(np.float64)).astype(F32)

            def binned(raw):
                return np.clip(np.nan_to_num(raw, nan=0.0, neginf=0.0, posinf=float(last)), 0, last).astype(np.int64)

            bins = [binned(np.floor(((lu - log_umin).astype(F32) * inv_dlog).astype(F32)))]
            for comp in (ux, uy, uz):
                bins.append(binned(np.floor((((comp * ig).astype(F32) + F32(1.0)).astype(F32) * inv_dbeta)
                                            .astype(F32))))
            xb = tc[0] * nxt + fx.astype(np.int64)
            for f in range(4):
                np.add.at(out[4 * s + f, tc[2], tc[1]], (xb, bins[f]), F32(1.0))
    return out
