"""GPU-side initial conditions for large worlds.

Host-side injection (Tile.batch_inject_to_cells with a Python generator per
tile) is the reference's way in (src/runko/pic/tile.c++:207-322); at 4.3e9
particles it would dominate the run, so this module fills a grid world's
pools directly on the device:
  * positions uniform in every cell (ppc per species), species 1 on top of
    species 0 as in projects/pic-turbulence/pic.py:141-156,
  * momenta from the Juttner-Synge distribution (Sobol rejection for
    theta > 0.2, Rayleigh magnitude below, isotropic direction), the same
    algorithm as runko/sample_thermal_distributions.py:7-127, drawn with
    torch's Philox generator (statistically, not bitwise, equal),
  * ids (tile_tag << 40) | ordinal, ordinals counted brick by brick,
  * a guide field B0 along z plus n x n sinusoidal dBx, dBy modes with
    seeded random phases (the discrete curl of a vector potential, so the
    discrete div B is zero), normalised to rms(dB) = dB_over_B0 * B0;
  * for the reflecting-wall shock: a cold flow along -x (rest-frame
    Juttner-Synge momenta boosted by the flow's Lorentz factor with the
    reference's flipping method), a density ramp in front of the wall, the
    upstream perpendicular field and its motional E.
"""

import math

import numpy as np
import torch
from scipy.special import kn

from .world import HALO, World, bin_caps


def plasma_parameters(cfl, c_omp, ppc, theta, sigma):
    """Charges and guide field of projects/pic-turbulence/pic.py:44-68
    (m0 = m1 = 1 mass-to-charge ratios, oppc = 2 ppc)."""
    oppc = 2 * ppc
    omp = cfl / c_omp
    q0 = -(omp**2) / (0.5 * oppc * (1.0 + 1.0))
    m0hat = abs(q0)
    B0 = math.sqrt((1.0 + 1.5 * theta) * oppc * m0hat * cfl**2 * sigma)
    return q0, -q0, B0


def juttner_momenta(n, theta, gen, device):
    """n samples (ux, uy, uz) float32 of the isotropic Juttner-Synge law."""
    if theta > 0.2:
        eff = kn(2, 1 / theta) / (2 * theta**2)
        parts, have = [], 0
        while have < n:
            m = int(1.05 * (n - have) / eff) + 4096
            x = torch.rand((4, m), generator=gen, device=device, dtype=torch.float32).clamp_min_(1e-30)
            p3 = x[0] * x[1] * x[2]
            u = -theta * torch.log(p3)
            eta = -theta * torch.log(p3 * x[3])
            keep = (eta * eta - u * u) > 1.0
            parts.append(u[keep])
            have += int(parts[-1].numel())
            del x, p3, eta, keep
        u = torch.cat(parts)[:n]
    else:
        vth = math.sqrt(2 * theta)
        u = vth * torch.sqrt(-2.0 * torch.log(torch.rand(n, generator=gen, device=device).clamp_min_(1e-30)))
    X1 = torch.rand(n, generator=gen, device=device)
    X2 = torch.rand(n, generator=gen, device=device)
    s = 2 * u * torch.sqrt(X1 * (1 - X1))
    return torch.stack([u * (2 * X1 - 1), s * torch.cos(2 * math.pi * X2), s * torch.sin(2 * math.pi * X2)])


def fill_uniform_plasma(world, ppc, theta, seed=1, chunk_bins=4096, slot_order="interleaved", layout=None):
    """Every cell gets ppc particles of every species (all species at the same
    positions), bins laid out for exactly that occupancy plus slack.

    slot_order: "interleaved" stores slot p of a bin in cell p mod (cells per
    brick), so the 64 lanes of a wave sit in 64 different, neighbouring cells
    (LDS gathers and current atomics spread over the banks); "random" uses a
    seeded random slot permutation.  layout: keyword arguments of
    World.presize_all (slack, pad, arena) for a non-default bin layout."""
    if slot_order not in ("interleaved", "random"):
        raise ValueError(f"unknown slot_order {slot_order!r}")
    dev = world.dev
    br = world.brick
    cells = br[0] * br[1] * br[2]
    per_bin = cells * ppc
    world.presize_all({s: np.full(world.nbricks, per_bin, dtype=np.int64) for s in range(world.nsp)},
                      **(layout or {}))
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    nbd = [world.n[a] // br[a] for a in range(3)]
    cl = torch.stack(torch.meshgrid(*(torch.arange(v, device=dev) for v in br), indexing="ij"), -1).reshape(-1, 3)
    # a fixed random slot order inside every bin: consecutive particles (and
    # so the lanes of a wave) sit in different cells
    slot_perm = torch.randperm(per_bin, generator=gen, device=dev)
    for b0 in range(0, world.nbricks, chunk_bins):
        b1 = min(world.nbricks, b0 + chunk_bins)
        b = torch.arange(b0, b1, device=dev)
        tile = b // world.bpt
        lb = b % world.bpt
        lbc = torch.stack([lb // (nbd[1] * nbd[2]), (lb // nbd[2]) % nbd[1], lb % nbd[2]], -1)
        tc = torch.stack([tile // (world.nt[1] * world.nt[2]), (tile // world.nt[2]) % world.nt[1],
                          tile % world.nt[2]], -1)
        origin = (tc + torch.tensor(world.toff, device=dev)) * torch.tensor(world.n, device=dev) + \
            lbc * torch.tensor(br, device=dev)
        nb = b1 - b0
        base = (origin[:, None, None, :] + cl[None, :, None, :]).expand(nb, cells, ppc, 3).to(torch.float64)
        pos = (base + torch.rand((nb, cells, ppc, 3), generator=gen, device=dev, dtype=torch.float64))
        # keep each float32 position inside its cell (and so inside its brick):
        # a draw just below the upper face can round up onto it
        upper = (base + 1.0).to(torch.float32)
        pos = pos.to(torch.float32)
        pos = torch.where(pos >= upper, torch.nextafter(upper, torch.zeros_like(upper)), pos)
        del upper
        if slot_order == "random":
            pos = pos.to(torch.float32).reshape(nb, per_bin, 3)[:, slot_perm, :]
        else:
            pos = pos.to(torch.float32).transpose(1, 2).reshape(nb, per_bin, 3)
        ordinal = (lb[:, None] * per_bin + torch.arange(per_bin, device=dev)[None, :])
        ids = (tile[:, None] << 40) | ordinal
        for s in range(world.nsp):
            pool = world.pools[s]
            cap = int(pool.cap[0].item())
            u = juttner_momenta(nb * per_bin, theta, gen, dev).reshape(3, nb, per_bin)
            for d in range(3):
                pool.pos[d][: world.nbricks * cap].view(world.nbricks, cap)[b0:b1, :per_bin] = pos[:, :, d]
                pool.vel[d][: world.nbricks * cap].view(world.nbricks, cap)[b0:b1, :per_bin] = u[d]
            pool.id[: world.nbricks * cap].view(world.nbricks, cap)[b0:b1, :per_bin] = ids
    for s in range(world.nsp):
        world.count[s].fill_(per_bin)
    torch.cuda.synchronize()


def set_turbulent_field(world, B0, n_modes=8, dB_over_B0=1.0, seed=2):
    """Bz = B0 everywhere; dBx, dBy are the discrete curl of
    A_z = sum_{m,n=1..n_modes} a_mn sin(m k x + n k y + phi_mn) (k = 2 pi / L,
    seeded phases) sampled on the Yee lattice (A_z on the z edges, Bx at
    (i, j+1/2, k+1/2), By at (i+1/2, j, k+1/2)):
        Bx(i, j+1/2) = A_z(i, j+1) - A_z(i, j),  By(i+1/2, j) = A_z(i, j) - A_z(i+1, j),
    so the discrete div B vanishes up to rounding.  a_mn = 1 / |k_mn| with the
    lattice wavenumber |k_mn|^2 = (2 sin(m k / 2))^2 + (2 sin(n k / 2))^2
    gives every mode the same B amplitude; the total is scaled to
    rms(dB) = dB_over_B0 * B0."""
    dev = world.dev
    N = world.N
    rng = np.random.default_rng(seed)
    phi = 2 * np.pi * rng.random((n_modes, n_modes))
    kx, ky = 2 * np.pi / N[0], 2 * np.pi / N[1]
    # A_z on the (periodic) nodes i = 0..N, j = 0..N
    X = torch.arange(N[0] + 1, device=dev, dtype=torch.float64)[:, None]
    Y = torch.arange(N[1] + 1, device=dev, dtype=torch.float64)[None, :]
    A = torch.zeros((N[0] + 1, N[1] + 1), device=dev, dtype=torch.float64)
    for m in range(1, n_modes + 1):
        for n in range(1, n_modes + 1):
            kd = math.sqrt((2 * math.sin(m * kx / 2)) ** 2 + (2 * math.sin(n * ky / 2)) ** 2)
            A += torch.sin(m * kx * X + n * ky * Y + float(phi[m - 1, n - 1])) / kd
    bx = A[:-1, 1:] - A[:-1, :-1]
    by = A[:-1, :-1] - A[1:, :-1]
    rms = math.sqrt(float((bx**2 + by**2).mean()))
    scale = dB_over_B0 * B0 / rms
    inner = (slice(HALO, HALO + N[0]), slice(HALO, HALO + N[1]), slice(HALO, HALO + N[2]))
    world.F.zero_()
    world.F[1][0][inner] = (scale * bx).to(torch.float32)[:, :, None]
    world.F[1][1][inner] = (scale * by).to(torch.float32)[:, :, None]
    world.F[1][2][inner] = float(B0)
    torch.cuda.synchronize()


def turbulence_world(N=512, tile=32, ppc=16, theta=0.3, sigma=10.0, cfl=0.45, c_omp=10.0, n_modes=8,
                     dB_over_B0=1.0, seed=1, brick=None, mig_fraction=0.01, mailbox_fraction=0.12,
                     slot_order="interleaved", layout=None):
    """Decaying-turbulence box: N^3 cells in tile^3 tiles, periodic.  layout:
    keyword arguments of World.presize_all (slack, pad, arena)."""
    nt = N // tile
    q0, q1, B0 = plasma_parameters(cfl, c_omp, ppc, theta, sigma)
    w = World((tile,) * 3, (nt,) * 3, (0, 0, 0), (nt,) * 3, periodic=True, cfl=cfl, charges=[q0, q1],
              masses=[1.0, 1.0], current_filter="binomial2", brick=brick)
    set_turbulent_field(w, B0, n_modes=n_modes, dB_over_B0=dB_over_B0, seed=seed + 1)
    # record sets and mailboxes first: the pools then size their spare arenas
    # from the device memory that is left
    cells = tile ** 3 * nt ** 3
    w.ensure_migration(max(1 << 16, int(mig_fraction * 2 * ppc * cells)))
    br = w.brick
    w.ensure_mailbox(mailbox_fraction, brick_capacity=2 * int(bin_caps([br[0] * br[1] * br[2] * ppc])[0]))
    fill_uniform_plasma(w, ppc, theta, seed=seed, slot_order=slot_order, layout=layout)
    w.comm_local(1)
    w.comm_local(2)
    torch.cuda.synchronize()
    return w, dict(q0=q0, q1=q1, B0=B0)


def shock_parameters(cfl, c_omp, ppc, gamma_flow, sigma, m0=1.0, m1=1.0):
    """Charges and upstream field of projects/pic-shock/pic.py:44-69 (field
    perpendicular to the flow, along z): (q0, q1, B0, Ey) with the motional
    E_y = -beta B0 of a flow along -x."""
    beta = math.sqrt(1.0 - 1.0 / gamma_flow**2)
    oppc = 2 * ppc
    omp = cfl / c_omp
    q0 = -(omp**2 * gamma_flow) / (0.5 * oppc * (1.0 + m0 / m1))
    q1 = abs(q0)
    m0hat, m1hat = m0 * abs(q0), m1 * abs(q1)
    B0 = math.sqrt(gamma_flow * oppc * 0.5 * cfl**2 * m0hat * (1.0 + m0hat / m1hat) * sigma)
    return q0, q1, B0, -beta * B0


def boost_flow(u, gamma_flow, gen):
    """Rest-frame 4-velocities [3, n] as the flow of a frame moving along -x
    with Lorentz factor gamma_flow, by Zenitani's flipping method
    (runko/sample_thermal_distributions.py:106-111, direction '-x'): u_x
    changes sign where -beta v_x > U (U uniform), which weights each
    particle by the flux factor of the boosted frame, then
    u_x' = -G (u_x + beta g)."""
    beta = math.sqrt(1.0 - 1.0 / gamma_flow**2)
    g = torch.sqrt(1.0 + (u * u).sum(0))
    vx = u[0] / g
    U = torch.rand(u.shape[1], generator=gen, device=u.device, dtype=u.dtype)
    ux = torch.where(-beta * vx > U, -u[0], u[0])
    out = u.clone()
    out[0] = -(gamma_flow * (ux + beta * g))
    return out


def fill_shock_plasma(world, ppc, theta, gamma_flow, walloc, ramp_len, ramp_factor, seed=1, chunk_bins=2048,
                      injloc=None):
    """Cold flow along -x in front of a wall at x = walloc: ppc particles per
    cell per species upstream, ramp_factor * ppc at the wall falling linearly
    to ppc over ramp_len cells (a compressed layer, so tiles near the wall hold
    several times the upstream occupancy), none behind the wall, and none from
    the injector front on (cells from ceil(injloc), the whole-cell rule of
    pic/tile.c++:239-249; None: the plasma fills the box).  Positions uniform
    in each cell (x >= walloc), species 1 on top of species 0, rest frame
    momenta from the Juttner-Synge law at theta boosted by gamma_flow."""
    dev = world.dev
    br = world.brick
    N = world.N
    xc = np.arange(N[0], dtype=np.float64)
    factor = 1.0 + (ramp_factor - 1.0) * np.clip(1.0 - (xc - walloc) / ramp_len, 0.0, 1.0)
    cnt_x = np.where(xc + 1.0 <= walloc, 0, np.rint(ppc * factor)).astype(np.int64)
    if injloc is not None:
        cnt_x[xc >= math.ceil(injloc)] = 0
    nbd = [world.n[a] // br[a] for a in range(3)]
    # particles per brick: bricks are tile-major, layout_right inside the tile
    b = np.arange(world.nbricks)
    tile = b // world.bpt
    lb = b % world.bpt
    tcx = tile // (world.nt[1] * world.nt[2])
    lbx = lb // (nbd[1] * nbd[2])
    x0 = (tcx + world.toff[0]) * world.n[0] + lbx * br[0]
    csum = np.concatenate([[0], np.cumsum(cnt_x)])
    demand = (csum[x0 + br[0]] - csum[x0]) * br[1] * br[2]
    world.presize_all({s: demand for s in range(world.nsp)})
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    cl = torch.stack(torch.meshgrid(*(torch.arange(v, device=dev) for v in br), indexing="ij"), -1).reshape(-1, 3)
    cnt_x_t = torch.from_numpy(cnt_x).to(dev)
    nbt = torch.tensor(nbd, device=dev)
    for b0 in range(0, world.nbricks, chunk_bins):
        b1 = min(world.nbricks, b0 + chunk_bins)
        bb = torch.arange(b0, b1, device=dev)
        t = bb // world.bpt
        lbb = bb % world.bpt
        lbc = torch.stack([lbb // (nbt[1] * nbt[2]), (lbb // nbt[2]) % nbt[1], lbb % nbt[2]], -1)
        tc = torch.stack([t // (world.nt[1] * world.nt[2]), (t // world.nt[2]) % world.nt[1], t % world.nt[2]], -1)
        origin = (tc + torch.tensor(world.toff, device=dev)) * torch.tensor(world.n, device=dev) + \
            lbc * torch.tensor(br, device=dev)
        cells = (origin[:, None, :] + cl[None, :, :]).reshape(-1, 3)      # [nb*cells, 3], brick-major
        per_cell = cnt_x_t[cells[:, 0]]
        total = int(per_cell.sum().item())
        if total == 0:
            continue
        # interleaved slot order inside every bin: slot p holds the particle
        # of rank p // (cells per brick) of cell p mod (cells per brick) where
        # the cell has one, so the lanes of a wave sit in different cells
        cpb = cl.shape[0]
        ci = torch.repeat_interleave(torch.arange(cells.shape[0], device=dev), per_cell)
        cfirst = torch.cumsum(per_cell, 0) - per_cell
        rank = torch.arange(total, device=dev) - cfirst[ci]
        key = (ci // cpb) * (int(per_cell.max().item()) * cpb) + rank * cpb + ci % cpb
        ci = ci[torch.argsort(key)]
        base = cells[ci].to(torch.float64)
        pos = base + torch.rand((total, 3), generator=gen, device=dev, dtype=torch.float64)
        pos[:, 0] = torch.clamp(pos[:, 0], min=walloc)
        # keep each float32 position inside its cell (and so inside its brick)
        upper = (base + 1.0).to(torch.float32)
        pos = pos.to(torch.float32)
        pos = torch.where(pos >= upper, torch.nextafter(upper, torch.zeros_like(upper)), pos)
        pos = pos.t().contiguous()
        brick_of = torch.repeat_interleave(bb, torch.as_tensor(demand[b0:b1], device=dev))
        first = torch.zeros(b1 - b0 + 1, dtype=torch.int64, device=dev)
        first[1:] = torch.cumsum(torch.as_tensor(demand[b0:b1], device=dev), 0)
        slot = torch.arange(total, device=dev) - first[brick_of - b0]
        # ids (tile_tag << 40) | ordinal, ordinal = (brick in tile << 20) | slot (< 2^20 per brick)
        ids = ((brick_of // world.bpt) << 40) | ((brick_of % world.bpt) << 20) | slot
        for s in range(world.nsp):
            pool = world.pools[s]
            dst = pool.start[brick_of] + slot
            u = boost_flow(juttner_momenta(total, theta, gen, dev).to(torch.float64), gamma_flow,
                           gen).to(torch.float32)
            for d in range(3):
                pool.pos[d][dst] = pos[d]
                pool.vel[d][dst] = u[d]
            pool.id[dst] = ids
    for s in range(world.nsp):
        world.count[s].copy_(torch.as_tensor(demand, dtype=torch.int32, device=dev))
    torch.cuda.synchronize()


def shock_world(nx=2048, ny=128, nz=128, tile=32, ppc=8, theta=1e-4, gamma_flow=10.0, sigma=0.1, cfl=0.45,
                c_omp=10.0, walloc=15.0, ramp_len=512.0, ramp_factor=4.0, seed=1, brick=None, mig_fraction=0.12,
                pusher="boris", stencil=None, current_filter="binomial2", injloc=None):
    """Reflecting-wall shock set-up: nx x ny x nz cells in tile^3 tiles, a cold
    gamma_flow flow along -x with a perpendicular field (B0 along z, motional
    E_y), a wall at walloc with the conducting and upstream edge conditions
    of projects/pic-shock/pic.py:124-133.  pusher, stencil ([3][3][5] float32
    matrices of the stencil B push, or None for fdtd2) and current_filter
    select the schemes.  Returns the world, the walls and the edge conditions
    (as rk structs) and the parameters."""
    from . import _lib as L
    nt = (nx // tile, ny // tile, nz // tile)
    q0, q1, B0, Ey = shock_parameters(cfl, c_omp, ppc, gamma_flow, sigma)
    w = World((tile,) * 3, nt, (0, 0, 0), nt, periodic=True, cfl=cfl, charges=[q0, q1], masses=[1.0, 1.0],
              current_filter=current_filter, brick=brick)
    if pusher != "boris" or stencil is not None:
        w.set_schemes(pusher, stencil)
    inner = (slice(HALO, HALO + w.N[0]), slice(HALO, HALO + w.N[1]), slice(HALO, HALO + w.N[2]))
    w.F.zero_()
    w.F[0][1][inner] = float(Ey)
    w.F[1][2][inner] = float(B0)
    fill_shock_plasma(w, ppc, theta, gamma_flow, walloc, ramp_len, ramp_factor, seed=seed, injloc=injloc)
    w.ensure_migration(max(1 << 16, int(mig_fraction * w.total_particles())))
    w.comm_local(1)
    w.comm_local(2)
    torch.cuda.synchronize()
    wall = L.ReflectorWall(float(np.float32(walloc)), 0.0, 1.0)
    Lx = float(nx)
    conducting = L.EdgeBC(0, 0, float(np.float32(walloc)), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0),
                          0b110, 0, 0b111)
    upstream = L.EdgeBC(0, 1, float(np.float32(Lx - 5)), (0.0, float(np.float32(Ey)), 0.0),
                        (0.0, 0.0, float(np.float32(B0))), (0.0, 0.0, 0.0), 0b111, 0b111, 0b111)
    return w, [wall], [conducting, upstream], dict(q0=q0, q1=q1, B0=B0, Ey=Ey)


class _StripeSink:
    """Stands in for the tiles of a simulation: records the stripe that
    MovingInjector.inject asks the tiles to fill."""

    def __init__(self):
        self.stripe = None

    def batch_inject_in_x_stripe(self, species, pgen, x_left, x_right):
        self.stripe = (x_left, x_right)


class _LapCounter:
    def __init__(self, lap, sink):
        self.lap = lap
        self._sink = sink

    def local_tiles(self):
        return [self._sink]


class WorldInjector:
    """The moving injector of runko/moving_injector.py:24-54 for a world
    stepped as a whole.  The front bookkeeping is runko_amd.MovingInjector's
    own (every n_inj laps the stripe [max(injloc - beta_flow n_inj cfl,
    walloc), injloc + beta_inj n_inj cfl), the front then at its right edge,
    and no more injection once the edge reaches Lx - Lx_margin); the stripe
    is filled on the GPU: the whole cells it touches (pic/tile.c++:239-249),
    ppc passes of one particle per cell per species, positions uniform in the
    cell, species 1 on top of species 0, rest-frame Juttner-Synge momenta at
    theta boosted along -x by gamma_flow (flipping method), as the shock
    driver's generators draw them (projects/pic-shock/pic.py:155-168).  Ids
    are (tile tag << 40) | a per-tile, per-species counter that starts at
    2^39, above every ordinal the set-up uses."""

    ORDINAL0 = 1 << 39

    def __init__(self, world, injloc, beta_flow, ppc, theta, gamma_flow, walloc=0.0, beta_inj=1.0, n_inj=50,
                 Lx=None, Lx_margin=10.0, seed=11):
        from .moving_injector import MovingInjector
        self.w = world
        self.front = MovingInjector(injloc=injloc, beta_inj=beta_inj, beta_flow=beta_flow, cfl=world.cfl, n_inj=n_inj,
                                    walloc=walloc, Lx_margin=Lx_margin, Lx=Lx)
        self.ppc, self.theta, self.gamma_flow = int(ppc), float(theta), float(gamma_flow)
        self.gen = torch.Generator(device=world.dev)
        self.gen.manual_seed(int(seed))
        self.next = np.full((world.ntiles, world.nsp), self.ORDINAL0, dtype=np.int64)
        self.injected = 0
        self.events = 0
        self.stripes = []

    @property
    def injloc(self):
        return self.front.injloc

    @property
    def moving(self):
        return self.front.moving

    def inject(self, lap):
        """Call once per lap with the lap counter; returns the number of
        particles injected."""
        sink = _StripeSink()
        self.front.inject(_LapCounter(lap, sink), [(s, None) for s in range(self.w.nsp)], 1)
        if sink.stripe is None:
            return 0
        n = self.fill_stripe(*sink.stripe)
        self.stripes.append(sink.stripe)
        self.injected += n
        self.events += 1
        return n

    def fill_stripe(self, x_left, x_right):
        w = self.w
        dev = w.dev
        N = w.N
        i0 = max(int(math.floor(x_left)), 0)
        i1 = min(int(math.ceil(x_right)), N[0])
        total = 0
        if i0 >= i1:
            return 0
        n1, n2 = w.n[1], w.n[2]
        for tx in range(i0 // w.n[0], (i1 - 1) // w.n[0] + 1):
            a0, a1 = max(i0, tx * w.n[0]), min(i1, (tx + 1) * w.n[0])
            for ty in range(w.nt[1]):
                for tz in range(w.nt[2]):
                    t = (tx * w.nt[1] + ty) * w.nt[2] + tz
                    cx = torch.arange(a0, a1, device=dev, dtype=torch.float64)
                    cy = torch.arange(ty * n1, (ty + 1) * n1, device=dev, dtype=torch.float64)
                    cz = torch.arange(tz * n2, (tz + 1) * n2, device=dev, dtype=torch.float64)
                    base = torch.stack(torch.meshgrid(cx, cy, cz, indexing="ij"), 0).reshape(3, 1, -1)
                    base = base.expand(3, self.ppc, base.shape[2]).reshape(3, -1)  # pass-major, like ppc calls
                    pos = base + torch.rand(base.shape, generator=self.gen, device=dev, dtype=torch.float64)
                    upper = (base + 1.0).to(torch.float32)
                    pos = pos.to(torch.float32)
                    pos = torch.where(pos >= upper, torch.nextafter(upper, torch.zeros_like(upper)), pos)
                    m = pos.shape[1]
                    w.ensure_migration(m)
                    for s in range(w.nsp):
                        u = boost_flow(juttner_momenta(m, self.theta, self.gen, dev).to(torch.float64),
                                       self.gamma_flow, self.gen).to(torch.float32)
                        ids = (int(t) << 40) | (int(self.next[t, s]) + torch.arange(m, device=dev, dtype=torch.int64))
                        self.next[t, s] += m
                        w.inject(t, s, pos, u, ids)
                    total += w.nsp * m
        return total
