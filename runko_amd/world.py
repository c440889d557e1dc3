"""World: the device state behind one tile grid (or one stand-alone tile).

A world owns, as torch tensors on the GPU:
  * F: float32 [4, 3, Lx, Ly, Lz] -- E, B, J and the filter target Jf on one
    haloed lattice of L = N+6 nodes per dimension (N = tiles * cells),
  * per species a brick-binned SoA pool (pos[3], vel[3] float32, id uint64)
    with bin_start (int64), bin_cap and bin_count (uint32 stored in int32),
  * two migration record sets (tile leavers, in-tile brick changes),
and a handle to the C-ABI world that borrows them.  Bins grow on demand: an
append that finds its bin full leaves the record pending, the world repacks
that species' pool with larger bins and re-appends.
"""

import ctypes
import math
import time
import os

import numpy as np
import torch

from . import _lib as L

HALO = 3
SLACK = 1.25
PAD = 64


def choose_brick(n_cells):
    """Largest divisor <= 8 per dimension (LDS window of one workgroup)."""
    env = os.environ.get("RUNKO_AMD_BRICK")
    if env:
        b = tuple(int(v) for v in env.split(","))
        if all(n % v == 0 for n, v in zip(n_cells, b)):
            return b
    out = []
    for n in n_cells:
        for d in range(8, 0, -1):
            if n % d == 0:
                out.append(d)
                break
    return tuple(out)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def free_device_bytes(dev):
    """Device memory a new tensor can take: free on the device plus what
    torch's caching allocator holds without using it."""
    free, _ = torch.cuda.mem_get_info(dev)
    return int(free) + int(torch.cuda.memory_reserved(dev)) - int(torch.cuda.memory_allocated(dev))


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("runko_amd needs a ROCm GPU (no CPU fallback)")


SLOT_BYTES = 32       # one particle slot: pos[3], vel[3] float32 + id uint64
ARENA_TARGET = 0.5    # spare pool capacity aimed at, fraction of the laid-out bins
RESERVE_MIN = 16 << 30  # HBM left free after a layout (torch temporaries, relayout staging)
RESERVE_FRAC = 0.04   # ... or this fraction of the device, whichever is larger
STAGE_MAX = 2 << 30   # relayout staging buffer


def round32(v):
    return (np.asarray(v, dtype=np.int64) + 31) // 32 * 32


def bin_caps(demand, slack=SLACK, pad=PAD):
    """Bin capacities for a per-bin demand: demand * slack + pad, multiples
    of 32; empty bins get none."""
    d = np.asarray(demand, dtype=np.int64)
    return np.where(d > 0, round32(np.ceil(d * slack).astype(np.int64) + pad), 0)


def plan_arena(used, free_bytes, reserve_bytes, target=ARENA_TARGET):
    """Spare slots per species for pools whose bins take `used` slots.

    Every species aims at target * used spare slots.  When the device cannot
    hold that beside the bins and the reserve, the slots that fit are shared
    in proportion to `used`.  Raises if the bins themselves do not fit."""
    used = np.asarray(used, dtype=np.int64)
    total = int(used.sum())
    budget = (int(free_bytes) - int(reserve_bytes)) // SLOT_BYTES - total
    if budget < 0:
        raise RuntimeError(f"runko_amd: particle bins need {total * SLOT_BYTES / 2**30:.1f} GiB, only "
                           f"{(free_bytes - reserve_bytes) / 2**30:.1f} GiB of device memory are free beyond the "
                           f"{reserve_bytes / 2**30:.1f} GiB reserve")
    want = np.ceil(target * used).astype(np.int64)
    if int(want.sum()) <= budget:
        return want
    return budget * used // max(total, 1)


RELAYOUT_SLACK = 1.5    # bin capacity per particle after a relayout, if the arena keeps ARENA_MIN
ARENA_MIN = 0.15        # fraction of the pool left spare by that first choice
ARENA_LOW = 0.05        # a relayout leaving less spare grows the pool, if the device has room


def plan_relayout(demand, caps, grown, capacity):
    """New capacities of every bin for an in-place relayout of a pool of
    `capacity` slots.  A relayout moves every bin anyway, so bins that
    emptied give their room back.  First choice: every bin its demand times
    RELAYOUT_SLACK plus a pad (a bin whose demand exceeds its capacity at
    least `grown`), if that leaves ARENA_MIN of the pool spare for bins that
    overflow later; then the same with the set-up's slack; then full bins
    grown and the others keeping theirs; then every bin its demand plus the
    set-up's slack, then plus a pad, then exactly its demand.  Returns None
    when even the demand exceeds the pool."""
    demand = np.asarray(demand, dtype=np.int64)
    caps = np.asarray(caps, dtype=np.int64)
    full = demand > caps
    loose = bin_caps(demand, RELAYOUT_SLACK)
    bc = bin_caps(demand)
    roomy = int((1.0 - ARENA_MIN) * capacity)
    for c, limit in ((np.where(full, np.maximum(grown, loose), loose), roomy),
                     (np.where(full, np.maximum(grown, bc), bc), capacity), (np.where(full, grown, caps), capacity),
                     (bc, capacity), (np.where(demand > 0, round32(demand + PAD), 0), capacity), (demand, capacity)):
        c = np.maximum(c, np.minimum(demand, caps))
        if int(c.sum()) <= limit:
            return c
    return None


class _Pool:
    def __init__(self, nb, cap_per_bin, dev, arena=None):
        caps = np.asarray(cap_per_bin, dtype=np.int64)
        self.set_layout(caps, dev, arena)

    def set_layout(self, caps, dev, arena=None):
        starts = np.zeros_like(caps)
        starts[1:] = np.cumsum(caps)[:-1]
        # bins fill [0, used); the arena [used, capacity) takes bins that
        # outgrow their place (World.settle_overflow moves them there; when it
        # is used up the pool is relaid out in place, so buffer addresses stay)
        self.used = int(caps.sum())
        if arena is None:
            arena = int(np.ceil(ARENA_TARGET * self.used))
        self.capacity = self.used + int(arena) if self.used else 0
        self.start_h, self.cap_h = starts.copy(), caps.copy()
        n = max(self.capacity, 1)
        self.pos = torch.empty((3, n), dtype=torch.float32, device=dev)
        self.vel = torch.empty((3, n), dtype=torch.float32, device=dev)
        self.id = torch.empty((n,), dtype=torch.int64, device=dev)
        self.start = torch.from_numpy(starts).to(dev)
        self.cap = torch.from_numpy(caps.astype(np.int32)).to(dev)

    def nbytes(self):
        return max(self.capacity, 1) * SLOT_BYTES


class World:
    def __init__(self, n_cells, n_tiles, tile_offset, grid_tiles, periodic, cfl, charges, masses,
                 current_filter="binomial2", brick=None, device="cuda"):
        _require_gpu()
        self.dev = torch.device(device)
        self.n = tuple(int(v) for v in n_cells)
        self.nt = tuple(int(v) for v in n_tiles)
        self.toff = tuple(int(v) for v in tile_offset)
        self.gt = tuple(int(v) for v in grid_tiles)
        self.periodic = bool(periodic)
        self.cfl = float(cfl)
        self.charges = [float(q) for q in charges]
        self.masses = [float(m) for m in masses]
        self.nsp = len(self.charges)
        self.brick = tuple(brick) if brick else choose_brick(self.n)
        self.N = tuple(a * b for a, b in zip(self.n, self.nt))
        self.Lh = tuple(v + 2 * HALO for v in self.N)
        d = L.WorldDesc()
        for a in range(3):
            d.n_cells[a] = self.n[a]
            d.n_tiles[a] = self.nt[a]
            d.tile_offset[a] = self.toff[a]
            d.grid_tiles[a] = self.gt[a]
            d.brick[a] = self.brick[a]
        d.periodic = 1 if periodic else 0
        d.n_species = self.nsp
        self.current_filter = current_filter or "binomial2"
        d.current_filter = {"binomial2": 0, "binomial2_unrolled": 1}[self.current_filter]
        d.cfl = self.cfl
        for s in range(self.nsp):
            d.charge[s] = self.charges[s]
            d.mass[s] = self.masses[s]
        h = ctypes.c_void_p()
        L.check(L.lib().rk_world_create(ctypes.byref(h), ctypes.byref(d)))
        self.h = h
        self.F = torch.zeros((4, 3) + self.Lh, dtype=torch.float32, device=self.dev)
        L.check(L.lib().rk_world_set_fields(self.h, L.p3(self.F[0]), L.p3(self.F[1]), L.p3(self.F[2]),
                                            L.p3(self.F[3])))
        self.nbricks = L.lib().rk_world_num_bricks(self.h)
        self.bpt = self.nbricks // (self.nt[0] * self.nt[1] * self.nt[2])
        # bumped whenever a buffer the kernels take as an argument is replaced
        # (captured HIP graphs hold the old one)
        self.layout_version = 0
        self.repacks = 0
        self.fused_launch = "grid"
        self.fused_workgroup = 512
        self.bin_moves = 0  # repacks done inside the arena (no new pool)
        self.relayouts = 0  # in-place relayouts of a pool (arena used up)
        self.new_pools = 0  # pools replaced by larger ones
        # a relayout that would leave less than ARENA_LOW of its pool spare
        # allocates a larger pool instead when the device has room (False:
        # always relayout in place, so buffer addresses never change)
        self.grow_pools = True
        self.settle_seconds = 0.0  # wall time spent repacking and re-appending
        self.count = [torch.zeros((self.nbricks,), dtype=torch.int32, device=self.dev) for _ in range(self.nsp)]
        self.pools = [_Pool(self.nbricks, np.zeros(self.nbricks, dtype=np.int64), self.dev) for _ in range(self.nsp)]
        for s in range(self.nsp):
            self._bind_species(s)
        self.mig = [None, None]
        self._mig_cap = [0, 0]
        self.ensure_migration(1 << 12)
        self.mbox = None
        self._mbox_per_brick = 0
        self.jstage = None
        self.pusher = "boris"
        self.stencil = None

    # ------------------------------------------------------------------ plumbing
    PUSHERS = {"boris": 0, "higuera_cary": 1, "faraday": 2}

    def set_schemes(self, pusher="boris", stencil=None):
        """Particle pusher of every tile and the B propagator (stencil
        matrices [3][3][5] float32, or None for fdtd2)."""
        if pusher not in self.PUSHERS:
            raise RuntimeError(f"{pusher} is not supported particle pusher.")
        self.pusher = pusher
        self.stencil = None if stencil is None else np.ascontiguousarray(stencil, dtype=np.float32).reshape(3, 3, 5)
        ptr = None if self.stencil is None else self.stencil.ctypes.data
        L.check(L.lib().rk_world_set_schemes(self.h, self.PUSHERS[pusher], ptr))
    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                torch.cuda.synchronize()
                L.lib().rk_world_destroy(h)
            except Exception:
                pass
            self.h = None

    def _bind_species(self, s):
        self.layout_version += 1
        p = self.pools[s]
        L.check(L.lib().rk_world_set_species(self.h, s, p.capacity, L.p3(p.pos), L.p3(p.vel), p.id.data_ptr(),
                                             p.start.data_ptr(), p.cap.data_ptr(), self.count[s].data_ptr()))

    def ensure_migration(self, cap):
        for st in range(2):
            if self._mig_cap[st] >= cap:
                continue
            c = int(cap)
            rec = dict(pos=torch.empty((3, c), dtype=torch.float32, device=self.dev),
                       vel=torch.empty((3, c), dtype=torch.float32, device=self.dev),
                       id=torch.empty((c,), dtype=torch.int64, device=self.dev),
                       meta=torch.empty((c,), dtype=torch.int32, device=self.dev))
            torch.cuda.synchronize()
            L.check(L.lib().rk_world_set_migration(self.h, st, c, L.p3(rec["pos"]), L.p3(rec["vel"]),
                                                   rec["id"].data_ptr(), rec["meta"].data_ptr()))
            self.mig[st] = rec
            self._mig_cap[st] = c
            self.layout_version += 1

    def ensure_mailbox(self, fraction=0.12, brick_capacity=None):
        """Per-brick mailboxes for the fused particle step, sized to `fraction`
        of the largest brick's total capacity (all species; `brick_capacity`
        gives it before the pools are laid out), and the per-brick current
        staging slots (3 (br+3)^3 floats each) of the same step."""
        if self.jstage is None:
            stride = 3 * (self.brick[0] + 3) * (self.brick[1] + 3) * (self.brick[2] + 3)
            self.jstage = torch.empty((self.nbricks, stride), dtype=torch.float32, device=self.dev)
            torch.cuda.synchronize()
            L.check(L.lib().rk_world_set_current_staging(self.h, self.jstage.data_ptr(), stride))
            self.layout_version += 1
        if brick_capacity is None:
            caps = sum(p.cap.to(torch.int64) for p in self.pools)
            brick_capacity = int(caps.max().item()) if self.nsp else 0
        per = int(math.ceil(fraction * int(brick_capacity)))
        per = max(64, (per + 31) // 32 * 32)
        if per <= self._mbox_per_brick:
            return
        c = per * self.nbricks
        rec = dict(pos=torch.empty((3, c), dtype=torch.float32, device=self.dev),
                   vel=torch.empty((3, c), dtype=torch.float32, device=self.dev),
                   id=torch.empty((c,), dtype=torch.int64, device=self.dev),
                   meta=torch.empty((c,), dtype=torch.int32, device=self.dev))
        torch.cuda.synchronize()
        L.check(L.lib().rk_world_set_mailbox(self.h, per, L.p3(rec["pos"]), L.p3(rec["vel"]), rec["id"].data_ptr(),
                                             rec["meta"].data_ptr()))
        self.mbox = rec
        self._mbox_per_brick = per
        self.layout_version += 1

    @property
    def ntiles(self):
        return self.nt[0] * self.nt[1] * self.nt[2]

    def tile_coords(self, t):
        return (t // (self.nt[1] * self.nt[2]), (t // self.nt[2]) % self.nt[1], t % self.nt[2])

    def tile_interior(self, t):
        c = self.tile_coords(t)
        return tuple(slice(HALO + c[a] * self.n[a], HALO + (c[a] + 1) * self.n[a]) for a in range(3))

    def tile_window(self, t):
        c = self.tile_coords(t)
        return tuple(slice(c[a] * self.n[a], c[a] * self.n[a] + self.n[a] + 2 * HALO) for a in range(3))

    @property
    def E(self):
        return self.F[0]

    @property
    def B(self):
        return self.F[1]

    @property
    def J(self):
        """The buffer currently holding J (filter passes swap F[2] and F[3])."""
        p = L.lib().rk_world_j_pointer(self.h, 0)
        return self.F[2] if p == self.F[2][0].data_ptr() else self.F[3]

    # ------------------------------------------------------------------ methods
    def call(self, name, tile):
        f = getattr(L.lib(), name)
        rc = f(self.h, int(tile), _stream())
        if rc == L.RK_EOVERFLOW:
            self.settle_overflow()
            return
        L.check(rc)

    def filter_current_passes(self, passes):
        """`passes` filter passes over every tile from a current J ghost layer,
        the ghost layer synced between passes (rk_filter_current_passes: one
        launch for three binomial2 passes on a periodic world)."""
        L.check(L.lib().rk_filter_current_passes(self.h, int(passes), _stream()))

    def comm_local(self, mode):
        rc = L.lib().rk_comm_local(self.h, int(mode), _stream())
        if rc == L.RK_EOVERFLOW:
            self.settle_overflow()
            return
        L.check(rc)

    def set_fused_launch(self, mode):
        """"grid": one workgroup per brick; "queue": persistent workgroups
        taking bricks from per-XCD queues."""
        modes = {"grid": 0, "queue": 1}
        if mode not in modes:
            raise ValueError(f"fused launch must be one of {sorted(modes)}")
        L.check(L.lib().rk_world_set_fused_launch(self.h, modes[mode]))
        self.fused_launch = mode

    def set_walls(self, walls):
        """Hold the reflecting walls (rk_reflector_wall structs) in device
        memory for particle_step_fused_device / advance_walls."""
        arr = (L.ReflectorWall * max(len(walls), 1))(*walls)
        L.check(L.lib().rk_world_set_walls(self.h, arr, len(walls)))
        self.n_device_walls = len(walls)

    def particle_step_fused_device(self):
        """The fused particle step with the walls held in device memory."""
        if self._mbox_per_brick == 0:
            self.ensure_mailbox()
        L.check(L.lib().rk_particle_step_fused_device_walls(self.h, _stream()))

    def advance_walls(self):
        L.check(L.lib().rk_advance_walls(self.h, _stream()))

    def read_walls(self):
        """The device walls (synchronous)."""
        n = getattr(self, "n_device_walls", 0)
        arr = (L.ReflectorWall * max(n, 1))()
        L.check(L.lib().rk_read_walls(self.h, arr, n, _stream()))
        return [arr[i] for i in range(n)]

    def set_fused_workgroup(self, threads):
        """Threads per workgroup of the fused step (512 or 256)."""
        L.check(L.lib().rk_world_set_fused_workgroup(self.h, int(threads)))
        self.fused_workgroup = int(threads)

    def particle_step_fused(self, walls=None):
        """walls: optional sequence of rk_reflector_wall structs, applied to
        the tiles they lie in between the push and the classification."""
        if self._mbox_per_brick == 0:
            self.ensure_mailbox()
        if walls:
            arr = (L.ReflectorWall * len(walls))(*walls)
            L.check(L.lib().rk_particle_step_fused_walls(self.h, arr, len(walls), _stream()))
        else:
            L.check(L.lib().rk_particle_step_fused(self.h, _stream()))

    def pending_overflow(self):
        n = ctypes.c_int64(0)
        L.check(L.lib().rk_pending_overflow(self.h, ctypes.byref(n), _stream()))
        return n.value

    def settle_overflow(self, pending=None):
        """Give every bin that overflowed room, then re-append the pending
        records (synchronous; no particle is lost).

        A full bin grows to its demand plus slack and at least 1.5x, so a bin
        that keeps filling (a compressed layer) moves a logarithmic number of
        times.  It moves into its pool's spare arena while the arena has room
        (rk_move_bins); when the arena is used up the pool is relaid out in
        place (rk_relayout), which reclaims the places moved bins left.  The
        pool buffers are never reallocated unless even the relayout cannot
        fit and the device has room for a larger pool."""
        t0 = time.perf_counter()
        for _ in range(4):
            # `pending`: the caller's count, read just before (saves a sync)
            if (self.pending_overflow() if pending is None else pending) == 0:
                break
            pending = None
            for s in range(self.nsp):
                demand = self.count[s].cpu().numpy().view(np.uint32).astype(np.int64)
                pool = self.pools[s]
                caps = pool.cap_h
                idx = np.flatnonzero(demand > caps)
                if idx.size == 0:
                    continue
                g = round32(np.maximum(np.ceil(demand[idx] * SLACK).astype(np.int64) + PAD, (3 * caps[idx] + 1) // 2))
                if pool.used + int(g.sum()) <= pool.capacity:
                    self._move_bins(s, idx, g)
                else:
                    grown = caps.copy()
                    grown[idx] = g
                    self._relayout(s, demand, grown)
            rc = L.lib().rk_retry_append(self.h, _stream())
            if rc != L.RK_EOVERFLOW:  # overflow again: the next pass gives those bins room
                L.check(rc)
        else:
            if self.pending_overflow():
                raise RuntimeError("runko_amd: particle bins still overflowing after four relayouts")
        self.settle_seconds += time.perf_counter() - t0
        # the mailboxes keep their size (leavers past it spill into the shared
        # record sets); replacing them here would orphan pending records

    def _move_bins(self, s, idx, g):
        """Bins idx of species s to new places in the pool's arena, with
        capacities g; the other bins stay.  Same buffers: captured graphs
        stay valid."""
        pool = self.pools[s]
        new = pool.used + np.concatenate([[0], np.cumsum(g)[:-1]]).astype(np.int64)
        d_idx = torch.from_numpy(idx.astype(np.int32)).to(self.dev)
        d_new = torch.from_numpy(new).to(self.dev)
        d_cap = torch.from_numpy(g.astype(np.int32)).to(self.dev)
        L.check(L.lib().rk_move_bins(self.h, s, len(idx), d_idx.data_ptr(), d_new.data_ptr(), d_cap.data_ptr(),
                                     _stream()))
        torch.cuda.synchronize()
        pool.used += int(g.sum())
        pool.start_h[idx] = new
        pool.cap_h[idx] = g
        self.repacks += 1
        self.bin_moves += 1

    def _relayout(self, s, demand, grown):
        """Relays out species s in place with every bin that overflowed grown
        (plan_relayout).  A larger pool replaces it when the relayout cannot
        hold the demand, or would leave less than ARENA_LOW of the pool spare
        (relayouts would then follow one another every few laps) and the
        device has room for one."""
        pool = self.pools[s]
        caps = plan_relayout(demand, pool.cap_h, grown, pool.capacity)
        if caps is None:
            self._grow_pool(s, demand)
            return
        if self.grow_pools and pool.capacity - int(caps.sum()) < ARENA_LOW * pool.capacity:
            plan = self._plan_grow(demand)
            if plan is not None:
                self._repack(s, *plan)
                return
        free = free_device_bytes(self.dev)
        stage_bytes = int(min(getattr(self, "stage_bytes", STAGE_MAX), max(free // 2, 1 << 20)))
        stage = torch.empty((stage_bytes,), dtype=torch.uint8, device=self.dev)
        caps32 = np.ascontiguousarray(caps.astype(np.uint32))
        starts = np.zeros(self.nbricks, dtype=np.int64)
        L.check(L.lib().rk_relayout(self.h, s, caps32.ctypes.data, stage.data_ptr(), stage_bytes, starts.ctypes.data,
                                    _stream()))
        del stage
        pool.start_h, pool.cap_h = starts, caps.astype(np.int64)
        pool.used = int(caps.sum())
        self.repacks += 1
        self.relayouts += 1

    def _plan_grow(self, demand):
        """(bin capacities, arena) of a new pool for `demand`, or None when the
        device cannot hold it beside the old one and the reserve."""
        caps = bin_caps(demand)
        used = int(caps.sum())
        _, total = torch.cuda.mem_get_info(self.dev)
        free = free_device_bytes(self.dev)
        reserve = max(RESERVE_MIN, int(RESERVE_FRAC * total))
        if free <= reserve + used * SLOT_BYTES:
            return None
        return caps, plan_arena([used], free, reserve)[0]

    def _grow_pool(self, s, demand):
        """A larger pool for species s: only when the device can hold it
        beside the old one and the reserve."""
        plan = self._plan_grow(demand)
        if plan is None:
            caps = bin_caps(demand)
            free = free_device_bytes(self.dev)
            raise RuntimeError(f"runko_amd: species {s} holds {int(demand.sum())} particles, more than its pool's "
                               f"{self.pools[s].capacity} slots, and the device has {free / 2**30:.1f} GiB free "
                               f"(a larger pool needs {int(caps.sum()) * SLOT_BYTES / 2**30:.1f} GiB beside the "
                               f"reserve)")
        self._repack(s, *plan)

    def _repack(self, s, newcaps, arena=None):
        newp = _Pool(self.nbricks, newcaps, self.dev, arena)
        L.check(L.lib().rk_repack(self.h, s, newp.capacity, L.p3(newp.pos), L.p3(newp.vel), newp.id.data_ptr(),
                                  newp.start.data_ptr(), newp.cap.data_ptr(), _stream()))
        torch.cuda.synchronize()
        self.pools[s] = newp
        self.layout_version += 1
        self.repacks += 1
        self.new_pools += 1

    def presize(self, s, demand):
        """Lay out species s with bins sized for `demand` particles per bin
        (only while the species is empty)."""
        self.presize_all({s: demand})

    def presize_all(self, demands, reserve_bytes=None, slack=SLACK, pad=PAD, arena=None):
        """Lay out the species of `demands` ({species: per-bin demand}, all
        empty) together: bins of demand * slack + pad, and spare arenas sized
        from the free device memory (plan_arena) so that the pools never need
        a second allocation while the arenas and in-place relayouts can hold
        the particles.  arena: spare slots per species instead of the plan."""
        for s in demands:
            if int(self.count[s].sum().item()) != 0:
                raise RuntimeError("presize on a non-empty species")
        caps = {s: bin_caps(d, slack, pad) for s, d in demands.items()}
        for s in demands:  # release the old (empty) pools first
            self.pools[s] = _Pool(self.nbricks, np.zeros(self.nbricks, dtype=np.int64), self.dev)
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        free, total = torch.cuda.mem_get_info(self.dev)
        reserve = max(RESERVE_MIN, int(RESERVE_FRAC * total)) if reserve_bytes is None else int(reserve_bytes)
        keys = list(caps)
        arenas = plan_arena([int(caps[s].sum()) for s in keys], free, reserve)
        if arena is not None:
            arenas = [int(arena)] * len(keys)
        for s, arena in zip(keys, arenas):
            self.pools[s] = _Pool(self.nbricks, caps[s], self.dev, int(arena))
            self._bind_species(s)

    def memory_report(self):
        """Pool occupancy and overflow statistics (host numbers)."""
        out = dict(repacks=self.repacks, bin_moves=self.bin_moves, relayouts=self.relayouts, new_pools=self.new_pools,
                   settle_seconds=self.settle_seconds, species=[])
        for s, p in enumerate(self.pools):
            out["species"].append(dict(particles=int(self.count[s].sum().item()), bins_used=p.used,
                                       capacity=p.capacity, gib=p.nbytes() / 2**30))
        return out

    def bricks_of(self, t, pos):
        """Brick index (world bin) of positions pos [3, n] (float32 tensor) in tile t."""
        c = self.tile_coords(t)
        lb = []
        for a in range(3):
            origo = np.float32((self.toff[a] + c[a]) * self.n[a]) - np.float32(3.0)
            cell = torch.floor(pos[a] - origo).to(torch.int64) - HALO
            cell = cell.clamp(0, self.n[a] - 1)
            lb.append(cell // self.brick[a])
        nbd = [self.n[a] // self.brick[a] for a in range(3)]
        return t * self.bpt + (lb[0] * nbd[1] + lb[1]) * nbd[2] + lb[2]

    def inject(self, t, s, pos, vel, ids):
        """pos, vel: float32 device tensors [3, n]; ids: int64 device tensor [n]."""
        n = int(ids.numel())
        if n == 0:
            return
        self.ensure_migration(n)
        pos = pos.contiguous()
        vel = vel.contiguous()
        ids = ids.contiguous()
        rc = L.lib().rk_inject(self.h, int(t), int(s), n, L.p3(pos), L.p3(vel), ids.data_ptr(), _stream())
        if rc == L.RK_EOVERFLOW:
            self.settle_overflow()
        else:
            L.check(rc)

    # ------------------------------------------------------------------ readback
    def tile_count(self, t, s):
        n = ctypes.c_int64(0)
        L.check(L.lib().rk_tile_particle_count(self.h, int(t), int(s), ctypes.byref(n), _stream()))
        return n.value

    def tile_particles(self, t, s):
        n = self.tile_count(t, s)
        pos = torch.empty((3, max(n, 1)), dtype=torch.float32, device=self.dev)
        vel = torch.empty((3, max(n, 1)), dtype=torch.float32, device=self.dev)
        ids = torch.empty((max(n, 1),), dtype=torch.int64, device=self.dev)
        got = ctypes.c_int64(0)
        L.check(L.lib().rk_tile_particles(self.h, int(t), int(s), L.p3(pos), L.p3(vel), ids.data_ptr(), max(n, 1),
                                          ctypes.byref(got), _stream()))
        return pos[:, :n], vel[:, :n], ids[:n]

    def outgoing(self):
        cap = self._mig_cap[0]
        pos = torch.empty((3, cap), dtype=torch.float32, device=self.dev)
        vel = torch.empty((3, cap), dtype=torch.float32, device=self.dev)
        ids = torch.empty((cap,), dtype=torch.int64, device=self.dev)
        src = np.empty(cap, dtype=np.int32)
        dirs = np.empty(cap, dtype=np.int32)
        sp = np.empty(cap, dtype=np.int32)
        n = ctypes.c_int64(0)
        L.check(L.lib().rk_outgoing(self.h, L.p3(pos), L.p3(vel), ids.data_ptr(), src.ctypes.data,
                                    dirs.ctypes.data, sp.ctypes.data, cap, ctypes.byref(n), _stream()))
        k = n.value
        return (pos[:, :k].cpu().numpy(), vel[:, :k].cpu().numpy(), ids[:k].cpu().numpy().view(np.uint64),
                src[:k].copy(), dirs[:k].copy(), sp[:k].copy())

    def energies(self, tile=-1):
        out = np.zeros(2 + self.nsp, dtype=np.float64)
        L.check(L.lib().rk_energies(self.h, int(tile), out.ctypes.data, _stream()))
        return out

    def gauss_residual(self, filter_passes):
        out = np.zeros(3, dtype=np.float64)
        L.check(L.lib().rk_gauss_residual(self.h, int(filter_passes), out.ctypes.data, _stream()))
        return out

    def error_counters(self):
        out = np.zeros(2, dtype=np.int64)
        L.check(L.lib().rk_error_counters(self.h, out.ctypes.data, _stream()))
        return out

    def total_particles(self):
        return sum(int(c.sum().item()) for c in self.count)
