"""The canonical PIC lap of projects/pic-turbulence/pic.py:187-221 on one world,
without the lap-function dispatch of Simulation:

* BatchedLap: one launch per method over every tile, the particle phase (push,
  pack_outgoing, comm pic_particle, sort, deposit_current) as the fused
  particle step, E += cfl curl B and E -= J in one pass (Simulation "batched").
* PlainLap: the same lap in the per-tile form of the reference driver, one call
  per tile per method, each method synchronous over all tiles (Simulation
  "plain").
* ShockLap / PlainShockLap: the reflecting-wall shock lap of
  projects/pic-shock/pic.py:250-299 (edge conditions after every field
  update, reflect between push and pack, the wall advanced per lap; IO and
  the injector excluded), batched (one launch per method over every tile;
  the particle phase unfused, since the reflect sits inside it) and per tile.
"""

import ctypes
import time

from . import _lib as L

import numpy as np
import torch

from .tools import comm_mode
from .world import _stream


class BatchedLap:
    def __init__(self, world, filter_passes=3):
        self.w = world
        self.filter_passes = filter_passes

    def __call__(self):
        w = self.w
        w.call("rk_push_half_b", -1)
        w.comm_local(int(comm_mode.emf_B))
        w.particle_step_fused()
        w.comm_local(int(comm_mode.emf_J_exchange))
        w.comm_local(int(comm_mode.emf_J))
        # filter, comm J, filter, filter (pic.py:208-212) as one call
        w.filter_current_passes(self.filter_passes)
        w.call("rk_push_half_b", -1)
        w.comm_local(int(comm_mode.emf_B))
        w.call("rk_push_e_add_current", -1)
        w.comm_local(int(comm_mode.emf_E))

    def settle(self):
        """Synchronous check after a lap: bins that overflowed are repacked and
        their pending particles appended (none are lost)."""
        n = self.w.pending_overflow()
        if n:
            self.w.settle_overflow(n)


class PlainLap:
    def __init__(self, world, filter_passes=3):
        self.w = world
        self.filter_passes = filter_passes

    def _each(self, name):
        for t in range(self.w.ntiles):
            self.w.call(name, t)
        torch.cuda.synchronize()

    def __call__(self):
        w = self.w
        self._each("rk_push_half_b")
        w.comm_local(int(comm_mode.emf_B))
        self._each("rk_push_particles")
        self._each("rk_pack_outgoing_particles")
        w.comm_local(int(comm_mode.pic_particle))
        self._each("rk_deposit_current")
        w.comm_local(int(comm_mode.emf_J_exchange))
        w.comm_local(int(comm_mode.emf_J))
        for p in range(self.filter_passes):
            self._each("rk_filter_current")
            if p == 0 and self.filter_passes > 1:
                w.comm_local(int(comm_mode.emf_J))
        self._each("rk_push_half_b")
        w.comm_local(int(comm_mode.emf_B))
        self._each("rk_push_e")
        self._each("rk_add_current")
        w.comm_local(int(comm_mode.emf_E))
        torch.cuda.synchronize()

    def settle(self):
        n = self.w.pending_overflow()
        if n:
            self.w.settle_overflow(n)


class GraphLap:
    """BatchedLap replayed from captured HIP graphs.

    Every filter pass swaps the roles of J and the filter target, so with an
    odd number of passes consecutive laps run on swapped buffers: two graphs
    are captured (even and odd laps) and replayed alternately, starting with
    the even one.  `lap` may instead be a fused ShockLap that has run one
    eager lap (that lap allocates the wall correction buffer), with walls that
    stand still (they travel as launch arguments) or walls held in device
    memory (device_walls=True; they may move).  Capture records
    kernels only (the library allocates nothing and never synchronises inside
    these calls; kernel timing must be off).
    After every replay settle() checks the overflow counter (synchronous), as
    for BatchedLap.
    """

    def __init__(self, world, filter_passes=3, lap=None):
        self.w = world
        if lap is None:
            lap = BatchedLap(world, filter_passes)
        elif not isinstance(lap, ShockLap) or not lap.fused:
            raise ValueError("only BatchedLap and the fused ShockLap are captured (the separate particle "
                             "methods check their counters synchronously)")
        elif not lap.device_walls and any(float(wall.betawall) != 0.0 for wall in lap._walls):
            raise ValueError("a moving wall given as launch arguments changes every lap: use "
                             "ShockLap(..., device_walls=True) to replay it from a graph")
        elif lap.laps_run == 0:
            raise ValueError("run one eager lap of the ShockLap before capturing it")
        self.lap = lap
        if world.mbox is None or world.jstage is None:
            world.ensure_mailbox()
        self.recaptures, self.recapture_seconds = 0, 0.0
        self._capture()

    def _capture(self):
        torch.cuda.synchronize()
        self.graphs = []
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2 if self.lap.filter_passes % 2 else 1):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=side):
                    self.lap()
                self.graphs.append(g)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.k = 0
        self.version = self.w.layout_version

    def __call__(self):
        self.graphs[self.k].replay()
        self.k = (self.k + 1) % len(self.graphs)

    def settle(self):
        """After a repack (or a grown mailbox) the graphs hold replaced
        buffers: they are captured again from the current state.  J between
        laps is scratch (every lap rebuilds it), so recapturing after an odd
        number of replays changes nothing but which buffer the next lap's J
        starts in."""
        n = self.w.pending_overflow()
        if n:
            self.w.settle_overflow(n)
        if self.w.layout_version != self.version:
            t0 = time.perf_counter()
            del self.graphs
            self._capture()
            self.recaptures += 1
            self.recapture_seconds += time.perf_counter() - t0


class ShockLap:
    """walls: list of rk_reflector_wall; bcs: list of rk_edge_bc (all tiles).

    fused (batched form): push, reflect, pack, comm pic_particle and deposit
    run as the fused particle step with the walls inside it; otherwise as
    separate launches.  device_walls (fused form): the walls live in device
    memory and advance with a kernel, so a moving wall replays from a graph;
    `walls` then reads them back."""

    def __init__(self, world, walls, bcs, filter_passes=3, tiles=None, fused=True, device_walls=False):
        self.w = world
        self._walls = list(walls)
        self.bcs = list(bcs)
        self.filter_passes = filter_passes
        self.tiles = tiles  # None: one launch per method over every tile
        self.fused = fused and tiles is None
        self.device_walls = bool(device_walls) and self.fused
        if self.device_walls:
            world.set_walls(self._walls)
        self.laps_run = 0

    @property
    def walls(self):
        return self.w.read_walls() if self.device_walls else self._walls

    def _targets(self):
        return [-1] if self.tiles is None else range(self.w.ntiles)

    def _each(self, name):
        for t in self._targets():
            self.w.call(name, t)
        if self.tiles is not None:
            torch.cuda.synchronize()

    def _bcs(self, mode):
        for t in self._targets():
            for bc in self.bcs:
                L.check(L.lib().rk_apply_edge_bc(self.w.h, t, ctypes.byref(bc), int(mode), _stream()))
        if self.tiles is not None:
            torch.cuda.synchronize()

    def _reflect(self):
        arr = (L.ReflectorWall * len(self._walls))(*self._walls)
        for t in self._targets():
            L.check(L.lib().rk_reflect_particles(self.w.h, t, arr, len(self._walls), _stream()))
        if self.tiles is not None:
            torch.cuda.synchronize()

    def __call__(self):
        w = self.w
        self._each("rk_push_half_b")
        self._bcs(comm_mode.emf_B)
        w.comm_local(int(comm_mode.emf_B))
        if self.device_walls:
            w.particle_step_fused_device()
        elif self.fused:
            w.particle_step_fused(self._walls)
        else:
            self._each("rk_push_particles")
            self._reflect()
            self._each("rk_pack_outgoing_particles")
            w.comm_local(int(comm_mode.pic_particle))
            self._each("rk_deposit_current")
        w.comm_local(int(comm_mode.emf_J_exchange))
        w.comm_local(int(comm_mode.emf_J))
        self._bcs(comm_mode.emf_J)
        if self.tiles is None and self.filter_passes == 3 and w.periodic and w.current_filter == "binomial2":
            # three passes with no ghost sync between them: one launch reading
            # J and writing the result once (the same interior bit for bit)
            w.filter_current_passes(3)
        else:
            for p in range(self.filter_passes):
                if p > 0 and p % 3 == 0:
                    w.comm_local(int(comm_mode.emf_J))
                self._each("rk_filter_current")
        self._bcs(comm_mode.emf_J)
        self._each("rk_push_half_b")
        self._bcs(comm_mode.emf_B)
        w.comm_local(int(comm_mode.emf_B))
        self._each("rk_push_e")
        self._bcs(comm_mode.emf_E)
        self._each("rk_add_current")
        self._bcs(comm_mode.emf_E)
        w.comm_local(int(comm_mode.emf_E))
        if self.device_walls:
            w.advance_walls()
        else:
            cfl = np.float32(w.cfl)
            for wall in self._walls:
                wall.walloc = float(np.float32(wall.walloc) + np.float32(wall.betawall) * cfl)
        self.laps_run += 1
        if self.tiles is not None:
            torch.cuda.synchronize()

    def settle(self):
        n = self.w.pending_overflow()
        if n:
            self.w.settle_overflow(n)


class PlainShockLap(ShockLap):
    def __init__(self, world, walls, bcs, filter_passes=3):
        super().__init__(world, walls, bcs, filter_passes, tiles="all", fused=False)
