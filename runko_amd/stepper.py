"""The canonical PIC lap of projects/pic-turbulence/pic.py:187-221 on one world,
without the lap-function dispatch of Simulation:

* BatchedLap: one launch per method over every tile, the particle phase (push,
  pack_outgoing, comm pic_particle, sort, deposit_current) as the fused
  particle step, E += cfl curl B and E -= J in one pass (Simulation "batched").
* PlainLap: the same lap in the per-tile form of the reference driver, one call
  per tile per method, each method synchronous over all tiles (Simulation
  "plain").
"""

import torch

from .tools import comm_mode


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
        for p in range(self.filter_passes):
            w.call("rk_filter_current", -1)
            if p == 0 and self.filter_passes > 1:
                w.comm_local(int(comm_mode.emf_J))
        w.call("rk_push_half_b", -1)
        w.comm_local(int(comm_mode.emf_B))
        w.call("rk_push_e_add_current", -1)
        w.comm_local(int(comm_mode.emf_E))

    def settle(self):
        """Synchronous check after a lap: bins that overflowed are repacked and
        their pending particles appended (none are lost)."""
        if self.w.pending_overflow():
            self.w.settle_overflow()


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
        if self.w.pending_overflow():
            self.w.settle_overflow()


class GraphLap:
    """BatchedLap replayed from captured HIP graphs.

    Every filter pass swaps the roles of J and the filter target, so with an
    odd number of passes consecutive laps run on swapped buffers: two graphs
    are captured (even and odd laps) and replayed alternately, starting with
    the even one.  Capture records kernels only (the library allocates nothing
    and never synchronises inside these calls; kernel timing must be off).
    After every replay settle() checks the overflow counter (synchronous), as
    for BatchedLap.
    """

    def __init__(self, world, filter_passes=3):
        self.w = world
        self.lap = BatchedLap(world, filter_passes)
        if world.mbox is None or world.jstage is None:
            world.ensure_mailbox()
        torch.cuda.synchronize()
        self.graphs = []
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2 if filter_passes % 2 else 1):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=side):
                    self.lap()
                self.graphs.append(g)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.k = 0

    def __call__(self):
        self.graphs[self.k].replay()
        self.k = (self.k + 1) % len(self.graphs)

    def settle(self):
        if self.w.pending_overflow():
            self.w.settle_overflow()
