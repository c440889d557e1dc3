"""Simulation: runs lap functions over the tile grid (runko/simulation.py:31-518).

Lap functions receive a MethodWrapper whose calls are dispatched by name:
  prtcl_<m>  -> per-tile particle method (push, sort, pack_outgoing, deposit_current)
  grid_<m>   -> per-tile emf method (push_half_b, push_e, add_current, filter_current)
  comm_local / comm_external (comm_mode arguments only, TypeError otherwise)
  io_<m>     -> average energies (FP64 reductions on the GPU)

Two execution modes share that interface:
  plain    -- the drop-in form: one call per tile per method, as the
              reference's Python loop over local tiles does;
  batched  -- one launch per method over every tile, and the particle phase
              push, pack_outgoing, comm(pic_particle), [sort], deposit_current
              is recognised and run as one fused kernel (rk_particle_step_fused).
              Bins are re-established every step there, which is what the
              periodic sort provides for memory locality, so a prtcl_sort inside
              that phase needs no pass of its own.  The filter sequence
              filter_current, comm_local(emf_J), filter_current, filter_current
              (or three filter_current in a row) runs as one three-pass
              launch (rk_filter_current_passes) on a periodic binomial2 world.
"""

import ctypes
import logging
import os
import pathlib
import time

import numpy as np
import torch

from . import _lib as L
from .method_wrapper import MethodWrapper
from .tiles import _walls_array
from .world import _stream
from .timer import Timer, timer_statistics
from .tools import comm_mode

_PRTCL = {"prtcl_push": "push_particles", "prtcl_sort": "sort_particles",
          "prtcl_pack_outgoing": "pack_outgoing_particles", "prtcl_deposit_current": "deposit_current",
          "prtcl_reflect_particles": "reflect_particles",
          "prtcl_advance_reflector_walls": "advance_reflector_walls"}
_RK_TILE = {"push_particles": "rk_push_particles", "sort_particles": "rk_sort_particles",
            "pack_outgoing_particles": "rk_pack_outgoing_particles", "deposit_current": "rk_deposit_current",
            "push_half_b": "rk_push_half_b", "push_e": "rk_push_e", "add_current": "rk_add_current",
            "filter_current": "rk_filter_current"}
_NEXT = {"push": ("pack", "reflect"), "reflect": ("pack",), "pack": ("comm",), "comm": ("sort", "deposit"),
         "sort": ("deposit",)}


class Simulation:
    _im_not_user = 42

    def __init__(self, tile_grid, prevent_user_init, **kwargs):
        if prevent_user_init != Simulation._im_not_user:
            raise RuntimeError("Don't instantiate this class directly!")
        self._tile_grid = tile_grid
        self._world = tile_grid.world
        self._lap = 0
        self._last_lap = kwargs["Nt"]
        self._io_config = dict(kwargs["io_config"])
        self.verbose_laps = kwargs.get("verbose", True)
        self.mode = os.environ.get("RUNKO_AMD_EXECUTION", kwargs.get("mode", "plain"))
        if self.mode not in ("plain", "batched"):
            raise RuntimeError(f"unknown execution mode {self.mode}")
        self._lap_timers = []
        self._lap_wall_times = []
        self._prev_elapsed_wall_time = 0
        self._total_interval_time = 0
        self._pending = []
        self._pending_walls = None
        self._pending_filter = ""  # batched filter sequence so far: "f" filter_current, "c" comm_local(emf_J)
        self._warned = set()
        self._emf_writer = None
        self._prtcl_writers = {}
        self._spectra_writer = None
        self._logger = logging.getLogger("runko_amd.Simulation")
        out = self._io_config["outdir"]
        self._io_config["kinetic_energy_path"] = f"{out}/average_kinetic_energy.txt"
        self._io_config["average_B_energy_density_path"] = f"{out}/average_B_energy_density.txt"
        self._io_config["average_E_energy_density_path"] = f"{out}/average_E_energy_density.txt"
        self._files_reset = False

    # ------------------------------------------------------------------ tiles
    def local_tiles(self):
        return iter(self._tile_grid.tiles())

    def virtual_tiles(self):
        return iter(())

    @property
    def lap(self):
        return self._lap

    @property
    def world(self):
        return self._world

    # ------------------------------------------------------------------ io
    def _reset_files(self):
        if self._files_reset:
            return
        pathlib.Path(self._io_config["outdir"]).mkdir(parents=True, exist_ok=True)
        for k in ("kinetic_energy_path", "average_B_energy_density_path", "average_E_energy_density_path"):
            pathlib.Path(self._io_config[k]).unlink(missing_ok=True)
        self._files_reset = True

    def _io(self, what):
        w = self._world
        if what in ("average_kinetic_energy", "average_B_energy_density", "average_E_energy_density"):
            self._reset_files()
            e = w.energies(-1)
            if what == "average_kinetic_energy":
                vals = []
                for s in range(w.nsp):
                    n = int(w.count[s].sum().item())
                    vals.append(0.0 if n == 0 else e[2 + s] / n)
                line = f"{self.lap} " + "".join(f"{v:g} " for v in vals) + "\n"
                path = self._io_config["kinetic_energy_path"]
            else:
                cells = float(np.prod(w.N))
                v = (e[1] if what == "average_B_energy_density" else e[0]) / cells
                line = f"{self.lap} {v:g}\n"
                path = self._io_config["average_B_energy_density_path" if "B" in what else "average_E_energy_density_path"]
            with open(path, "a") as f:
                f.write(line)
        elif what in ("emf_snapshot", "prtcl_snapshot", "spectra_snapshot"):
            # runko/simulation.py:96-146,257-276: writers built on first use
            from .snapshots import MpiioFieldsWriter, MpiioParticlesWriter, MpiioSpectraWriter
            io, g = self._io_config, self._tile_grid
            dims = (g._Nx, g._NxMesh, g._Ny, g._NyMesh, g._Nz, g._NzMesh)
            pathlib.Path(io["outdir"]).mkdir(parents=True, exist_ok=True)
            if what == "emf_snapshot":
                if self._emf_writer is None:
                    self._emf_writer = MpiioFieldsWriter(io["outdir"], *dims, io["stride"], io.get("nspecies", 2))
                self._emf_writer.write(g, self.lap)
            elif what == "prtcl_snapshot":
                if not self._prtcl_writers:
                    self._prtcl_writers = {sp: MpiioParticlesWriter(io["outdir"], io.get("n_prtcls", 0), sp)
                                           for sp in range(io.get("nspecies", 2))}
                for wr in self._prtcl_writers.values():
                    wr.write(g, self.lap)
            else:
                if self._spectra_writer is None:
                    self._spectra_writer = MpiioSpectraWriter(
                        io["outdir"], *dims, io.get("spectra_stride", io["stride"]), io["spectra_nbins"],
                        io["spectra_umin"], io["spectra_umax"], io.get("nspecies", 2))
                self._spectra_writer.write(g, self.lap)
        elif what == "ram_usage":
            if what not in self._warned:
                self._logger.warning("io_ram_usage: host memory diagnostics are not part of runko_amd; skipped")
                self._warned.add(what)
        else:
            raise AttributeError(f"io_{what} is not supported IO type.")

    # ------------------------------------------------------------------ batched particle phase
    def _run_unfused(self, op):
        w = self._world
        if op == "push":
            w.call("rk_push_particles", -1)
        elif op == "pack":
            w.call("rk_pack_outgoing_particles", -1)
        elif op == "comm":
            w.comm_local(int(comm_mode.pic_particle))
        elif op == "sort":
            w.call("rk_sort_particles", -1)
        elif op == "deposit":
            w.call("rk_deposit_current", -1)
        elif op == "reflect":
            # one launch per tile that a wall overlaps, issued by the library
            arr = self._pending_walls
            L.check(L.lib().rk_reflect_particles(w.h, -1, arr, len(arr), _stream()))

    def _flush(self):
        ops, self._pending = self._pending, []
        for op in ops:
            self._run_unfused(op)
        self._flush_filter()

    def _flush_filter(self):
        seq, self._pending_filter = self._pending_filter, ""
        w = self._world
        for head in ("fcff", "fff"):
            if seq.startswith(head) and w.periodic and w.current_filter == "binomial2":
                w.filter_current_passes(3)
                seq = seq[len(head):]
                break
        for op in seq:
            if op == "f":
                w.call("rk_filter_current", -1)
            else:
                w.comm_local(int(comm_mode.emf_J))

    def _particle_op(self, op):
        if self._pending and op in _NEXT.get(self._pending[-1], ()):
            self._pending.append(op)
            if op == "deposit":
                walls = list(self._pending_walls) if "reflect" in self._pending else None
                self._pending = []
                w = self._world
                w.particle_step_fused(walls)
                n = w.pending_overflow()
                if n:
                    w.settle_overflow(n)
            return
        self._flush()
        if op == "push":
            self._pending = ["push"]
        else:
            self._run_unfused(op)

    # ------------------------------------------------------------------ dispatch
    def _execute_lap_function(self, lap_function, disable_timing=False):
        t_begin = time.time()
        lap_timer = None if disable_timing else Timer()

        def name_of(method, kwargs):
            return kwargs.get("name", method)

        def pre(method, *vargs, **kwargs):
            lap_timer.start(name_of(method, kwargs))

        def post(method, *vargs, **kwargs):
            if self.mode == "plain":
                torch.cuda.synchronize()
            lap_timer.stop(name_of(method, kwargs))

        def action(method, *vargs):
            batched = self.mode == "batched"
            if method.startswith("prtcl_"):
                m = _PRTCL.get(method, method)
                if batched and m in _RK_TILE:
                    op = {"push_particles": "push", "pack_outgoing_particles": "pack",
                          "sort_particles": "sort", "deposit_current": "deposit"}[m]
                    self._particle_op(op)
                elif batched and m == "reflect_particles" and self._same_on_all_tiles("_walls"):
                    # joins a pending fused particle phase (push -> reflect ->
                    # pack -> comm -> deposit), else runs at once
                    walls = self._tile_grid.tiles()[0]._walls
                    if walls:
                        self._pending_walls = _walls_array(walls)
                        self._particle_op("reflect")
                else:
                    self._flush()
                    for tile in self.local_tiles():
                        getattr(tile, m)()
            elif method.startswith("grid_"):
                m = method[5:]
                if batched and m == "apply_edge_bcs" and self._same_on_all_tiles("_edge_bcs"):
                    self._flush()
                    for bc in self._tile_grid.tiles()[0]._edge_bcs:
                        L.check(L.lib().rk_apply_edge_bc(self._world.h, -1, ctypes.byref(bc._struct()),
                                                         int(vargs[0]), _stream()))
                elif batched and m == "filter_current" and not vargs:
                    if self._tile_grid.tiles()[0].current_filter is None:
                        raise RuntimeError("Trying to filter current without specifying `current_filter`!")
                    ops, self._pending = self._pending, []
                    for op in ops:
                        self._run_unfused(op)
                    self._pending_filter += "f"
                elif batched and m in _RK_TILE and not vargs:
                    self._flush()
                    self._world.call(_RK_TILE[m], -1)
                else:
                    self._flush()
                    for tile in self.local_tiles():
                        getattr(tile, m)(*vargs)
            elif method.startswith("io_"):
                self._flush()
                self._io(method[3:])
            elif method.startswith("comm_"):
                for mode in vargs:
                    if type(mode) != comm_mode:
                        raise TypeError("Communications only accept runko.comm_mode arguments.\n"
                                        f"Received {mode} of type {type(mode)}.")
                kind = method[5:]
                if kind == "external":
                    return  # one process owns every tile: nothing crosses a process boundary
                if kind != "local":
                    raise AttributeError(f"{method} is not supported communication type.")
                for mode in vargs:
                    if batched and mode == comm_mode.pic_particle and self._pending:
                        self._particle_op("comm")
                    elif batched and mode == comm_mode.emf_J and self._pending_filter:
                        self._pending_filter += "c"
                    else:
                        self._flush()
                        self._world.comm_local(int(mode.value))
            else:
                raise RuntimeError(f"{method} is not supported!")

        hooks = {} if disable_timing else dict(pre=pre, post=post)
        lap_function(MethodWrapper(action, **hooks))
        self._flush()
        t_end = time.time()
        if not disable_timing:
            self._lap_timers.append(lap_timer)
            self._lap_wall_times.append(t_end - t_begin)

    def _same_on_all_tiles(self, attr):
        """True when every tile holds the same registered walls / edge bcs."""
        key = lambda objs: [tuple(sorted(o.__dict__.items())) for o in objs]
        tiles = self._tile_grid.tiles()
        first = key(getattr(tiles[0], attr))
        return all(key(getattr(t, attr)) == first for t in tiles[1:])

    def prelude(self, lap_function):
        self._execute_lap_function(lap_function, disable_timing=True)

    def for_one_lap(self, lap_function):
        self._execute_lap_function(lap_function)
        self._lap += 1

    def for_each_lap(self, lap_function):
        while self._lap < self._last_lap:
            self.for_one_lap(lap_function)

    # ------------------------------------------------------------------ timers
    def get_time_statistics(self, n_latests=None):
        if n_latests:
            return timer_statistics(self._lap_timers[-n_latests:])
        return timer_statistics(self._lap_timers)

    def reset_timers(self):
        self._lap_timers = []
        self._lap_wall_times = []
        self._prev_elapsed_wall_time = self._total_interval_time

    def log_timer_statistics(self, level=logging.INFO):
        stats = self.get_time_statistics(self._io_config.get("laps_in_timer_statistics"))
        if not stats:
            self._logger.warning("Trying to log timer non-existing statistic.")
            return
        items = sorted(stats.items(), key=lambda kv: -kv[1].total)
        total = sum(s.total for _, s in items)
        self._total_interval_time = total
        width = max(len(k) for k, _ in items)
        msg = "Simulation execution time statistics:\n"
        msg += f"{'name':<{width}} | total [s] | % of total | average [s] | std [s] | count\n"
        for name, s in items:
            pct = 100 * s.total / total if total else 0.0
            msg += f"{name:<{width}} | {s.total:>9.1e} | {pct:>10.4} | {s.average:>11.3e} | {s.std_dev:>7.1e} | {s.count:>5}\n"
        laps = len(self._lap_wall_times)
        if laps:
            msg += f"Lap wall times: {np.sum(self._lap_wall_times):.4} s / {laps} laps = {np.mean(self._lap_wall_times):.4} s / lap\n"
        msg += f"Current lap: {self._lap} / {self._last_lap}"
        self._logger.log(level, msg)
