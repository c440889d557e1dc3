#!/usr/bin/env python
"""Benchmark of the PIC time step on one MI355X.

Headline (`value`): particle-steps/s of the full PIC step in the batched form
(one launch per method over every tile, fused particle phase) on the 3D
decaying-turbulence workload: 512^3 cells in 32^3-cell tiles, 16
particles/cell/species (4.29e9 particles), theta = 0.3, sigma = 10, guide
field + 8x8 sinusoidal dB modes, FP32.  Inputs are generated on the GPU and
resident in HBM before the timed region.  Each step = half-B push, B ghost
sync, fused interpolation + Boris + migration + zigzag deposit, J fold and
sync, 3 binomial passes, half-B push, E push with E -= J, E sync.

Also reported: the plain per-tile form (one call per tile per method through
the runko-compatible Simulation) vs the batched form on the 64^3 workload,
the roofline of the dominant kernel from HIP events around its launches, the
differentiation workload's adjoint/forward time (96^3 pair plasma, 16
ppc/species, 200 steps, checkpoints every 20), and the oracle (FP32
C++/OpenMP restatement) timed on the host cores.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s/GPU full PIC step, batched vs plain; adjoint/forward time"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(args):
    """HBM bytes per k_fused launch from the committed rocprofv3 summary of
    this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    if (args.cells, args.tile, args.ppc) != (512, 32, 16):
        return None
    path = os.path.join(ROOT, "profiles", "k_fused_turbulence512.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch k_fused ="):
                return float(line.split("=")[1])
    except OSError:
        return None
    return None


def release_device_memory():
    """Return the cached blocks of a finished workload to the device, so the
    next workload (and the library's own allocations) find them free."""
    import gc

    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def run_turbulence(args):
    import torch

    from runko_amd import _lib
    from runko_amd.init import turbulence_world
    from runko_amd.stepper import BatchedLap

    N, tile, ppc = args.cells, args.tile, args.ppc
    tuned = None
    brick = None
    if not args.no_autotune:
        # brick shape from the autotuner (fused kernel timed per shape on a
        # 2x2x2-tile world of the same tile, ppc and theta)
        from runko_amd.autotune import tune_shape
        t0 = time.time()
        brick, wg, table = tune_shape(tile, ppc=ppc, theta=0.3)
        tuned = dict(tile=tile, ppc=ppc, theta=0.3, best=list(brick), workgroup=wg, seconds=time.time() - t0,
                     ns_per_particle={f"{','.join(map(str, b))}/{t}": v
                                      for (b, t), v in sorted(table.items(), key=lambda kv: kv[1])})
        log(f"[bench] autotune: brick {brick}, {wg} threads, of {len(table)} shapes in {tuned['seconds']:.1f}s")
    t0 = time.time()
    w, info = turbulence_world(N=N, tile=tile, ppc=ppc, theta=0.3, sigma=10.0, cfl=0.45, c_omp=10.0, n_modes=8,
                               dB_over_B0=1.0, seed=1, mig_fraction=args.mig_fraction,
                               slot_order=args.slot_order, brick=brick)
    if tuned:
        w.set_fused_workgroup(tuned["workgroup"])
    npart = w.total_particles()
    ncell = N ** 3
    log(f"[bench] turbulence {N}^3, tiles {tile}^3, {npart} particles, brick {w.brick}, setup {time.time() - t0:.1f}s")
    mem0 = w.memory_report()
    log("[bench] pools: " + ", ".join(f"species {s}: {p['particles']} particles, {p['bins_used']} slots in bins, "
                                      f"{p['capacity']} capacity ({p['gib']:.1f} GiB)"
                                      for s, p in enumerate(mem0["species"])))
    lap = BatchedLap(w)
    for i in range(args.warmup):
        lap()
        lap.settle()
        log(f"[bench] warmup step {i} done (moves {w.bin_moves}, relayouts {w.relayouts})")
    launch = None
    if not args.no_autotune:
        # launch form of the fused step, timed on this world (one lap each)
        from runko_amd.autotune import tune_fused_launch
        best, times = tune_fused_launch(w, lap)
        launch = dict(chosen=best, k_fused_ms=times)
        log(f"[bench] fused launch: {best} ({', '.join(f'{k} {v:.2f} ms' for k, v in times.items())})")
    graph = None
    if not args.no_graph:
        # the same lap replayed from captured HIP graphs (one launch per lap,
        # two graphs alternating since three filter passes swap J and its
        # target), with the synchronous overflow check after every lap: the
        # headline form
        from runko_amd.stepper import GraphLap
        try:
            gl = GraphLap(w)
        except RuntimeError as e:  # capture refused: report the eager form only
            log(f"[bench] graph capture failed ({e}); headline from the eager laps")
            gl = None
        if gl is not None:
            for _ in range(2):
                gl()
                gl.settle()
            torch.cuda.synchronize()
            mem_g0 = w.memory_report()
            tg = time.perf_counter()
            for i in range(args.steps):
                gl()
                gl.settle()
            torch.cuda.synchronize()
            gwall = time.perf_counter() - tg
            mem_g1 = w.memory_report()
            graph = dict(steps=args.steps, wall=gwall, ms_per_step=1e3 * gwall / args.steps,
                         particle_steps_per_s=npart * args.steps / gwall, recaptures=gl.recaptures,
                         bin_moves=mem_g1["bin_moves"] - mem_g0["bin_moves"],
                         relayouts=mem_g1["relayouts"] - mem_g0["relayouts"])
            log(f"[bench] graph laps: {graph['ms_per_step']:.2f} ms/step ({graph})")
            del gl
    torch.cuda.synchronize()
    mem_t0 = w.memory_report()
    _lib.check(_lib.lib().rk_profile_enable(w.h, 1))
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    ev0.record()
    for i in range(args.steps):
        lap()
        lap.settle()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    _lib.check(_lib.lib().rk_profile_enable(w.h, 0))
    mem = w.memory_report()
    mem["timed"] = dict(repacks=mem["repacks"] - mem_t0["repacks"], bin_moves=mem["bin_moves"] - mem_t0["bin_moves"],
                        relayouts=mem["relayouts"] - mem_t0["relayouts"],
                        settle_ms_per_step=1e3 * (mem["settle_seconds"] - mem_t0["settle_seconds"]) / args.steps)
    log(f"[bench] pools after the timed laps: {mem['bin_moves']} bin moves, {mem['relayouts']} relayouts, "
        f"{mem['new_pools']} new pools; in the timed laps {mem['timed']}")
    import numpy as np
    ms = np.zeros(9)
    cnt = np.zeros(9, dtype=np.int64)
    _lib.check(_lib.lib().rk_profile_read(w.h, ms.ctypes.data, cnt.ctypes.data))
    prof = {name: {"ms_total": float(ms[i]), "launches": int(cnt[i])} for i, name in enumerate(_lib.PROF_CATEGORIES)}
    assert w.total_particles() == npart, "particle count changed"
    err = w.error_counters()
    sort_effect = None
    if not args.no_sort_check:
        # the reference driver cell-sorts every bin every 5 laps
        # (projects/pic-turbulence/pic.py:198-199): time that sort once, and
        # the fused kernel on the lap right after it against the timed laps'
        # average (bins are kept compact every step, so only the in-bin order
        # changes)
        ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ea.record()
        w.call("rk_sort_particles", -1)
        eb.record()
        torch.cuda.synchronize()
        _lib.check(_lib.lib().rk_profile_enable(w.h, 1))
        lap()
        lap.settle()
        ms2, cnt2 = np.zeros(9), np.zeros(9, dtype=np.int64)
        _lib.check(_lib.lib().rk_profile_read(w.h, ms2.ctypes.data, cnt2.ctypes.data))
        _lib.check(_lib.lib().rk_profile_enable(w.h, 0))
        sort_effect = dict(sort_ms=ea.elapsed_time(eb), k_fused_ms_after_sort=float(ms2[0] / max(cnt2[0], 1)),
                           k_fused_ms_timed_laps=float(ms[0] / max(cnt[0], 1)),
                           sort_ms_per_lap_every_5=ea.elapsed_time(eb) / 5)
        log(f"[bench] in-bin cell sort: {sort_effect}")
        assert w.total_particles() == npart, "particle count changed (sort)"
    plain = None
    if args.plain_steps > 0:
        # the same world advanced in the plain per-tile form (one call per tile
        # per method, each method synchronous), as the reference driver issues it
        from runko_amd.stepper import PlainLap
        w.ensure_migration(int(0.03 * npart))
        pl = PlainLap(w)
        pl()
        pl.settle()
        log("[bench] plain warm-up step done")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(args.plain_steps):
            pl()
            pl.settle()
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t1
        assert w.total_particles() == npart, "particle count changed (plain)"
        plain = dict(steps=args.plain_steps, ms_per_step=1e3 * pwall / args.plain_steps,
                     particle_steps_per_s=npart * args.plain_steps / pwall, launches_per_step=11 * w.ntiles)
    fused_ms = ms[0] / max(cnt[0], 1)
    alg_bytes = 48.0 * npart + 36.0 * ncell  # per fused launch: pos/u r+w, E/B staging read, J write
    achieved = alg_bytes / (fused_ms * 1e-3) / 1e9
    field_ms = float(ms[5] + ms[6] + ms[7])
    return dict(
        npart=npart, ncell=ncell, steps=args.steps, wall=wall, gpu_ms=ev0.elapsed_time(ev1),
        ms_per_step=1e3 * wall / args.steps, prof=prof, fused_ms=float(fused_ms), alg_bytes=alg_bytes,
        achieved_gbs=achieved, field_ms_per_step=field_ms / args.steps,
        cell_updates_per_s=ncell * args.steps / (field_ms * 1e-3) if field_ms > 0 else None,
        error_counters=[int(v) for v in err], brick=list(w.brick), B0=info["B0"], plain=plain, pools=mem, graph=graph,
        sort_effect=sort_effect,
        autotune=dict(brick=tuned, fused_launch=launch))


def run_turbulence_long(args):
    """Configuration 2 at its stated length: the 512^3 turbulence world for
    args.long_run laps of the batched lap replayed from HIP graphs, with the
    overflow settle after every lap.  Every 50 laps (and at the end): field
    and kinetic energies (FP64 reductions), the alive particle count and the
    error counters; every 250 laps the Gauss residual (its FP64 charge
    deposit takes about 1 s at this size).  The diagnostics are timed apart
    from the laps; the run stops at the first error."""
    import numpy as np
    import torch

    from runko_amd.init import turbulence_world
    from runko_amd.stepper import GraphLap

    N, tile, ppc = args.cells, args.tile, args.ppc
    t0 = time.time()
    w, info = turbulence_world(N=N, tile=tile, ppc=ppc, theta=0.3, sigma=10.0, cfl=0.45, c_omp=10.0, n_modes=8,
                               dB_over_B0=1.0, seed=1, mig_fraction=args.mig_fraction, slot_order=args.slot_order)
    npart = w.total_particles()
    log(f"[long] turbulence {N}^3, {npart} particles, set-up {time.time() - t0:.1f}s")
    gl = GraphLap(w)
    hist = []

    diag = [0.0]

    def sample(lap, wall):
        td = time.perf_counter()
        e = w.energies()
        err = [int(v) for v in w.error_counters()]
        alive = w.total_particles()
        # total energy in code units: sum(E^2 + B^2)/2 + sum_s m_s |q_s| c^2
        # sum_p (gamma_p - 1) (projects/pic-turbulence/pic.py:52-53 macro-masses)
        total = float(e[0] + e[1] + sum(m * abs(q) * w.cfl ** 2 * k for m, q, k in zip(w.masses, w.charges, e[2:])))
        rec = dict(lap=lap, wall=wall, field_E=float(e[0]), field_B=float(e[1]), total=total,
                   kinetic=[float(v) for v in e[2:]], particles=alive, errors=err,
                   bin_moves=w.bin_moves, relayouts=w.relayouts, recaptures=gl.recaptures)
        if lap % 250 == 0 or lap == args.long_run:
            g = w.gauss_residual(3)
            rec["gauss_max_over_rms_rho"] = float(g[0] / g[1]) if g[1] > 0 else 0.0
        torch.cuda.synchronize()
        diag[0] += time.perf_counter() - td
        hist.append(rec)
        log(f"[long] lap {lap}: total energy {total:.9e}; field B {rec['field_B']:.6e}, E {rec['field_E']:.6e}; kinetic "
            f"{', '.join(f'{k:.6e}' for k in rec['kinetic'])}; particles {alive}, errors {err}, "
            f"bin moves {w.bin_moves}, relayouts {w.relayouts}, {wall:.1f}s")
        return alive == npart and err == [0, 0]

    torch.cuda.synchronize()
    ok = sample(0, 0.0)
    diag[0] = 0.0
    tl = time.perf_counter()
    laps = 0
    settle_s = w.settle_seconds
    for i in range(1, args.long_run + 1):
        gl()
        gl.settle()
        laps = i
        if i % 50 == 0 or i == args.long_run:
            torch.cuda.synchronize()
            if not sample(i, time.perf_counter() - tl):
                ok = False
                break
    torch.cuda.synchronize()
    wall = time.perf_counter() - tl - diag[0]
    out = dict(workload=f"3D decaying turbulence, {N}^3 cells in {tile}^3-cell tiles, {ppc} particles/cell/species, "
                        f"theta 0.3, sigma 10, 8x8 dB modes, batched lap from HIP graphs",
               laps=laps, ok=ok, particles=npart, wall_s=wall, ms_per_lap=1e3 * wall / max(laps, 1),
               settle_ms_per_lap=1e3 * (w.settle_seconds - settle_s) / max(laps, 1),
               diagnostics_s=diag[0], bin_moves=w.bin_moves, relayouts=w.relayouts,
               total_energy_change=(hist[-1]["total"] - hist[0]["total"]) / hist[0]["total"],
               total_energy_max_deviation=max(abs(r["total"] - hist[0]["total"]) for r in hist) / hist[0]["total"],
               history=hist)
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


def run_shock_long(args):
    """Configuration 3 at its stated length: the reflecting-wall shock of
    run_shock for args.long_shock laps of the fused lap replayed from HIP
    graphs, the moving injector every 50 laps (between replays) and the
    overflow settle after every lap.  Every 100 laps (and at the end): the
    particle count against the injected total, the error counters and the
    energies, timed apart from the laps; the run stops at the first error."""
    import numpy as np
    import torch

    from runko_amd.init import WorldInjector, shock_world
    from runko_amd.stepper import GraphLap, ShockLap

    nx, ny, nz = args.shock_cells
    injloc = args.shock_injloc if args.shock_injloc > 0 else None
    t0 = time.perf_counter()
    w, walls, bcs, par = shock_world(nx=nx, ny=ny, nz=nz, tile=args.tile, ppc=8, theta=1e-4, gamma_flow=10.0,
                                     sigma=0.1, seed=3, injloc=injloc)
    n0 = w.total_particles()
    log(f"[long-shock] {nx}x{ny}x{nz}, {n0} particles, set-up {time.perf_counter() - t0:.1f}s")
    beta = (1.0 - 1.0 / 100.0) ** 0.5
    inj = None
    if injloc is not None:
        inj = WorldInjector(w, injloc=injloc, beta_flow=beta, ppc=8, theta=1e-4, gamma_flow=10.0, walloc=15.0,
                            n_inj=50, Lx=float(nx))
    fused = ShockLap(w, walls, bcs)
    fused()
    fused.settle()
    lap = GraphLap(w, lap=fused)
    hist, diag = [], [0.0]

    def sample(i, wall):
        td = time.perf_counter()
        torch.cuda.synchronize()
        e = w.energies()
        err = [int(v) for v in w.error_counters()]
        alive = w.total_particles()
        injected = inj.injected if inj else 0
        rec = dict(lap=i, wall=wall, particles=alive, injected=injected, errors=err, field_E=float(e[0]),
                   field_B=float(e[1]), kinetic=[float(v) for v in e[2:]], bin_moves=w.bin_moves,
                   relayouts=w.relayouts, new_pools=w.new_pools, recaptures=lap.recaptures,
                   front=float(inj.injloc) if inj else None)
        hist.append(rec)
        diag[0] += time.perf_counter() - td
        log(f"[long-shock] lap {i}: particles {alive} (+{injected} injected), errors {err}, "
            f"bin moves {w.bin_moves}, relayouts {w.relayouts}, new pools {w.new_pools}, "
            f"recaptures {lap.recaptures}, {wall:.1f}s")
        return alive == n0 + injected and err == [0, 0]

    ok = sample(0, 0.0)
    diag[0] = 0.0
    settle0 = w.settle_seconds
    tl = time.perf_counter()
    laps = 0
    for i in range(1, args.long_shock + 1):
        lap()
        lap.settle()
        if inj is not None:
            inj.inject(i)
        laps = i
        if i % 100 == 0 or i == args.long_shock:
            if not sample(i, time.perf_counter() - tl):
                ok = False
                break
    torch.cuda.synchronize()
    wall = time.perf_counter() - tl - diag[0]
    nm = np.mean([r["particles"] for r in hist])
    out = dict(workload=f"reflecting-wall shock, {nx}x{ny}x{nz} cells in {args.tile}^3 tiles, fused lap from HIP "
                        f"graphs, moving injector every 50 laps", laps=laps, ok=ok, particles_start=n0,
               particles_end=hist[-1]["particles"], wall_s=wall, ms_per_lap=1e3 * wall / max(laps, 1),
               particle_steps_per_s=nm * laps / wall, settle_ms_per_lap=1e3 * (w.settle_seconds - settle0) / max(laps, 1),
               recaptures=lap.recaptures, diagnostics_s=diag[0], history=hist)
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


def config0_inputs(seed=42):
    """64^3 cells, 4^3 tiles of 16^3, 8 ppc/species, theta 0.3, sigma 1, E = 0,
    B = B0 z (projects/pic-turbulence/pic.py conventions)."""
    from runko_amd.workloads import pair_plasma_inputs
    return pair_plasma_inputs(seed=seed)


def run_config0_gpu(inputs, laps, execution):
    import torch

    from runko_amd.workloads import build_simulation as build_gpu
    from runko_amd.workloads import turbulence_lap as canonical_lap

    release_device_memory()
    nt, n, Q, M, tiles = inputs
    graph = execution == "graph"
    conf, grid, sim = build_gpu(nt, n, Q, M, tiles, execution="batched" if graph else execution)
    npart = grid.world.total_particles()
    step = canonical_lap(sim)
    sim.for_one_lap(step)  # warm-up (includes the lap-0 sort)
    if graph:
        # the same lap replayed from captured HIP graphs (runko_amd.stepper.GraphLap)
        from runko_amd.stepper import GraphLap
        gl = GraphLap(grid.world)
        for _ in range(2):
            gl()
            gl.settle()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(laps):
        if graph:
            gl()
            gl.settle()
        else:
            sim.for_one_lap(step)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dict(particle_steps_per_s=npart * laps / dt, ms_per_step=1e3 * dt / laps, laps=laps, npart=npart)


def run_pair_plasma(inputs=None):
    """The 64^3 pair plasma through the Simulation API: batched, replayed from
    HIP graphs, and plain per tile."""
    if inputs is None:
        inputs = config0_inputs()
    b = run_config0_gpu(inputs, laps=20, execution="batched")
    gph = run_config0_gpu(inputs, laps=100, execution="graph")
    p = run_config0_gpu(inputs, laps=5, execution="plain")
    step_bytes = 48.0 * b["npart"] + 228.0 * 64 ** 3
    log(f"[bench] 64^3: graph {gph['particle_steps_per_s']:.3e}, batched {b['particle_steps_per_s']:.3e}, "
        f"plain {p['particle_steps_per_s']:.3e} p-s/s")
    return dict(
        workload="3D periodic pair plasma, 64^3 cells in 4^3 tiles of 16^3, 8 ppc/species, theta 0.3, "
                 "sigma 1 (Simulation API, canonical lap)",
        batched=b, graph=gph, plain_per_tile=p,
        graph_vs_plain=gph["particle_steps_per_s"] / p["particle_steps_per_s"],
        batched_vs_plain=b["particle_steps_per_s"] / p["particle_steps_per_s"],
        graph_step_roofline_frac=step_bytes / (gph["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(seconds):
    """The oracle (FP32 C++/OpenMP restatement of the reference's step, same
    lap order) timed on this host's cores on a sample of the headline
    workload (a 64^3 periodic sub-box of the turbulence run: 2^3 tiles of
    32^3 cells, 16 ppc/species, theta 0.3, sigma 10, 8x8 dB modes): built
    -O3 -march=native here when g++ can, on every thread OpenMP offers and
    on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from runko_amd.workloads import turbulence_sample_inputs

    native = oracle.use_native_build()
    nt, n, Q, M, tiles = turbulence_sample_inputs()
    w = oracle.OracleWorld(nt, n, 0.45, Q, M, prec=32)
    for idx, (F, parts) in tiles.items():
        t = w.tile_index(*idx)
        w.set_fields(t, F)
        for s, st in enumerate(parts):
            w.add_particles(t, s, st)
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    npart = sum(w.num_alive(t, s) for t in range(w.ntiles) for s in range(2))
    threads = oracle.max_threads()

    def timed(limit, max_laps, start):
        laps, t0 = 0, time.perf_counter()
        while True:
            w.lap(start + laps)
            laps += 1
            if time.perf_counter() - t0 > limit or laps >= max_laps:
                break
        return laps, time.perf_counter() - t0

    w.lap(0)  # warm-up
    laps, dt = timed(seconds, 200, 1)
    oracle.set_threads(1)
    laps1, dt1 = timed(max(seconds / 3, 1.0), 20, 1 + laps)
    oracle.set_threads(threads)
    build = "-O3 -march=native" if native else "-O3 -march=x86-64-v3"
    return dict(value=npart * laps / dt, unit="particle-steps/s", cores=threads, kind="port",
                one_thread=npart * laps1 / dt1, cpu_model=cpu_model(), build=build,
                sample=f"oracle FP32 C++/OpenMP restatement ({build}, -ffp-contract=off) on a 64^3 periodic sub-box of "
                       f"the turbulence workload (2^3 tiles of 32^3 cells, 16 ppc/species, theta 0.3, sigma 10, 8x8 "
                       f"dB modes), {npart} particles: {laps} full laps on {threads} threads ({dt:.1f} s), then "
                       f"{laps1} laps on 1 thread ({dt1:.1f} s); same step order as the GPU run")


# the stencil coefficients of the reference's 3D shock configuration
# (projects/pic-shock/3d_shock.ini:48-70, "StencilCylindricalDivFree_2x2")
SHOCK_INI_STENCIL = {
    "stencil_x_delta": -0.13512912586798165, "stencil_x_gamma": 0.016100341594971104,
    "stencil_x_beta_p1": 0.022350417588794424, "stencil_x_beta_p2": 0.022350417588794424,
    "stencil_x_zeta_p1": 0.012206124727660462, "stencil_x_zeta_p2": 0.012206124727660462,
    "stencil_y_delta": 0.010144292861133962, "stencil_y_gamma": 0.012206124727660462,
    "stencil_y_beta_p1": 0.022350417588794424, "stencil_y_beta_p2": -0.11902878427301054,
    "stencil_y_zeta_p1": 0.012206124727660462, "stencil_y_zeta_p2": 0.016100341594971104,
    "stencil_z_delta": 0.010144292861133962, "stencil_z_gamma": 0.012206124727660462,
    "stencil_z_beta_p1": -0.11902878427301054, "stencil_z_beta_p2": 0.022350417588794424,
    "stencil_z_zeta_p1": 0.016100341594971104, "stencil_z_zeta_p2": 0.012206124727660462,
}


def run_shock(args):
    """Reflecting-wall shock: nx x ny x nz cells in tile^3 tiles, gamma 10 cold
    flow along -x, 8 ppc/species upstream, density ramped 4x toward the wall
    (512 cells), theta 1e-4, sigma 0.1 with B perpendicular to the flow, the
    plasma reaching the moving injector's front (walls, edge conditions and
    injector of projects/pic-shock/pic.py:133-152).  The pic-shock lap (edge
    conditions, reflect, the rest of the PIC step, then the injector every 50
    laps), on the same world, back to back, in four forms: replayed from
    captured HIP graphs (fused particle step with the wall inside), the same
    eagerly, the batched form with the particle methods as separate launches,
    and the plain per-tile form (one call per tile per method, each method
    synchronous).  Bins that fill up are regrown inside the timed region
    (counted in `repacks`); the injector runs inside it too (eagerly, between
    replays)."""
    import gc

    import numpy as np
    import torch

    from runko_amd.init import WorldInjector, shock_world
    from runko_amd.stepper import GraphLap, PlainShockLap, ShockLap

    gc.collect()
    torch.cuda.empty_cache()
    nx, ny, nz = args.shock_cells
    injloc = args.shock_injloc if args.shock_injloc > 0 else None
    t0 = time.perf_counter()
    w, walls, bcs, par = shock_world(nx=nx, ny=ny, nz=nz, tile=args.tile, ppc=8, theta=1e-4, gamma_flow=10.0,
                                     sigma=0.1, seed=3, injloc=injloc)
    npart = w.total_particles()
    per_tile = np.array([w.tile_count(t, 0) + w.tile_count(t, 1) for t in range(w.ntiles)])
    occupied = per_tile[per_tile > 0]
    log(f"[bench] shock {nx}x{ny}x{nz}, {w.ntiles} tiles, {npart} particles, tile occupancy "
        f"{occupied.min()}..{occupied.max()} ({occupied.max() / occupied.min():.1f}x, "
        f"+{int((per_tile == 0).sum())} empty), setup {time.perf_counter() - t0:.1f}s")
    beta = (1.0 - 1.0 / 100.0) ** 0.5
    inj = None
    if injloc is not None:
        inj = WorldInjector(w, injloc=injloc, beta_flow=beta, ppc=8, theta=1e-4, gamma_flow=10.0, walloc=15.0,
                            n_inj=50, Lx=float(nx))
    res = dict(workload=f"reflecting-wall shock, {nx}x{ny}x{nz} cells in {args.tile}^3 tiles, gamma 10 flow in -x, "
                        f"8 ppc/species upstream, 4x density ramp over 512 cells at the wall, theta 1e-4, "
                        f"sigma 0.1 perpendicular B, conducting + upstream edge conditions, FP32"
                        + (f", plasma up to the moving injector at x = {injloc} (every 50 laps)" if inj else ""),
               particles=npart, tiles=w.ntiles, tile_particles_min=int(occupied.min()),
               tile_particles_max=int(occupied.max()), occupancy_spread=float(occupied.max() / occupied.min()),
               tiles_empty=int((per_tile == 0).sum()))
    lapno = [0]

    def after_lap():
        lapno[0] += 1
        if inj is not None:
            inj.inject(lapno[0])

    fused = ShockLap(w, walls, bcs)
    fused()
    fused.settle()
    after_lap()
    if not args.no_autotune:
        from runko_amd.autotune import tune_fused_launch
        best, times = tune_fused_launch(w, fused)
        lapno[0] += 2  # the tuner ran one lap per launch form
        res["fused_launch"] = dict(chosen=best, k_fused_ms=times)
        log(f"[bench] shock fused launch: {best} ({', '.join(f'{k} {v:.2f} ms' for k, v in times.items())})")
    runs = [("graph", lambda: GraphLap(w, lap=fused), args.shock_steps),
            ("batched_fused", lambda: fused, args.shock_steps),
            ("batched_unfused", lambda: ShockLap(w, walls, bcs, fused=False), max(1, args.shock_steps // 4))]
    if args.plain_steps > 0:
        runs.append(("plain_per_tile", lambda: PlainShockLap(w, walls, bcs), max(1, args.shock_plain_steps)))
    for name, make, steps in runs:
        lap = make()
        lap()
        lap.settle()
        after_lap()
        torch.cuda.synchronize()
        r0, s0 = w.repacks, w.settle_seconds
        e0, i0 = (inj.events, inj.injected) if inj else (0, 0)
        n_start = w.total_particles()
        t1 = time.perf_counter()
        for i in range(steps):
            lap()
            lap.settle()
            after_lap()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        events, added = (inj.events - e0, inj.injected - i0) if inj else (0, 0)
        n_end = w.total_particles()
        assert n_end == n_start + added, "particle count changed (shock)"
        nmean = 0.5 * (n_start + n_end)
        res[name] = dict(steps=steps, ms_per_step=1e3 * dt / steps, particle_steps_per_s=nmean * steps / dt,
                         particles_start=n_start, particles_end=n_end, injector_events=events,
                         injected=added, repacks=w.repacks - r0,
                         repack_ms_per_step=1e3 * (w.settle_seconds - s0) / steps)
        log(f"[bench] shock {name}: {1e3 * dt / steps:.2f} ms/step, {nmean * steps / dt:.3e} p-s/s, "
            f"{w.repacks - r0} repacks ({1e3 * (w.settle_seconds - s0) / steps:.2f} ms/step), "
            f"{events} injector events (+{added} particles)")
        if hasattr(lap, "recaptures"):
            res[name].update(recaptures=lap.recaptures, recapture_ms_per_step=1e3 * lap.recapture_seconds / steps)
            log(f"[bench] shock {name}: {lap.recaptures} graph recaptures "
                f"({1e3 * lap.recapture_seconds / steps:.2f} ms/step)")
        del lap
    if "plain_per_tile" in res:
        for name in ("graph", "batched_fused", "batched_unfused"):
            res[f"{name}_vs_plain"] = res["plain_per_tile"]["ms_per_step"] / res[name]["ms_per_step"]
    res["laps_run"] = lapno[0]
    res["error_counters"] = [int(v) for v in w.error_counters()]
    res["pools"] = w.memory_report()
    del runs, fused, w, inj
    gc.collect()
    torch.cuda.empty_cache()
    if not args.no_shock_schemes:
        # the same shock with the schemes of the reference's 3D shock
        # configuration: Faraday pusher, extended-stencil B push,
        # binomial2_unrolled filter (projects/pic-shock/3d_shock.ini:40-70)
        from runko_amd.stepper import GraphLap
        from runko_amd.tiles import stencil_from_config
        w, walls, bcs, par = shock_world(nx=nx, ny=ny, nz=nz, tile=args.tile, ppc=8, theta=1e-4, gamma_flow=10.0,
                                         sigma=0.1, seed=3, pusher="faraday",
                                         stencil=stencil_from_config(SHOCK_INI_STENCIL),
                                         current_filter="binomial2_unrolled")
        lap = ShockLap(w, walls, bcs)
        lap()
        lap.settle()
        g = GraphLap(w, lap=lap)
        g()
        g.settle()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.shock_steps):
            g()
            g.settle()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        res["graph_faraday_stencil"] = dict(steps=args.shock_steps, ms_per_step=1e3 * dt / args.shock_steps,
                                            particle_steps_per_s=w.total_particles() * args.shock_steps / dt,
                                            schemes="faraday pusher, stencil B push (3d_shock.ini coefficients), "
                                                    "binomial2_unrolled filter")
        log(f"[bench] shock graph (faraday, stencil, binomial2_unrolled): {1e3 * dt / args.shock_steps:.2f} ms/step, "
            f"{w.total_particles() * args.shock_steps / dt:.3e} p-s/s")
        del g, lap, w
        gc.collect()
        torch.cuda.empty_cache()
    return res


def run_ensemble(args):
    """Batched ensemble + gradients: n_boxes independent periodic pair plasmas
    (sides from {32, 48, 64, 96}, ppc/species from {4, 8, 16}, theta
    log-uniform in [0.05, 3], sigma log-uniform in [0.1, 10], 6 seeded dB
    modes each with learnable amplitudes), stepped as one ensemble: every
    kernel of a step covers all boxes in one launch.  Loss = sum of the boxes'
    final magnetic energies; forward rollout + adjoint with checkpoints.  The
    step kernels are also timed box by box (one launch set per box) for the
    ensemble-vs-per-box ratio."""
    import gc

    import torch

    from runko_amd.diff import DiffBox, EnsembleProblem, rollout

    gc.collect()
    torch.cuda.empty_cache()
    from runko_amd.diff import checkpoint_states, choose_checkpoints

    steps = args.ensemble_steps
    t0 = time.perf_counter()
    prob = EnsembleProblem(n_boxes=args.ensemble_boxes, seed=5)
    box = prob.box
    dev = prob.dev
    nb = box.n_boxes
    amp = torch.full((nb, prob.n_modes), 0.5, dtype=torch.float64, device=dev).requires_grad_(True)
    prob.loss(amp, 2, 1).backward()  # warm-up
    torch.cuda.synchronize()
    gc.collect()
    torch.cuda.empty_cache()
    state_bytes = sum(t.numel() * t.element_size() for t in box.empty_state())
    free = torch.cuda.mem_get_info()[0]
    # checkpoints of the rollout: flat if they fit 90 % of the free memory, else nested
    every = args.ensemble_every if args.ensemble_every > 0 else choose_checkpoints(steps, state_bytes, 0.9 * free)
    log(f"[bench] ensemble: {nb} boxes, {box.np} particles, {box.ncell} cells, setup {time.perf_counter() - t0:.1f}s; "
        f"state {state_bytes / 2**30:.2f} GiB, {free / 2**30:.0f} GiB free: checkpoints {every} "
        f"({checkpoint_states(steps, every)} states at the peak)")
    amp.grad = None
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    E, B, x, u = prob.initial_state(amp)
    out = rollout(box, E, B, x, u, steps, every)
    loss = prob.box_magnetic_energies(out[1]).sum()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = t1 - t0, t2 - t1
    log(f"[bench] ensemble: forward {fwd:.2f}s, backward {bwd:.2f}s for {steps} steps "
        f"(peak {torch.cuda.max_memory_allocated() / 2**30:.0f} GiB)")
    state = tuple(t.detach().contiguous() for t in out)
    del out, loss, E, B, x, u
    spare = box.empty_state()
    lam = tuple(torch.ones_like(t) * 1e-3 for t in state)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    k = 5
    ev[0].record()
    for _ in range(k):
        box.step(state, spare)
    ev[1].record()
    for _ in range(k):
        box.step_adjoint(state, lam, spare)
    ev[2].record()
    torch.cuda.synchronize()
    f_ms, a_ms = ev[0].elapsed_time(ev[1]) / k, ev[1].elapsed_time(ev[2]) / k
    del lam, spare
    # the same boxes one at a time (each its own launches)
    per_box_ms = 0.0
    for i, p in enumerate(prob.params):
        n = (p["side"],) * 3
        cnt = box.nps[i] // 2
        single = DiffBox(n, [p["q0"], -p["q0"]], [1.0, 1.0], [cnt, cnt], cfl=box.cfl,
                         filter_passes=box.filter_passes, dtype=box.dtype)
        st = (box.field_view(state[0], i).contiguous(), box.field_view(state[1], i).contiguous(),
              box.particle_view(state[2], i).contiguous(), box.particle_view(state[3], i).contiguous())
        sp = single.empty_state()
        single.step(st, sp)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        single.step(st, sp)
        e1.record()
        torch.cuda.synchronize()
        per_box_ms += e0.elapsed_time(e1)
        del single, st, sp
    grad = amp.grad
    res = dict(
        workload=f"batched ensemble: {nb} periodic pair plasmas, sides from {{32,48,64,96}}, ppc/species from "
                 f"{{4,8,16}}, theta log-uniform [0.05,3], sigma log-uniform [0.1,10], 6 seeded dB modes each; "
                 f"loss = sum of final magnetic energies after {steps} steps, checkpoints {every}, FP32",
        boxes=nb, particles=box.np, cells=box.ncell, steps=steps, checkpoint_every=every,
        checkpoint_states=checkpoint_states(steps, every), state_gib=state_bytes / 2**30,
        forward_s=fwd, backward_s=bwd, adjoint_over_forward=bwd / fwd,
        forward_ms_per_step=f_ms, adjoint_ms_per_step=a_ms, adjoint_step_over_forward_step=a_ms / f_ms,
        forward_particle_steps_per_s=box.np * steps / fwd,
        per_box_launches_forward_ms_per_step=per_box_ms, ensemble_vs_per_box=per_box_ms / f_ms,
        grad_norm=float(grad.norm()), error_count=box.error_count())
    del prob, box, state, amp
    gc.collect()
    torch.cuda.empty_cache()
    return res


def run_differentiation(args):
    """Differentiation workload: periodic pair plasma on a cells^3 box, ppc
    particles/cell/species, theta 0.3, initial dB from 12 seeded Fourier modes
    with |k| <= 3 (learnable amplitudes), counter-drift along x (learnable);
    loss = magnetic energy in the |k| band [4, 8] after `steps` steps; gradient
    by the adjoint kernels with checkpoints every `every` steps."""
    import gc

    import torch

    from runko_amd.diff import PairPlasmaProblem, band_magnetic_energy, rollout

    gc.collect()
    torch.cuda.empty_cache()
    n, steps, every = args.adjoint_cells, args.adjoint_steps, args.adjoint_every
    t0 = time.perf_counter()
    prob = PairPlasmaProblem(n=(n, n, n), ppc=args.adjoint_ppc, theta=0.3, seed=11)
    dev = prob.dev
    amp = torch.linspace(-1.0, 1.0, 12, dtype=torch.float64, device=dev).requires_grad_(True)
    drift = torch.tensor(0.05, dtype=torch.float64, device=dev).requires_grad_(True)
    prob.loss(amp, drift, 2, 1).backward()  # warm-up (kernel loads, allocator)
    torch.cuda.synchronize()
    log(f"[bench] differentiation {n}^3, {prob.box.np} particles, setup {time.perf_counter() - t0:.1f}s")
    amp.grad = None
    drift.grad = None
    t0 = time.perf_counter()
    E, B, x, u = prob.initial_state(amp, drift)
    out = rollout(prob.box, E, B, x, u, steps, every)
    loss = band_magnetic_energy(out[1])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = t1 - t0, t2 - t1
    log(f"[bench] differentiation: forward {fwd:.2f}s, backward {bwd:.2f}s for {steps} steps")
    # the two step kernels alone, on the final state, timed with HIP events
    state = tuple(t.detach().contiguous() for t in out)
    spare = prob.box.empty_state()
    lam = tuple(torch.ones_like(t) * 1e-3 for t in state)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    k = 10
    ev[0].record()
    for _ in range(k):
        prob.box.step(state, spare)
    ev[1].record()
    for _ in range(k):
        prob.box.step_adjoint(state, lam, spare)
    ev[2].record()
    torch.cuda.synchronize()
    f_ms, a_ms = ev[0].elapsed_time(ev[1]) / k, ev[1].elapsed_time(ev[2]) / k
    grad = torch.cat([amp.grad, drift.grad.view(1)])
    return dict(
        workload=f"differentiation: periodic pair plasma {n}^3 cells, {args.adjoint_ppc} ppc/species, theta 0.3, "
                 f"12 dB modes |k|<=3 + drift, loss = B energy in |k| in [4,8] after {steps} steps, "
                 f"checkpoints every {every}, FP32",
        particles=prob.box.np, steps=steps, checkpoint_every=every,
        forward_s=fwd, backward_s=bwd, adjoint_over_forward=bwd / fwd,
        forward_ms_per_step=f_ms, adjoint_ms_per_step=a_ms, adjoint_step_over_forward_step=a_ms / f_ms,
        forward_particle_steps_per_s=prob.box.np * steps / fwd,
        loss=float(loss.detach()), grad_norm=float(grad.norm()), error_count=prob.box.error_count())


def run_copy_probe(mib):
    """Plain device copies of a known size: the scale for rocprof's FETCH_SIZE
    and WRITE_SIZE counters (KiB per dispatch) when they are read for k_fused."""
    import torch

    n = int(mib) * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: 3 copies of {4 * n} bytes (read {4 * n / 1024:.0f} KiB, write {4 * n / 1024:.0f} KiB each)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=512)
    ap.add_argument("--tile", type=int, default=32)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--mig-fraction", type=float, default=0.01)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="time only the eager batched laps (no HIP-graph replay of the turbulence lap)")
    ap.add_argument("--only-pair", action="store_true",
                    help="run only the 64^3 pair-plasma legs (batched, graph, plain per tile)")
    ap.add_argument("--no-sort-check", action="store_true",
                    help="skip timing the in-bin cell sort and the fused kernel after it")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="steps of the plain per-tile form on the same world (0: skip)")
    ap.add_argument("--slot-order", default="interleaved", choices=["interleaved", "random"],
                    help="initial particle order inside a bin")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation workload")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="run only the differentiation workload and print its result object")
    ap.add_argument("--adjoint-cells", type=int, default=96)
    ap.add_argument("--adjoint-ppc", type=int, default=16)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true", help="brick shape by the rule, not the autotuner")
    ap.add_argument("--no-shock", action="store_true", help="skip the reflecting-wall shock workload")
    ap.add_argument("--only-shock", action="store_true",
                    help="run only the shock workload and print its result object")
    ap.add_argument("--no-shock-schemes", action="store_true",
                    help="skip the shock run with the reference configuration's schemes")
    ap.add_argument("--shock-cells", type=int, nargs=3, default=[2048, 128, 128], metavar=("NX", "NY", "NZ"))
    ap.add_argument("--shock-steps", type=int, default=50)
    ap.add_argument("--shock-injloc", type=float, default=1484.0,
                    help="initial injector front of the shock (0: the plasma fills the box, no injector)")
    ap.add_argument("--shock-plain-steps", type=int, default=3)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched ensemble workload")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="run only the ensemble workload and print its result object")
    ap.add_argument("--ensemble-boxes", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=300)
    ap.add_argument("--ensemble-every", type=int, default=0,
                    help="flat checkpoint spacing of the ensemble rollout (0: chosen from free memory, nested if "
                         "flat does not fit)")
    ap.add_argument("--long-shock", type=int, default=0, metavar="LAPS",
                    help="only the reflecting-wall shock for LAPS graph laps with the injector and diagnostics")
    ap.add_argument("--long-run", type=int, default=0, metavar="LAPS",
                    help="only the 512^3 turbulence world for LAPS graph laps with diagnostics every 50 laps")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run plain copies of MIB MiB (scale for rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU: every box here has a single MI355X, and this build has no "
                 "multi-GPU path (--gpus must be 1)")

    if args.only_adjoint:
        d = run_differentiation(args)
        log(f"[bench] adjoint/forward {d['adjoint_over_forward']:.2f} (rollout with recompute), "
            f"step kernels {d['forward_ms_per_step']:.2f} / {d['adjoint_ms_per_step']:.2f} ms")
        print(json.dumps(d), flush=True)
        return
    if args.only_shock:
        print(json.dumps(run_shock(args)), flush=True)
        return
    if args.only_pair:
        print(json.dumps(run_pair_plasma()), flush=True)
        return
    if args.only_ensemble:
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    if args.long_run:
        sys.exit(run_turbulence_long(args))
    if args.long_shock:
        sys.exit(run_shock_long(args))
    if args.copy_probe:
        run_copy_probe(args.copy_probe)
    res = run_turbulence(args)
    log(f"[bench] batched turbulence: {res['ms_per_step']:.2f} ms/step, fused kernel {res['fused_ms']:.2f} ms, "
        f"{res['achieved_gbs']:.0f} GB/s algorithmic")
    release_device_memory()
    # headline: the graph-replayed batched lap when it ran, else the eager one
    head = res["graph"] if res["graph"] is not None else dict(wall=res["wall"], ms_per_step=res["ms_per_step"])
    value = res["npart"] * res["steps"] / head["wall"]

    extra = {}
    inputs = None
    if not args.no_plain:
        inputs = config0_inputs()
    if not args.no_plain:
        extra["pair_plasma_64cubed"] = run_pair_plasma(inputs)
    if not args.no_shock:
        release_device_memory()
        extra["shock"] = run_shock(args)
    if not args.no_adjoint:
        release_device_memory()
        extra["differentiation"] = run_differentiation(args)
        d = extra["differentiation"]
        log(f"[bench] adjoint/forward {d['adjoint_over_forward']:.2f} (rollout with recompute), "
            f"step kernels {d['forward_ms_per_step']:.2f} / {d['adjoint_ms_per_step']:.2f} ms")
    if not args.no_ensemble:
        release_device_memory()
        extra["ensemble"] = run_ensemble(args)
        d = extra["ensemble"]
        log(f"[bench] ensemble adjoint/forward {d['adjoint_over_forward']:.2f}, step kernels "
            f"{d['forward_ms_per_step']:.2f} / {d['adjoint_ms_per_step']:.2f} ms, "
            f"{d['ensemble_vs_per_box']:.2f}x the per-box launches")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = run_cpu_baseline(args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.3e} particle-steps/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded on-GPU initial conditions)",
        "config": {
            "workload": f"3D decaying turbulence, {args.cells}^3 cells in {args.tile}^3-cell tiles, "
                        f"{args.ppc} particles/cell/species, theta 0.3, sigma 10, 8x8 dB modes, batched form"
                        + (" replayed from HIP graphs" if res["graph"] is not None else ""),
            "particles": res["npart"],
            "cells": res["ncell"],
            "brick": res["brick"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_fused (interpolation + Boris + migration classify + zigzag deposit)",
            "achieved": res["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": res["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": profiled_traffic(args),
            "traffic_source": "profiles/k_fused_turbulence512.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes_per_launch": res["alg_bytes"],
            "avg_launch_ms": res["fused_ms"],
        },
        "step_roofline": {
            "algorithmic_bytes_per_step": 48.0 * res["npart"] + 228.0 * res["ncell"],
            "achieved": (48.0 * res["npart"] + 228.0 * res["ncell"]) / (head["ms_per_step"] * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (48.0 * res["npart"] + 228.0 * res["ncell"]) / (head["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "eager_batched": dict(ms_per_step=res["ms_per_step"], particle_steps_per_s=res["npart"] * res["steps"] / res["wall"]),
        "autotune": res["autotune"],
        "cpu_baseline": cpu,
        "fdtd_cell_updates_per_s": res["cell_updates_per_s"],
        "kernel_ms": res["prof"],
        "error_counters": res["error_counters"],
        "pools": res["pools"],
        "sort_effect": res["sort_effect"],
    }
    if res["plain"] is not None:
        line["batched_vs_plain"] = dict(workload=line["config"]["workload"].split(", batched form")[0],
                                        batched_ms_per_step=res["ms_per_step"], graph_ms_per_step=head["ms_per_step"],
                                        plain=res["plain"], ratio=res["plain"]["ms_per_step"] / res["ms_per_step"],
                                        graph_ratio=res["plain"]["ms_per_step"] / head["ms_per_step"])
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
