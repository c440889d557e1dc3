"""Named workloads built through the runko-compatible API.

* pair_plasma_inputs: the 3D periodic e-/e+ pair plasma of the turbulence
  driver's conventions (projects/pic-turbulence/pic.py:44-68,141-156):
  positions uniform in every cell from one seeded numpy generator, species 1
  on top of species 0, isotropic Juttner-Synge momenta, E = 0, B = B0 z.
* build_simulation: tiles with those inputs in a TileGrid and its Simulation,
  with the E/B halo sync as prelude (projects/pic-turbulence/pic.py:160-175).
* turbulence_lap: the lap function of projects/pic-turbulence/pic.py:187-221
  (sort every `sort_every` laps), without snapshot IO.
"""

import numpy as np

from .configuration import Configuration
from .init import plasma_parameters
from .sampling import sample_boosted_juttner_synge

CFL = 0.45


def pair_plasma_inputs(nt=(4, 4, 4), n=(16, 16, 16), ppc=8, theta=0.3, sigma=1.0, c_omp=10.0, seed=42):
    """Per tile (i, j, k): (fields [9, *n] float64, [states [k, 6] per species]).
    The defaults are the 64^3 workload: 4^3 tiles of 16^3 cells, 8 ppc per
    species (4.19e6 particles), theta 0.3, sigma 1."""
    q0, q1, B0 = plasma_parameters(CFL, c_omp, ppc, theta, sigma)
    rng = np.random.default_rng(seed)
    tiles = {}
    cells = np.stack(np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij"), -1).reshape(-1, 3)
    for i in range(nt[0]):
        for j in range(nt[1]):
            for k in range(nt[2]):
                lo = np.array([i * n[0], j * n[1], k * n[2]], dtype=np.float64)
                pos = np.repeat(cells, ppc, 0) + lo + rng.random((len(cells) * ppc, 3))
                parts = []
                for s in range(2):
                    ux, uy, uz = sample_boosted_juttner_synge(len(pos), theta, beta=0, gen=rng)
                    parts.append(np.concatenate([pos, np.stack([ux, uy, uz], 1)], 1))
                F = np.zeros((9,) + tuple(n))
                F[5] = B0
                tiles[(i, j, k)] = (F, parts)
    return tuple(nt), tuple(n), [q0, q1], [1.0, 1.0], tiles


def turbulence_sample_inputs(nt=(2, 2, 2), n=(32, 32, 32), ppc=16, theta=0.3, sigma=10.0, c_omp=10.0, n_modes=8,
                             dB_over_B0=1.0, seed=1):
    """A periodic sub-box of the decaying-turbulence workload, in the layout of
    pair_plasma_inputs: its 32^3-cell tiles, 16 ppc per species, theta 0.3,
    sigma 10 (q and B0 of the full box), Bz = B0 plus the 8x8 dBx, dBy modes
    of init.set_turbulent_field laid over the sub-box (so they stay periodic
    in it).  The defaults (2^3 tiles, 8.39e6 particles) are the CPU timing
    sample of the 512^3 workload."""
    q0, q1, B0 = plasma_parameters(CFL, c_omp, ppc, theta, sigma)
    N = [a * b for a, b in zip(nt, n)]
    rng = np.random.default_rng(seed)
    phi = 2 * np.pi * rng.random((n_modes, n_modes))
    kx, ky = 2 * np.pi / N[0], 2 * np.pi / N[1]
    X = np.arange(N[0] + 1, dtype=np.float64)[:, None]
    Y = np.arange(N[1] + 1, dtype=np.float64)[None, :]
    A = np.zeros((N[0] + 1, N[1] + 1))
    for m in range(1, n_modes + 1):
        for k in range(1, n_modes + 1):
            kd = np.sqrt((2 * np.sin(m * kx / 2)) ** 2 + (2 * np.sin(k * ky / 2)) ** 2)
            A += np.sin(m * kx * X + k * ky * Y + phi[m - 1, k - 1]) / kd
    bx = A[:-1, 1:] - A[:-1, :-1]
    by = A[:-1, :-1] - A[1:, :-1]
    scale = dB_over_B0 * B0 / np.sqrt(np.mean(bx ** 2 + by ** 2))
    bx = (scale * bx).astype(np.float32).astype(np.float64)
    by = (scale * by).astype(np.float32).astype(np.float64)
    cells = np.stack(np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij"), -1).reshape(-1, 3)
    tiles = {}
    for i in range(nt[0]):
        for j in range(nt[1]):
            for k in range(nt[2]):
                lo = np.array([i * n[0], j * n[1], k * n[2]], dtype=np.float64)
                pos = np.repeat(cells, ppc, 0) + lo + rng.random((len(cells) * ppc, 3))
                parts = []
                for s in range(2):
                    ux, uy, uz = sample_boosted_juttner_synge(len(pos), theta, beta=0, gen=rng)
                    parts.append(np.concatenate([pos, np.stack([ux, uy, uz], 1)], 1))
                F = np.zeros((9,) + tuple(n))
                sx = slice(i * n[0], (i + 1) * n[0])
                sy = slice(j * n[1], (j + 1) * n[1])
                F[3] = bx[sx, sy][:, :, None]
                F[4] = by[sx, sy][:, :, None]
                F[5] = B0
                tiles[(i, j, k)] = (F, parts)
    return tuple(nt), tuple(n), [q0, q1], [1.0, 1.0], tiles


def make_configuration(nt, n, charges, masses, filt="binomial2", cfl=CFL, execution="plain", pusher="boris",
                       stencil=None):
    c = Configuration(None)
    c.n_tiles = list(nt)
    c.n_cells_per_tile = list(n)
    c.cfl = cfl
    c.field_propagator = "fdtd2" if stencil is None else "stencil"
    for key, v in (stencil or {}).items():
        setattr(c, key, v)
    c.particle_pusher = pusher
    c.field_interpolator = "linear_1st"
    c.current_depositer = "zigzag_1st_atomic"
    c.current_filter = filt
    c.tile_partitioning = "hilbert_curve"
    c.n_laps = 10
    c.execution = execution
    for s, (q, m) in enumerate(zip(charges, masses)):
        setattr(c, f"q{s}", q)
        setattr(c, f"m{s}", m)
    return c


def build_simulation(nt, n, charges, masses, inputs, filt="binomial2", execution="plain", comm_prelude=True,
                     pusher="boris", stencil=None, brick_shape=None):
    """(configuration, tile grid, simulation) holding `inputs` (see
    pair_plasma_inputs for the layout)."""
    import runko_amd as runko

    conf = make_configuration(nt, n, charges, masses, filt=filt, execution=execution, pusher=pusher,
                              stencil=stencil)
    if brick_shape is not None:
        conf.brick_shape = brick_shape
    grid = runko.TileGrid(conf)
    for idx in grid.local_tile_indices():
        F, parts = inputs[idx]
        tile = runko.pic.threeD.Tile(idx, conf)
        fs = [lambda x, y, z, a=F[c]: a for c in range(9)]
        tile.batch_set_EBJ(*fs)
        for s, st in enumerate(parts):
            tile.batch_inject_to_cells(
                s, lambda x, y, z, st=st: runko.pic.threeD.ParticleStateBatch(pos=(st[:, 0], st[:, 1], st[:, 2]),
                                                                            vel=(st[:, 3], st[:, 4], st[:, 5])))
        grid.add_tile(tile, idx)
    sim = grid.configure_simulation(conf)
    if comm_prelude:
        def sync(x):
            x.comm_local(runko.tools.comm_mode.emf_E)
            x.comm_local(runko.tools.comm_mode.emf_B)
        sim.prelude(sync)
    return conf, grid, sim


def turbulence_lap(sim, io=False, sort_every=5):
    """The lap function of projects/pic-turbulence/pic.py:187-221."""
    import runko_amd as runko

    cm = runko.tools.comm_mode

    def step(x):
        x.grid_push_half_b()
        x.comm_external(cm.emf_B)
        x.comm_local(cm.emf_B)
        x.prtcl_push()
        x.prtcl_pack_outgoing()
        x.comm_external(cm.pic_particle)
        x.comm_local(cm.pic_particle)
        if sort_every and sim.lap % sort_every == 0:
            x.prtcl_sort()
        x.prtcl_deposit_current()
        x.comm_external(cm.emf_J)
        x.comm_local(cm.emf_J_exchange)
        x.comm_external(cm.emf_J)
        x.comm_local(cm.emf_J)
        x.grid_filter_current()
        x.comm_external(cm.emf_J)
        x.comm_local(cm.emf_J)
        x.grid_filter_current()
        x.grid_filter_current()
        x.grid_push_half_b()
        x.comm_external(cm.emf_B)
        x.comm_local(cm.emf_B)
        x.grid_push_e()
        x.grid_add_current()
        x.comm_external(cm.emf_E)
        x.comm_local(cm.emf_E)
        if io:
            x.io_average_kinetic_energy()
            x.io_average_B_energy_density()
            x.io_average_E_energy_density()

    return step
