"""Long-run agreement: the total-energy history of 1000 steps of the pair
plasma (the 64^3 workload's physics on a 16^3 box) within 0.5% of the
oracle at every step, for the plain and the batched forms.

Total energy in code units: sum(E^2 + B^2)/2 + sum_s m_s |q_s| c^2 sum_p (gamma_p - 1)
(projects/pic-turbulence/pic.py:52-53 macro-masses)."""

import numpy as np
import pytest

import runko_amd as runko
from gpu_common import CFL, build_gpu, build_oracle, canonical_lap
from runko_amd.init import plasma_parameters
from runko_amd.sampling import sample_boosted_juttner_synge

pytestmark = pytest.mark.gpu


def pair_plasma_inputs(nt, n, ppc, theta, sigma, seed):
    q0, q1, B0 = plasma_parameters(CFL, 10.0, ppc, theta, sigma)
    rng = np.random.default_rng(seed)
    cells = np.stack(np.meshgrid(*(np.arange(v, dtype=np.float64) for v in n), indexing="ij"), -1).reshape(-1, 3)
    tiles = {}
    for i in range(nt[0]):
        for j in range(nt[1]):
            for k in range(nt[2]):
                lo = np.array([i * n[0], j * n[1], k * n[2]], dtype=np.float64)
                pos = np.repeat(cells, ppc, 0) + rng.random((len(cells) * ppc, 3)) + lo
                parts = []
                for s in range(2):
                    u = np.stack(sample_boosted_juttner_synge(len(pos), theta, beta=0.0, gen=rng), 1)
                    parts.append(np.concatenate([pos, u], 1))
                F = np.zeros((9,) + tuple(n))
                F[5] = B0
                tiles[(i, j, k)] = (F, parts)
    return (q0, q1), tiles


def total_energy(e, charges):
    return e[0] + e[1] + sum(abs(q) * 1.0 * CFL**2 * e[2 + s] for s, q in enumerate(charges))


@pytest.mark.parametrize("execution", ["plain", "batched"])
def test_energy_history_1000_steps(orc, execution):
    nt, n, ppc, steps = (2, 2, 2), (8, 8, 8), 8, 1000
    charges, inputs = pair_plasma_inputs(nt, n, ppc, theta=0.3, sigma=1.0, seed=31)
    conf, grid, sim = build_gpu(nt, n, list(charges), [1.0, 1.0], inputs, execution=execution)
    w = build_oracle(orc, nt, n, list(charges), [1.0, 1.0], inputs)
    step = canonical_lap(sim)
    e0 = total_energy(grid.world.energies(), charges)
    assert total_energy(w.energies(), charges) == pytest.approx(e0, rel=1e-6)
    worst = worst_kin = 0.0
    for lap in range(steps):
        sim.for_one_lap(step)
        w.lap(lap)
        if lap % 10 == 9 or lap == steps - 1:
            ag, ao = grid.world.energies(), w.energies()
            eg, eo = total_energy(ag, charges), total_energy(ao, charges)
            worst = max(worst, abs(eg - eo) / abs(eo))
            assert abs(eg - eo) <= 5e-3 * abs(eo), (lap, eg, eo)
            # the kinetic part alone (the B0 guide field dominates the total)
            kg, ko = ag[2:].sum(), ao[2:].sum()
            worst_kin = max(worst_kin, abs(kg - ko) / ko)
            assert abs(kg - ko) <= 5e-3 * ko, (lap, kg, ko)
    print(f"[energy history] {execution}: worst relative difference {worst:.2e} (total), "
          f"{worst_kin:.2e} (kinetic) over {steps} steps")
    assert np.all(grid.world.error_counters() == 0)
