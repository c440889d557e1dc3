"""Writes tests/golden/sampler_vectors.npz: seeded outputs of the reference's
thermal samplers (runko/sample_thermal_distributions.py, hel-astro-lab/runko
v5.0.0), the one part of the initial conditions whose exact draw sequence
nothing else pins (moments and closed forms pin the distribution, not which
numbers a seed gives).

Run once, in a checkout that has the reference at /root/reference; the
module is loaded from its file alone (it needs only numpy and scipy), not
through the runko package.  Only the inputs and outputs are kept.

    python tests/golden/make_sampler_vectors.py
"""

import importlib.util
import os

import numpy as np

REF = "/root/reference/runko/sample_thermal_distributions.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sampler_vectors.npz")

# (size, theta, beta, Gamma, direction, seed)
BOOSTED = [
    (300, 0.3, 0.0, None, "x", 11),       # Sobol branch, no boost
    (300, 1.5, 0.5, None, "-x", 12),      # Sobol branch, boost along -x
    (300, 0.05, 0.0, None, "y", 13),      # Box-Muller branch
    (300, 1e-4, None, 10.0, "-x", 14),    # shock upstream: cold, gamma 10 along -x
    (300, 0.8, 0.3, None, "-z", 15),
    (257, 3.0, None, 2.0, "z", 16),
]
SOBOL = [(400, 0.3, 21), (400, 2.0, 22), (400, 0.6, 23)]       # theta > 0.2: the branch the drivers use it on
BOXMULLER = [(400, 0.3, 31), (400, 0.01, 32), (400, 1e-4, 33)]


def main():
    spec = importlib.util.spec_from_file_location("ref_sampler", REF)
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)
    arrays = {}
    for i, (size, theta, beta, gamma, direction, seed) in enumerate(BOOSTED):
        gen = np.random.default_rng(seed)
        ux, uy, uz = ref.sample_boosted_juttner_synge(size, theta, beta=beta, Gamma=gamma, direction=direction,
                                                      gen=gen)
        arrays[f"boosted{i}_u"] = np.stack([ux, uy, uz])
    for i, (size, theta, seed) in enumerate(SOBOL):
        arrays[f"sobol{i}"] = ref.sample_juttner_synge_sobol_method(size, theta, np.random.default_rng(seed))
    for i, (size, theta, seed) in enumerate(BOXMULLER):
        arrays[f"boxmuller{i}"] = ref.sample_maxwellian_box_muller_method(size, theta, np.random.default_rng(seed))
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {sorted(arrays)}")


if __name__ == "__main__":
    main()
