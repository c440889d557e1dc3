"""The thermal samplers of runko_amd.sampling against seeded outputs of the
reference's own samplers (tests/golden/sampler_vectors.npz, written by
tests/golden/make_sampler_vectors.py): the same generator seed must give the
same numbers, bit for bit, which pins the draw sequence and not only the
distribution."""

import os

import numpy as np
import pytest

from runko_amd import sampling

HERE = os.path.dirname(os.path.abspath(__file__))
V = np.load(os.path.join(HERE, "golden", "sampler_vectors.npz"))  # allow_pickle stays False

# the cases of make_sampler_vectors.py
BOOSTED = [
    (300, 0.3, 0.0, None, "x", 11),
    (300, 1.5, 0.5, None, "-x", 12),
    (300, 0.05, 0.0, None, "y", 13),
    (300, 1e-4, None, 10.0, "-x", 14),
    (300, 0.8, 0.3, None, "-z", 15),
    (257, 3.0, None, 2.0, "z", 16),
]
SOBOL = [(400, 0.3, 21), (400, 2.0, 22), (400, 0.6, 23)]
BOXMULLER = [(400, 0.3, 31), (400, 0.01, 32), (400, 1e-4, 33)]


@pytest.mark.parametrize("i", range(len(BOOSTED)))
def test_boosted_juttner_synge_matches_reference_vectors(i):
    size, theta, beta, gamma, direction, seed = BOOSTED[i]
    u = sampling.sample_boosted_juttner_synge(size, theta, beta=beta, Gamma=gamma, direction=direction,
                                              gen=np.random.default_rng(seed))
    assert np.array_equal(np.stack(u), V[f"boosted{i}_u"])


@pytest.mark.parametrize("i", range(len(SOBOL)))
def test_sobol_method_matches_reference_vectors(i):
    size, theta, seed = SOBOL[i]
    assert np.array_equal(sampling.sample_juttner_synge_sobol_method(size, theta, np.random.default_rng(seed)),
                          V[f"sobol{i}"])


@pytest.mark.parametrize("i", range(len(BOXMULLER)))
def test_box_muller_method_matches_reference_vectors(i):
    size, theta, seed = BOXMULLER[i]
    assert np.array_equal(sampling.sample_maxwellian_box_muller_method(size, theta, np.random.default_rng(seed)),
                          V[f"boxmuller{i}"])


def test_vectors_are_not_trivial():
    # a plausible slip (another seed, a dropped draw) changes every value
    u = sampling.sample_boosted_juttner_synge(300, 0.3, beta=0.0, direction="x", gen=np.random.default_rng(12))
    assert not np.array_equal(np.stack(u), V["boosted0_u"])
    assert V["sobol0"].shape == (400,) and np.all(V["sobol0"] > 0)


def test_boosted_sampler_matches_lab_frame_moments():
    # the flipping method gives the lab-frame Juttner-Synge law
    # exp(-G (gamma + beta u_x) / theta) d^3u of a flow along -x; its <u_x>
    # and <gamma> by quadrature in (u_par, u_perp)
    import math

    from scipy import integrate

    theta, G = 0.5, 3.0
    beta = math.sqrt(1 - 1 / G**2)

    def moment(f):
        def dens(up, ux):
            g = math.sqrt(1 + ux * ux + up * up)
            return 2 * math.pi * up * math.exp(-(G * (g + beta * ux) - 1) / theta)
        z = integrate.dblquad(dens, -40, 40, 0, 40, epsabs=1e-12)[0]
        return integrate.dblquad(lambda up, ux: f(ux, up) * dens(up, ux), -40, 40, 0, 40, epsabs=1e-12)[0] / z

    n = 1_000_000
    ux, uy, uz = sampling.sample_boosted_juttner_synge(n, theta, Gamma=G, direction="-x",
                                                       gen=np.random.default_rng(3))
    g = np.sqrt(1 + ux**2 + uy**2 + uz**2)
    for sample, expect in ((ux, moment(lambda a, b: a)), (g, moment(lambda a, b: math.sqrt(1 + a * a + b * b)))):
        assert abs(sample.mean() - expect) < 5 * sample.std() / math.sqrt(n)
