"""Thermal momentum samplers for initial conditions (host side, numpy).

Restates runko/sample_thermal_distributions.py:7-127: the Juttner-Synge
distribution by the Sobol rejection method for theta > 0.2 (Zenitani 2015,
eq. 2), a Maxwellian (Rayleigh inverse transform) below, an isotropic
direction, and the Zenitani flipping + boost.  The generator calls are made
in the same order and the expressions keep the same evaluation order, so a
seeded numpy Generator gives the same samples as the reference sampler.
"""

import numpy as np
from scipy.special import kn

_default_gen = np.random.default_rng()


def sample_juttner_synge_sobol_method(size, theta, gen=_default_gen):
    eff = kn(2, 1 / theta) / (2 * theta**2)
    draws_per_accept = 1.001
    accepted = []
    n_have = 0
    while n_have < size:
        n_try = int(draws_per_accept * (size - n_have) / eff)
        x4 = gen.random(size=n_try)
        x5 = gen.random(size=n_try)
        x6 = gen.random(size=n_try)
        x7 = gen.random(size=n_try)
        u = -theta * np.log(x4 * x5 * x6)
        eta = -theta * np.log(x4 * x5 * x6 * x7)
        ok = (eta**2 - u**2) > 1.0
        accepted.append(u[ok])
        n_have += len(accepted[-1])
    return np.concatenate(accepted)[:size]


def sample_maxwellian_box_muller_method(size, theta, gen=_default_gen):
    vth = np.sqrt(2 * theta)
    return vth * np.sqrt(-2.0 * np.log(gen.random(size=size)))


def sample_boosted_juttner_synge(size, theta, beta=None, Gamma=None, direction="-x", gen=_default_gen):
    """Spatial 4-velocity components (ux, uy, uz) of a boosted Juttner-Synge
    distribution with temperature theta = kT/mc^2; the frame moves with
    4-velocity (Gamma, -Gamma*beta) along `direction`."""
    if (beta is not None) == (Gamma is not None):
        raise RuntimeError("Either beta or Gamma has to be given but not both.")
    if beta is not None:
        Gamma = 1.0 / np.sqrt(1.0 - beta**2)
    else:
        beta = np.sqrt(1 - Gamma**(-2))

    if theta > 0.2:
        u = sample_juttner_synge_sobol_method(size, theta, gen)
    else:
        u = sample_maxwellian_box_muller_method(size, theta, gen)

    X1 = gen.random(size=size)
    X2 = gen.random(size=size)
    ux = u * (2 * X1 - 1)
    uy = 2 * u * np.sqrt(X1 * (1 - X1)) * np.cos(2 * np.pi * X2)
    uz = 2 * u * np.sqrt(X1 * (1 - X1)) * np.sin(2 * np.pi * X2)

    vx = ux / np.sqrt(1.0 + u**2)
    flip = (-beta * vx) > gen.random(size=size)
    ux[flip] *= -1
    ux = Gamma * (ux + beta * np.sqrt(1 + u**2))

    perm = {
        "x": lambda: (ux, uy, uz),
        "-x": lambda: (-ux, uy, uz),
        "y": lambda: (uz, ux, uy),
        "-y": lambda: (uz, -ux, uy),
        "z": lambda: (uy, uz, ux),
        "-z": lambda: (uy, uz, -ux),
    }
    if direction not in perm:
        raise RuntimeError("Invalid direction given!")
    return perm[direction]()
