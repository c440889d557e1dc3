"""GPU-side initial conditions of the turbulence workload (runko_amd.init):
positions uniform with exactly ppc per cell, the Juttner-Synge momenta of the
Sobol rejection branch against the distribution's own moments (closed form
and quadrature), isotropy, the guide field and the dB modes (rms dB/B0 and a
discrete div B that vanishes to rounding), and the flipping boost of the
shock flow against the moments of the boosted distribution."""

import math

import numpy as np
import pytest
import torch
from scipy import integrate
from scipy.special import kn

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from runko_amd.init import boost_flow, juttner_momenta, turbulence_world  # noqa: E402

THETA = 0.3


def _all(w, s):
    ps, vs = [], []
    for t in range(w.ntiles):
        p, v, _ = w.tile_particles(t, s)
        ps.append(p.double().cpu())
        vs.append(v.double().cpu())
    return torch.cat(ps, 1).numpy(), torch.cat(vs, 1).numpy()


def _within(sample, expected, k=5.0):
    """|mean - expected| within k standard errors of the sample mean."""
    m = sample.mean()
    se = sample.std() / math.sqrt(sample.size)
    return abs(m - expected) <= k * se, (m, expected, se)


def _mj_moment(f, theta):
    """<f(gamma)> of the Juttner-Synge law gamma sqrt(gamma^2-1) exp(-gamma/theta)."""
    w = lambda g: g * math.sqrt(g * g - 1.0) * math.exp(-(g - 1.0) / theta)
    top = 1.0 + 60 * theta
    z = integrate.quad(w, 1.0, top, limit=200)[0]
    return integrate.quad(lambda g: f(g) * w(g), 1.0, top, limit=200)[0] / z


@pytest.fixture(scope="module")
def world():
    w, info = turbulence_world(N=64, tile=32, ppc=16, theta=THETA, sigma=10.0, seed=4)
    return w, info


def test_positions_fill_every_cell_with_ppc(world):
    w, _ = world
    for s in range(2):
        p, _ = _all(w, s)
        cells = np.floor(p).astype(np.int64)
        assert cells.min() >= 0 and cells.max() < 64
        key = (cells[0] * 64 + cells[1]) * 64 + cells[2]
        cnt = np.bincount(key, minlength=64 ** 3)
        assert cnt.min() == 16 and cnt.max() == 16
    p0, _ = _all(w, 0)
    p1, _ = _all(w, 1)
    assert np.array_equal(np.sort(p0[0]), np.sort(p1[0]))


def test_juttner_momenta_moments(world):
    w, _ = world
    g_closed = kn(1, 1 / THETA) / kn(2, 1 / THETA) + 3 * THETA
    g_quad = _mj_moment(lambda g: g, THETA)
    assert g_quad == pytest.approx(g_closed, rel=1e-9)
    u2_quad = _mj_moment(lambda g: g * g - 1.0, THETA)
    for s in range(2):
        _, v = _all(w, s)
        u2 = (v * v).sum(0)
        gam = np.sqrt(1.0 + u2)
        ok, info = _within(gam, g_closed)
        assert ok, f"<gamma> {info}"
        ok, info = _within(u2, u2_quad)
        assert ok, f"<u^2> {info}"
        # isotropy: zero mean per component, each component a third of <u^2>
        for d in range(3):
            ok, info = _within(v[d], 0.0)
            assert ok, f"<u_{d}> {info}"
            ok, info = _within(v[d] ** 2, u2_quad / 3)
            assert ok, f"<u_{d}^2> {info}"
        # the distribution function, not only its moments: the fraction of
        # gamma below a few quantiles of the quadrature CDF
        for q in (1.2, 1.5, 2.0, 3.0):
            frac = _mj_moment(lambda g, q=q: float(g <= q), THETA)
            got = (gam <= q).mean()
            se = math.sqrt(frac * (1 - frac) / gam.size)
            assert abs(got - frac) <= 5 * se, (q, got, frac)


def test_turbulent_field(world):
    w, info = world
    B0 = info["B0"]
    inner = (slice(3, 67),) * 3
    B = w.B[(slice(None),) + inner].double().cpu().numpy()
    E = w.E[(slice(None),) + inner].cpu().numpy()
    assert np.all(E == 0.0)
    assert np.all(B[2] == np.float32(B0))
    rms = math.sqrt(float((B[0] ** 2 + B[1] ** 2).mean()))
    assert rms / B0 == pytest.approx(1.0, rel=1e-5)
    # dBx, dBy do not depend on z, and carry modes 1..8 in x and y only
    assert np.all(B[0] == B[0][:, :, :1]) and np.all(B[1] == B[1][:, :, :1])
    spec = np.abs(np.fft.fft2(B[0][:, :, 0]))
    m, n = np.nonzero(spec > 1e-6 * spec.max())
    assert set(m.tolist()) <= set(range(1, 9)) | set(range(56, 64))
    assert set(n.tolist()) <= set(range(1, 9)) | set(range(56, 64))
    # discrete divergence on the Yee stagger (periodic): zero up to FP32 rounding
    div = (np.roll(B[0], -1, 0) - B[0]) + (np.roll(B[1], -1, 1) - B[1]) + (np.roll(B[2], -1, 2) - B[2])
    assert np.abs(div).max() < 1e-6 * B0


def test_boost_flow_flipping_moments():
    # the lab-frame distribution of a flow along -x with Lorentz factor G is
    # exp(-G (gamma + beta u_x) / theta) d^3u; its <u_x> and <gamma> by
    # quadrature in (u_par, u_perp)
    theta, G = 0.5, 3.0
    beta = math.sqrt(1 - 1 / G**2)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    n = 1 << 22
    u = boost_flow(juttner_momenta(n, theta, gen, "cuda").double(), G, gen).cpu().numpy()

    def moment(f):
        def dens(up, ux):
            g = math.sqrt(1 + ux * ux + up * up)
            return 2 * math.pi * up * math.exp(-(G * (g + beta * ux) - 1) / theta)
        lim = 40.0
        z = integrate.dblquad(dens, -lim, lim, 0, lim, epsabs=1e-12)[0]
        return integrate.dblquad(lambda up, ux: f(ux, up) * dens(up, ux), -lim, lim, 0, lim,
                                 epsabs=1e-12)[0] / z

    ux_q = moment(lambda ux, up: ux)
    g_q = moment(lambda ux, up: math.sqrt(1 + ux * ux + up * up))
    gam = np.sqrt(1 + (u * u).sum(0))
    ok, info = _within(u[0], ux_q)
    assert ok, f"<u_x> {info}"
    ok, info = _within(gam, g_q)
    assert ok, f"<gamma> {info}"
    # the plain boost (no flipping) would give <u_x> = -G beta <gamma>_rest, which differs
    plain = -G * beta * (kn(1, 1 / theta) / kn(2, 1 / theta) + 3 * theta)
    se = u[0].std() / math.sqrt(n)
    assert abs(plain - ux_q) > 10 * se
