"""The oracle's Higuera-Cary and Faraday pushers and the extended-stencil B
push (oracle/pic_oracle.cpp), pinned by properties a plausible mistake breaks:

* Higuera-Cary with B = 0 gives Boris's momentum to one rounding; Faraday
  with E = 0 gives Boris's momentum to round-off (both reduce to Boris there);
* all three pushers converge at second order to an RK4 solution of
  du/dt = (q/m)(E + v x B), dx/dt = v (fields scaled by cfl^2: the code-unit
  fields of the push are (q/m) E dt^2 / dx in these units);
* Higuera-Cary and Faraday keep a particle at the relativistic E x B drift
  velocity, where Boris does not (the property the schemes were built for);
* the stencil push with zero coefficients is fdtd2 bit for bit, its curl of
  linear fields is exact for any coefficients (the alpha normalisation), and
  its transverse columns act along the documented axes (perp1 of z is x)."""

import numpy as np
import pytest

from oracle import OracleWorld, stencil_matrices

ORIGIN = (8.3, 8.6, 8.1)


def push_world(pusher, cfl, E, B, u0, prec=64, n=16):
    w = OracleWorld((1, 1, 1), (n, n, n), cfl, [1e-30], [1.0], prec=prec)
    F = np.zeros((9, n, n, n))
    for c in range(3):
        F[c] = E[c]
        F[3 + c] = B[c]
    w.set_fields(0, F)
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    w.add_particles(0, 0, np.array([[*ORIGIN, *u0]]))
    w.set_pusher(pusher)
    return w


def state(w):
    st, _ = w.get_particles(0, 0)
    return st[0]


def run(pusher, cfl, E, B, u0, steps, prec=64):
    w = push_world(pusher, cfl, E, B, u0, prec)
    for _ in range(steps):
        w.push_particles(0)
    return state(w)


def rk4(Ep, Bp, x0, u0, T, h=1e-3):
    """(x, u) at time T of du/dt = E + v x B, dx/dt = v (T may be negative)."""
    u = np.array(u0, dtype=float)
    x = np.array(x0, dtype=float)
    n = max(1, int(round(abs(T) / h)))
    dt = T / n

    def f(x, u):
        g = np.sqrt(1 + u @ u)
        v = u / g
        return v, Ep + np.cross(v, Bp)

    for _ in range(n):
        k1x, k1u = f(x, u)
        k2x, k2u = f(x + 0.5 * dt * k1x, u + 0.5 * dt * k1u)
        k3x, k3u = f(x + 0.5 * dt * k2x, u + 0.5 * dt * k2u)
        k4x, k4u = f(x + dt * k3x, u + dt * k3u)
        x = x + dt / 6 * (k1x + 2 * k2x + 2 * k3x + k4x)
        u = u + dt / 6 * (k1u + 2 * k2u + 2 * k3u + k4u)
    return x, u


@pytest.mark.parametrize("prec", [32, 64])
def test_higuera_cary_without_b_is_boris(prec):
    E, u0 = (0.01, -0.02, 0.005), (0.4, -0.3, 0.7)
    a = run("higuera_cary", 0.45, E, (0, 0, 0), u0, 5, prec)
    b = run("boris", 0.45, E, (0, 0, 0), u0, 5, prec)
    # same u2; Higuera-Cary scales it by the rounded 1/cfl, Boris divides by cfl
    ulp = 2.0**-23 if prec == 32 else 2.0**-52
    assert np.all(np.abs(a[3:] - b[3:]) <= 2 * ulp * np.abs(b[3:]))
    assert np.allclose(a[:3], b[:3], rtol=0, atol=1e-5 if prec == 32 else 1e-13)


def test_faraday_without_e_is_boris():
    B, u0 = (0.02, -0.05, 0.09), (0.4, -0.3, 0.7)
    a = run("faraday", 0.45, (0, 0, 0), B, u0, 20)
    b = run("boris", 0.45, (0, 0, 0), B, u0, 20)
    assert np.allclose(a, b, rtol=0, atol=1e-12)


@pytest.mark.parametrize("pusher", ["boris", "higuera_cary", "faraday"])
def test_second_order_convergence(pusher):
    # leapfrog: positions at whole steps, momenta at half steps; the run
    # starts from u(-dt/2) and ends with x(T), u(T - dt/2)
    Ep, Bp, u0, T = np.array([0.3, 0.1, -0.2]), np.array([0.1, 0.2, 1.0]), (0.5, -0.2, 0.3), 3.2
    xT, _ = rk4(Ep, Bp, ORIGIN, u0, T)
    errs = []
    for cfl in (0.2, 0.1, 0.05):
        _, u_start = rk4(Ep, Bp, ORIGIN, u0, -cfl / 2)
        _, u_end = rk4(Ep, Bp, ORIGIN, u0, T - cfl / 2)
        st = run(pusher, cfl, Ep * cfl**2, Bp * cfl**2, u_start, int(round(T / cfl)))
        errs.append(np.linalg.norm(st - np.concatenate([xT, u_end])))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.3 < r1 < 4.8 and 3.3 < r2 < 4.8, (errs, r1, r2)
    assert errs[2] < 1e-3


def test_pushers_differ():
    Ep, Bp, u0 = np.array([0.3, 0.1, -0.2]), np.array([0.1, 0.2, 1.0]), (0.5, -0.2, 0.3)
    s = [run(p, 0.2, Ep * 0.04, Bp * 0.04, u0, 16) for p in ("boris", "higuera_cary", "faraday")]
    assert np.abs(s[0] - s[1]).max() > 1e-8 and np.abs(s[0] - s[2]).max() > 1e-8 and np.abs(s[1] - s[2]).max() > 1e-8


def test_higuera_cary_keeps_exb_drift():
    # E = 0.9 y, B = z: v_d = E x B / B^2 = 0.9 x, a particle moving with it
    # feels no force; Higuera-Cary keeps it there, Boris does not
    cfl = 0.45
    E, B = np.array([0.0, 0.9, 0.0]) * cfl**2, np.array([0.0, 0.0, 1.0]) * cfl**2
    ud = 0.9 / np.sqrt(1 - 0.81)
    # 10 steps: the particle moves 4 cells and stays inside the 16-cell tile
    hc = run("higuera_cary", cfl, E, B, (ud, 0.0, 0.0), 10)
    fa = run("faraday", cfl, E, B, (ud, 0.0, 0.0), 10)
    bo = run("boris", cfl, E, B, (ud, 0.0, 0.0), 10)
    assert np.abs(hc[3:] - [ud, 0, 0]).max() < 1e-10
    assert np.abs(fa[3:] - [ud, 0, 0]).max() < 1e-10
    assert np.abs(bo[3:] - [ud, 0, 0]).max() > 1e-4


COEFFS = [dict(delta=-0.06, gamma=0.004, beta_p1=0.011, beta_p2=-0.007, beta2_p1=0.002, beta2_p2=0.001,
               beta3_p1=-0.0005, beta3_p2=0.0003, zeta_p1=0.003, zeta_p2=-0.004, zeta2_p1=0.0007,
               zeta2_p2=-0.0002, zeta3_p1=0.0001, zeta3_p2=0.00015),
          dict(delta=-0.05, beta_p1=0.02, beta_p2=-0.01, zeta_p2=0.005),
          dict(gamma=0.01, beta_p2=0.03, zeta_p1=-0.006, beta2_p1=0.004)]


def field_world(F, prec, stencil=None):
    n = F.shape[1:]
    w = OracleWorld((1, 1, 1), n, 0.45, [], [], prec=prec)
    w.set_fields(0, F)
    w.comm_local("emf_E")
    w.comm_local("emf_B")
    if stencil is not None:
        w.set_stencil(stencil)
    return w


@pytest.mark.parametrize("prec", [32, 64])
def test_stencil_with_zero_coefficients_is_fdtd2(prec):
    rng = np.random.default_rng(71)
    F = rng.standard_normal((9, 8, 9, 10))
    a = field_world(F, prec, stencil_matrices([{}, {}, {}]))
    b = field_world(F, prec)
    a.push_half_b(0)
    b.push_half_b(0)
    assert np.array_equal(a.global_fields(), b.global_fields())


def test_stencil_normalisation():
    M = stencil_matrices(COEFFS)
    for a in range(3):
        w = np.array([1.0, 3.0, 5.0]).reshape(3, 1) * np.array([1.0, 2, 2, 2, 2]).reshape(1, 5)
        assert abs(float((M[a] * w).sum()) - 1.0) < 1e-6


def test_stencil_curl_of_linear_fields_is_exact():
    # E_c = sum_a G[c][a] x_a: curl is the constant antisymmetric part; read
    # away from the periodic seams (the stencil reaches -2 .. +3)
    n = (12, 12, 12)
    G = np.array([[0.3, -0.7, 0.2], [0.5, 0.1, -0.4], [0.9, 0.6, -0.2]])
    X = np.stack(np.meshgrid(*(np.arange(v, dtype=float) for v in n), indexing="ij"))
    F = np.zeros((9,) + n)
    # staggered positions: E_c at +1/2 along c (a constant offset, exact for linear fields)
    for c in range(3):
        F[c] = sum(G[c][a] * (X[a] + (0.5 if a == c else 0.0)) for a in range(3))
    w = field_world(F, 64, stencil_matrices(COEFFS))
    w.push_half_b(0)
    B = w.global_fields()[3:6]
    # B advances by -dt curl E (Faraday's law), dt = cfl / 2
    curl = -np.array([G[2][1] - G[1][2], G[0][2] - G[2][0], G[1][0] - G[0][1]]) * (0.45 / 2)
    core = (slice(2, -3),) * 3
    for c in range(3):
        # exact up to the float32 normalisation of alpha (sum of weights = 1 +- 1e-7)
        assert np.allclose(B[c][core], curl[c], rtol=0, atol=1e-7), c


def test_stencil_transverse_axes():
    # E_y = g(x) z only: Bx = dt D*_z E_y, whose columns 1 / 3 (perp1 of z) see
    # g(x -+ 1) / g(x -+ 2) and columns 2 / 4 (perp2 = y) see 2 g(x)
    n = (10, 6, 12)
    rng = np.random.default_rng(72)
    g = rng.standard_normal(n[0])
    z = np.arange(n[2], dtype=float)
    F = np.zeros((9,) + n)
    F[1] = g[:, None, None] * z[None, None, :]
    coeffs = [dict(), dict(), COEFFS[0]]
    M = stencil_matrices(coeffs)[2].astype(float)
    w = field_world(F, 64, stencil_matrices(coeffs))
    w.push_half_b(0)
    Bx = w.global_fields()[3]
    gi = lambda d: np.roll(g, -d)
    expect = np.zeros(n[0])
    for r in range(3):
        odd = 2 * r + 1
        expect += odd * (M[r, 0] * g + M[r, 1] * (gi(1) + gi(-1)) + M[r, 2] * 2 * g + M[r, 3] * (gi(2) + gi(-2))
                         + M[r, 4] * 2 * g)
    expect *= 0.45 / 2
    core = Bx[:, :, 2:-3]
    assert np.allclose(core, expect[:, None, None], rtol=0, atol=1e-12)
