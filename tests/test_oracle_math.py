"""Closed forms, limits and invariants that pin the oracle's numerics
independently of the reference."""
import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import wofz

import pyoracle


@pytest.mark.parametrize("a", [1e-4, 1e-3, 0.01, 0.1, 0.5, 1.0, 3.0, 20.0])
def test_humlicek_is_the_faddeeva_function(a):
    """H + iF = w(v + i a) within Humlicek's stated 1e-4 relative accuracy
    (rh/humlicek.c:16-20); F has the sign convention of w(z) with z = a - iv."""
    for v in np.concatenate([np.linspace(-30, 30, 241), [0.0, 5.4, 5.6, 14.9, 15.1]]):
        H, F = pyoracle.voigt_humlicek(a, v)
        w = wofz(v + 1j * a)
        assert abs(H - w.real) <= 1e-4 * abs(w.real) + 1e-9
        assert abs(F - w.imag) <= 1e-4 * abs(w) + 1e-9


def test_humlicek_known_values():
    H, F = pyoracle.voigt_humlicek(0.0, 0.0)
    assert abs(H - 1.0) < 1e-4 and abs(F) < 1e-12
    for v in (40.0, 100.0):   # Lorentzian wing a/(sqrt(pi) v^2)
        a = 0.01
        H, _ = pyoracle.voigt_humlicek(a, v)
        assert abs(H / (a / (math.sqrt(math.pi) * v * v)) - 1) < 1e-3


@pytest.mark.parametrize("a", [1e-3, 0.05, 0.3, 0.9, 1.5, 2.4, 5.0])
def test_armstrong_voigt(a):
    """Armstrong (1967) series and quadratures, accurate to ~1e-4 relative
    (the v > 5 asymptotic series is the loosest)."""
    for v in np.linspace(0, 12, 97):
        w = wofz(v + 1j * a).real
        assert abs(pyoracle.voigt_armstrong(a, v) - w) <= 3e-4 * w + 1e-10


def _bezier_quadrature(dt):
    """dt * int_0^1 B(u) exp(-dt (1-u)) du for the Bernstein weights of
    S0 (u^3), Su ((1-u)^3), c0 (3u^2(1-u)), cu (3u(1-u)^2)."""
    f = lambda g: dt * quad(lambda u: g(u) * math.exp(-dt * (1 - u)), 0, 1,
                            epsabs=0, epsrel=1e-13)[0]
    return (f(lambda u: u ** 3), f(lambda u: (1 - u) ** 3),
            f(lambda u: 3 * u * u * (1 - u)), f(lambda u: 3 * u * (1 - u) ** 2))


@pytest.mark.parametrize("dt", [0.01, 0.049, 0.05, 0.2, 1.0, 5.0, 30.0])
def test_bezier3_coefficients_are_the_interpolant_integrals(dt):
    al, be, ga, th, eps = pyoracle.bezier3_coeffs(dt)
    qa, qb, qg, qt = _bezier_quadrature(dt)
    tol = 1e-8 if dt >= 0.05 else 0.05 * dt ** 4   # 4th-order Taylor below 0.05
    for c, q in ((al, qa), (be, qb), (ga, qg), (th, qt)):
        assert abs(c - q) <= tol * max(q, 1e-300) + 1e-15
    assert abs(eps - math.exp(-dt)) <= 1e-9 * math.exp(-dt) + (dt ** 5 / 100)


def test_bezier3_branches_meet_at_switch():
    lo = pyoracle.bezier3_coeffs(0.05 * (1 - 1e-12))
    hi = pyoracle.bezier3_coeffs(0.05)
    for a, b in zip(lo, hi):
        assert abs(a - b) <= 1e-6 * abs(b)


def test_w3_closed_form():
    for dt in (1e-5, 4.9e-4, 5.1e-4, 0.3, 2.0, 49.0):
        w = pyoracle.w3(dt)
        e = math.exp(-dt)
        # below 5e-4 the weights are second-order series (error ~ dt^3/6)
        rel = 1e-9 if dt >= 5e-4 else dt * dt / 5
        assert abs(w[0] - (1 - e)) <= rel * (1 - e) + 1e-16
        assert abs(w[1] - (1 - e - dt * e)) <= max(rel, 1e-6) * abs(1 - e - dt * e) + 1e-16


def test_planck_limits():
    # Wien limit and the closed form 2hc^2/lambda^5/(exp(hc/lambda kT)-1) per unit frequency
    h, c, k = 6.6260755E-34, 2.99792458E+08, 1.380658E-23
    lam = 630.0e-9
    for T in (3000.0, 6000.0, 20000.0):
        B = pyoracle.planck(T, 630.0)
        ref = 2 * h * c / lam ** 3 / (math.exp(h * c / (lam * k * T)) - 1)
        assert abs(B / ref - 1) < 1e-12


def _slab(n=120):
    z = np.linspace(1.0e6, 0.0, n)          # heights, decreasing index -> deeper
    tau = np.logspace(-7, 3, n)             # optical depth at mu = 1
    dz = -np.diff(z)
    chi = np.empty(n)
    chi[:-1] = np.diff(tau) / dz
    chi[-1] = chi[-2]
    return z, tau, chi


@pytest.mark.parametrize("mu", [1.0, 0.5])
def test_formal_constant_source(mu):
    z, tau, chi = _slab()
    S = np.full(z.size, 3.7)
    I = pyoracle.formal_bezier(z, mu, chi, S, B2=[3.7, 3.7])
    assert abs(I[0] / 3.7 - 1) < 1e-9
    S4 = np.zeros((4, z.size)); S4[0] = 3.7
    I4 = pyoracle.formal_stokes(z, mu, chi, S4, np.zeros((3, z.size)), B2=[3.7, 3.7])
    assert abs(I4[0, 0] / 3.7 - 1) < 1e-9 and np.all(I4[1:, 0] == 0)
    # Feautrier: second order inside, first-order upper boundary condition
    F = pyoracle.formal_feautrier(z, mu, chi, S, B2=[3.7, 3.7])
    assert abs(F / 3.7 - 1) < 3e-3


def test_formal_eddington_barbier_linear_source():
    """S = a + b tau  =>  I(0, mu) = a + b mu (exact for a linear source)."""
    z, tau, chi = _slab(200)
    a, b = 1.0, 2.5
    # the solvers integrate tau from chi and z with trapezoids; build the
    # source on that same tau
    t = np.concatenate([[0.0], np.cumsum(0.5 * (chi[1:] + chi[:-1]) * -np.diff(z))]) + tau[0]
    S = a + b * t
    B2 = [S[-2], S[-1]]
    I = pyoracle.formal_bezier(z, 1.0, chi, S, B2=B2)
    assert abs(I[0] - (a + b * 1.0)) / (a + b) < 2e-3
    S4 = np.zeros((4, z.size)); S4[0] = S
    I4 = pyoracle.formal_stokes(z, 1.0, chi, S4, np.zeros((3, z.size)), B2=B2)
    assert abs(I4[0, 0] - (a + b)) / (a + b) < 2e-3
    F = pyoracle.formal_feautrier(z, 1.0, chi, S, B2=B2)
    assert abs(F - (a + b)) / (a + b) < 2e-3


def test_unpolarized_delo_matches_scalar_bezier_for_smooth_source():
    """With K' = 0 the DELO-Bezier and the scalar Bezier solutions agree to
    discretisation error on a smooth source function."""
    z, tau, chi = _slab(150)
    S = 1.0 + 0.3 * np.log1p(tau) + 0.1 * np.sin(np.log(tau))
    I = pyoracle.formal_bezier(z, 1.0, chi, S, B2=[S[-2], S[-1]])
    S4 = np.zeros((4, z.size)); S4[0] = S
    I4 = pyoracle.formal_stokes(z, 1.0, chi, S4, np.zeros((3, z.size)), B2=[S[-2], S[-1]])
    assert abs(I4[0, 0] / I[0] - 1) < 2e-3
