"""Pins of the oracle's adjoint and El Karoui fit functions (CPU).

The adjoint of the at-points sums is checked against central finite
differences of the (separately pinned) forward sums, against two exact
invariances of S(x, lam, eta) = sum_j 1/((x - lam_j) - i eta) -- translation
(S depends on x - lam only) and homogeneity of degree -1 -- and, for the eta
and query parts, against the reference's own dS/dx (autovec.rs:38-59) through
dS/deta = -i dS/dx.

The El Karoui residual is checked against the reference's golden master
(deconvolution/mod.rs:281-397: w = z/a and m = a g) and against the
Marchenko-Pastur law, whose Stieltjes transform is computed here by
quadrature of the MP density, independently of the equation's code: the
residual must vanish there.  The fit's gradients are checked by central
finite differences of its loss.
"""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


def _case(seed=0, p=40, nq=25, eta=0.05):
    rng = np.random.default_rng(seed)
    ev = np.sort(rng.uniform(0.5, 2.0, p))
    q = np.linspace(0.3, 2.2, nq) + rng.uniform(-0.01, 0.01, nq)
    gre = rng.standard_normal(nq)
    gim = rng.standard_normal(nq)
    return ev, q, eta, gre, gim


def _loss(orc, q, ev, eta, gre, gim):
    sr, si = orc.at_points(q, ev, eta)
    return float(np.dot(gre, sr) + np.dot(gim, si))


def _rel(a, b):
    a, b = np.atleast_1d(a), np.atleast_1d(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_adjoint_matches_central_differences(orc):
    ev, q, eta, gre, gim = _case()
    gev, gq, ge = orc.adjoint_at_points(q, ev, eta, gre, gim)
    h = 1e-6
    fd_ev = np.empty_like(ev)
    for j in range(ev.size):
        e1, e0 = ev.copy(), ev.copy()
        e1[j] += h
        e0[j] -= h
        fd_ev[j] = (_loss(orc, q, e1, eta, gre, gim) - _loss(orc, q, e0, eta, gre, gim)) / (2 * h)
    fd_q = np.empty_like(q)
    for m in range(q.size):
        q1, q0 = q.copy(), q.copy()
        q1[m] += h
        q0[m] -= h
        fd_q[m] = (_loss(orc, q1, ev, eta, gre, gim) - _loss(orc, q0, ev, eta, gre, gim)) / (2 * h)
    he = 1e-7
    fd_e = (_loss(orc, q, ev, eta + he, gre, gim) - _loss(orc, q, ev, eta - he, gre, gim)) / (2 * he)
    assert _rel(gev, fd_ev) < 1e-6
    assert _rel(gq, fd_q) < 1e-6
    assert abs(ge - fd_e) / abs(fd_e) < 1e-6


def test_adjoint_invariances(orc):
    ev, q, eta, gre, gim = _case(seed=1, p=300, nq=90, eta=0.02)
    gev, gq, ge = orc.adjoint_at_points(q, ev, eta, gre, gim)
    scale = np.sum(np.abs(gev)) + np.sum(np.abs(gq))
    # translation: L(x + s, lam + s) = L  =>  sum grad_lam + sum grad_x = 0
    assert abs(gev.sum() + gq.sum()) <= 1e-12 * scale
    # homogeneity: L(a x, a lam, a eta) = L / a  =>  lam.g_lam + x.g_x + eta.g_eta = -L
    L = _loss(orc, q, ev, eta, gre, gim)
    lhs = np.dot(ev, gev) + np.dot(q, gq) + eta * ge
    assert abs(lhs + L) <= 1e-11 * (np.dot(np.abs(ev), np.abs(gev)) + abs(L))


def test_adjoint_eta_and_query_parts_match_reference_derivative(orc):
    """All points (x = lam): with the reference's dS/dx = (D_re, D_im)
    (autovec.rs:38-59), grad_x = G_re D_re + G_im D_im and
    grad_eta = sum G_re D_im - G_im D_re (dS/deta = -i dS/dx)."""
    rng = np.random.default_rng(2)
    ev = np.sort(rng.uniform(0.3, 3.0, 200))
    eta = 0.03
    gre, gim = rng.standard_normal(200), rng.standard_normal(200)
    _, gq, ge = orc.adjoint_at_points(ev, ev, eta, gre, gim)
    _, _, dre, dim = orc.with_deriv(ev, eta)
    assert _rel(gq, gre * dre + gim * dim) < 1e-13
    want = float(np.sum(gre * dim - gim * dre))
    assert abs(ge - want) <= 1e-13 * np.sum(np.abs(gre * dim) + np.abs(gim * dre))


def test_ek_residual_a_matches_golden_master(orc):
    """a(z) = 1 - c - c z g as the reference forms it: its golden master holds
    Im m = Im(a g) and w_re = Re(z / a) (deconvolution/mod.rs:206-248, 281-397)."""
    from shrinkers_amd.deconvolution import deconvolution_grid
    g = json.load(open(GOLDEN))["golden_master"]
    ev = np.asarray(g["evals"], dtype=np.float64)
    grid = deconvolution_grid(ev, g["n_points"])
    sr, si = orc.at_points(grid, ev, g["eta"])
    p = ev.size
    g_re, g_im = -sr / p, si / p
    np.testing.assert_allclose(g_im, g["sample_stieltjes_imag"], rtol=1e-13, atol=1e-15)
    _, a = orc.ek_residual(grid, g_re, g_im, g["eta"], g["c"], [1.0], [1.0])
    gz = g_re + 1j * g_im
    m = a * gz
    z = grid + 1j * g["eta"]
    np.testing.assert_allclose(m.imag, g["population_stieltjes_imag"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose((z / a).real, g["w_re"], rtol=1e-12, atol=1e-14)


def _mp_stieltjes_quadrature(z, c, sigma2):
    """g(z) = int rho(x)/(x - z) dx for the MP(c, sigma2) law, by adaptive
    quadrature in the angle x = sigma2 (m - r cos th), where
    rho dx = sigma2 (r sin th)^2 / (2 pi c x/sigma2) dth is smooth."""
    from scipy.integrate import quad
    m0, r = 1.0 + c, 2.0 * math.sqrt(c)

    def parts(th):
        xs = m0 - r * math.cos(th)
        w = (r * math.sin(th)) ** 2 / (2.0 * math.pi * c * xs)
        return w / (sigma2 * xs - z)

    re = quad(lambda th: parts(th).real, 0.0, math.pi, epsabs=1e-13, epsrel=1e-12, limit=400)[0]
    im = quad(lambda th: parts(th).imag, 0.0, math.pi, epsabs=1e-13, epsrel=1e-12, limit=400)[0]
    return re + 1j * im


@pytest.mark.parametrize("c,sigma2", [(0.25, 1.0), (0.5, 1.6)])
def test_ek_residual_vanishes_on_marchenko_pastur(orc, c, sigma2):
    eta = 0.05
    lo, hi = sigma2 * (1 - math.sqrt(c)) ** 2, sigma2 * (1 + math.sqrt(c)) ** 2
    x = np.linspace(lo - 0.3, hi + 0.3, 23)
    g = np.array([_mp_stieltjes_quadrature(xx + 1j * eta, c, sigma2) for xx in x])
    # H = delta_{sigma2}, also split over two coincident atoms
    r1, _ = orc.ek_residual(x, g.real, g.imag, eta, c, [sigma2], [1.0])
    r2, _ = orc.ek_residual(x, g.real, g.imag, eta, c, [sigma2, sigma2], [0.3, 0.7])
    assert np.max(np.abs(r1)) < 1e-10 * np.max(np.abs(g))
    assert np.max(np.abs(r2)) < 1e-10 * np.max(np.abs(g))
    # the check discriminates: a wrong ratio or population leaves a residual
    r3, _ = orc.ek_residual(x, g.real, g.imag, eta, c * 1.1, [sigma2], [1.0])
    r4, _ = orc.ek_residual(x, g.real, g.imag, eta, c, [sigma2 * 1.05], [1.0])
    assert np.max(np.abs(r3)) > 1e-3 and np.max(np.abs(r4)) > 1e-3


def test_ek_fit_gradients_match_central_differences(orc):
    rng = np.random.default_rng(4)
    c, eta = 0.25, 0.04
    x = np.linspace(0.2, 2.4, 31)
    g = np.array([_mp_stieltjes_quadrature(xx + 1j * eta, c, 1.0) for xx in x])
    g = g * (1.0 + 0.05 * rng.standard_normal(x.size))   # off the law: non-zero residual
    t = np.exp(np.linspace(np.log(0.2), np.log(5.0), 12))
    w = rng.uniform(0.2, 1.0, t.size)
    w /= w.sum()
    loss, ggr, ggi, gw, ge = orc.ek_fit(x, g.real, g.imag, eta, c, t, w)
    r, _ = orc.ek_residual(x, g.real, g.imag, eta, c, t, w)
    assert abs(loss - np.mean(np.abs(r) ** 2)) <= 1e-15 * max(loss, 1e-300) * 10

    def L(gr=g.real, gi=g.imag, ww=w, e=eta):
        return orc.ek_fit(x, gr, gi, e, c, t, ww)[0]

    h = 1e-6
    fd = lambda f: (f(h) - f(-h)) / (2 * h)
    fd_gr = np.array([fd(lambda s: L(gr=g.real + s * np.eye(x.size)[m])) for m in range(x.size)])
    fd_gi = np.array([fd(lambda s: L(gi=g.imag + s * np.eye(x.size)[m])) for m in range(x.size)])
    fd_w = np.array([fd(lambda s: L(ww=w + s * np.eye(t.size)[k])) for k in range(t.size)])
    fd_e = fd(lambda s: L(e=eta + s))
    assert _rel(ggr, fd_gr) < 1e-6
    assert _rel(ggi, fd_gi) < 1e-6
    assert _rel(gw, fd_w) < 1e-6
    assert abs(ge - fd_e) / abs(fd_e) < 1e-6


def _all_points_loss(orc, ev, eta, gre, gim):
    sr, si = orc.at_points(ev, ev, eta)
    return float(np.dot(gre, sr) + np.dot(gim, si))


def test_all_points_adjoint_matches_two_roles_and_fd(orc):
    """orc_adjoint_all_points at a moderate eta: the sum of the two roles of
    orc_adjoint_at_points at queries = eigenvalues (where leaving the self
    pair out or in differs only by rounding), central finite differences of
    the all-points loss with every eigenvalue moving in both roles, and the
    same eta gradient."""
    rng = np.random.default_rng(7)
    ev = np.sort(rng.uniform(0.5, 2.0, 60))
    eta = 0.05
    gre, gim = rng.standard_normal(60), rng.standard_normal(60)
    g, ge = orc.adjoint_all_points(ev, eta, gre, gim)
    gev, gq, ge2 = orc.adjoint_at_points(ev, ev, eta, gre, gim)
    assert _rel(g, gev + gq) <= 1e-13
    assert ge == ge2
    h = 1e-6
    fd = np.empty_like(ev)
    for i in range(ev.size):
        e1, e0 = ev.copy(), ev.copy()
        e1[i] += h
        e0[i] -= h
        fd[i] = (_all_points_loss(orc, e1, eta, gre, gim)
                 - _all_points_loss(orc, e0, eta, gre, gim)) / (2 * h)
    assert _rel(g, fd) < 1e-6
    # translation invariance of the all-points sums: the gradient sums to zero
    assert abs(g.sum()) <= 1e-12 * np.sum(np.abs(g))


def test_all_points_adjoint_exact_at_tiny_eta(orc):
    """At eta = 1e-10 the two roles' self terms are +-G_re/eta^2 = 1e20: the
    oracle's all-points gradient equals the exact rational sum over j != i
    (Python fractions on the same doubles) to 1e-15, while adding the two
    roles of orc_adjoint_at_points loses most of its digits to the rounding
    of 1e20 (the failure the all-points form avoids)."""
    from fractions import Fraction as F
    rng = np.random.default_rng(11)
    p = 24
    ev = np.sort(rng.uniform(0.5, 2.0, p))
    eta = 1e-10
    gre, gim = rng.standard_normal(p), rng.standard_normal(p)
    g, _ = orc.adjoint_all_points(ev, eta, gre, gim)
    fe = F(eta)
    exact = []
    for i in range(p):
        acc = F(0)
        for j in range(p):
            if j == i:
                continue
            d = F(ev[i]) - F(ev[j])
            den = d * d + fe * fe
            acc += ((F(gre[j]) - F(gre[i])) * (d * d - fe * fe) - 2 * fe * d * (F(gim[i]) + F(gim[j]))) / (den * den)
        exact.append(float(acc))
    exact = np.array(exact)
    assert _rel(g, exact) <= 1e-15
    gev, gq, _ = orc.adjoint_at_points(ev, ev, eta, gre, gim)
    assert _rel(gev + gq, exact) > 1e-9
