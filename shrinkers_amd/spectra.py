"""Synthetic eigenvalue spectra for the measured workloads.

* ``mp_inverse_cdf_spectrum`` -- Marchenko-Pastur bulk drawn by inverse-CDF
  sampling of seeded uniforms, plus outliers at the forward-BBP images
  lambda_hat = l (1 + c/(l - 1)) of population spikes l
  (src/spiked/estimation.rs:8-11).  The MP CDF is evaluated in closed form
  on the angle parametrisation x = m - r cos(theta) of the support and
  inverted by bisection in f64.
* ``spiked_wishart_spectrum`` -- eigenvalues of X^T X / n for X = Z diag(sqrt(Sigma)).
* ``harness_spectrum`` / ``log_spectrum`` -- the reference's bit-reproducible
  benchmark spectra (benches/support/mod.rs:65-82, src/stieltjes/testutil.rs:10-12).
"""
from __future__ import annotations

import math

import numpy as np

_LCG_A = 6364136223846793005
_LCG_C = 1442695040888963407
_MASK = (1 << 64) - 1


def lcg_uniforms(seed: int, n: int) -> np.ndarray:
    """64-bit LCG uniforms in [0,1) (benches/support/mod.rs:17-35)."""
    out = np.empty(n, np.float64)
    s = seed & _MASK
    for k in range(n):
        s = (s * _LCG_A + _LCG_C) & _MASK
        out[k] = float(s >> 11) / float(1 << 53)
    return out


def harness_spectrum(p: int) -> np.ndarray:
    """Uniform bulk on [(1-sqrt c)^2, (1+sqrt c)^2], c = 0.5, plus two outliers."""
    c = 0.5
    lo = (1.0 - math.sqrt(c)) ** 2
    hi = (1.0 + math.sqrt(c)) ** 2
    if p < 3:
        return lo + lcg_uniforms(42, p) * (hi - lo)
    v = lo + lcg_uniforms(42, p - 2) * (hi - lo)
    v = np.concatenate([v, [hi * 2.3, lo * 0.35]])
    return np.sort(v)


def log_spectrum(p: int) -> np.ndarray:
    return np.log(np.arange(1, p + 1, dtype=np.float64))


def forward_bbp(ell: float, c: float, sigma2: float = 1.0) -> float:
    """Sample image of a population spike above the BBP threshold:
    sigma2 * l' (1 + c/(l' - 1)) with l' = l/sigma2 (src/spiked/estimation.rs:8-11)."""
    x = ell / sigma2
    return sigma2 * x * (1.0 + c / (x - 1.0))


def _mp_cdf_theta(theta: np.ndarray, c: float) -> np.ndarray:
    """MP CDF (sigma^2 = 1) at x = m - r cos(theta), m = 1 + c, r = 2 sqrt(c).

    F = r^2/(2 pi c) [ sin(theta)/r + m theta/r^2
                       + (1 - m^2/r^2) * 2/(1 - c) * atan(sqrt((m+r)/(m-r)) tan(theta/2)) ]
    from integrating sqrt((b-x)(x-a))/(2 pi c x) with the substitution above;
    F(0) = 0 and F(pi) = 1.  Requires 0 < c < 1.
    """
    m = 1.0 + c
    r = 2.0 * math.sqrt(c)
    k = math.sqrt((m + r) / (m - r))
    t = np.arctan(k * np.tan(0.5 * theta))
    t = np.where(theta >= math.pi, 0.5 * math.pi, t)
    val = np.sin(theta) / r + m * theta / r**2 + (1.0 - m**2 / r**2) * 2.0 / (1.0 - c) * t
    return r**2 / (2.0 * math.pi * c) * val


def mp_inverse_cdf(u: np.ndarray, c: float, sigma2: float = 1.0) -> np.ndarray:
    """Inverse MP CDF in theta: a bracket from a 4097-point table of F, then
    safeguarded Newton steps (F'(theta) = r^2 sin^2(theta) / (2 pi c x)),
    falling back to bisection whenever a step leaves the bracket."""
    u = np.asarray(u, dtype=np.float64)
    m = 1.0 + c
    r = 2.0 * math.sqrt(c)
    grid = np.linspace(0.0, math.pi, 4097)
    table = _mp_cdf_theta(grid, c)
    k = np.clip(np.searchsorted(table, u, side="right") - 1, 0, grid.size - 2)
    lo = grid[k]
    hi = grid[k + 1]
    theta = lo + (u - table[k]) / np.maximum(table[k + 1] - table[k], 1e-300) * (hi - lo)
    theta = np.clip(theta, lo, hi)
    act = np.arange(u.size)
    for _ in range(80):  # converged elements leave the active set
        t, a, b, uu = theta[act], lo[act], hi[act], u[act]
        f = _mp_cdf_theta(t, c) - uu
        a = np.where(f < 0.0, t, a)
        b = np.where(f < 0.0, b, t)
        x = m - r * np.cos(t)
        fp = r * r * np.sin(t) ** 2 / (2.0 * math.pi * c * x)
        with np.errstate(divide="ignore", invalid="ignore"):
            step = t - f / fp
        conv = (f == 0.0) | (np.abs(step - t) <= 4e-16 * np.maximum(t, 1e-300))
        inside = (step >= a) & (step <= b) & np.isfinite(step)
        tn = np.where(conv, t, np.where(inside, step, 0.5 * (a + b)))
        theta[act], lo[act], hi[act] = tn, a, b
        done = conv | (b - a <= 1e-15)
        act = act[~done]
        if act.size == 0:
            break
    return sigma2 * (m - r * np.cos(theta))


def mp_inverse_cdf_spectrum(p: int, c: float = 0.25, sigma2: float = 1.0, seed: int = 0,
                            spikes=(12.0, 7.0, 4.0)) -> np.ndarray:
    """MP(c, sigma2) bulk of p - len(spikes) inverse-CDF draws plus the
    forward-BBP images of ``spikes``, sorted ascending."""
    k = len(spikes)
    u = np.random.default_rng(seed).random(p - k)
    bulk = mp_inverse_cdf(u, c, sigma2)
    out = [forward_bbp(s, c, sigma2) for s in spikes]
    return np.sort(np.concatenate([bulk, np.asarray(out, np.float64)]))


def spiked_wishart_spectrum(n: int = 4096, p: int = 1024, spikes=(12.0, 7.0, 4.0),
                            sigma2: float = 1.0, seed: int = 0) -> np.ndarray:
    """Eigenvalues (ascending) of X^T X / n, X = Z diag(sqrt(Sigma)),
    Sigma = diag(spikes..., sigma2, ..., sigma2), Z i.i.d. N(0, 1)."""
    rng = np.random.default_rng(seed)
    sig = np.full(p, sigma2)
    sig[: len(spikes)] = spikes
    x = rng.standard_normal((n, p)) * np.sqrt(sig)[None, :]
    return np.linalg.eigvalsh(x.T @ x / n)


def batched_workload(n: int = 512, seed: int = 0, p_range=(500, 20000), c_range=(0.05, 0.5),
                     sigma2_range=(0.5, 2.0), max_spikes: int = 5, spike_range=(2.0, 20.0)):
    """Independent spectra of uneven size: p ~ U[p_range], c ~ U[c_range],
    sigma^2 log-uniform on sigma2_range, 0..max_spikes population spikes
    l ~ U[spike_range]*sigma^2, each placed at its forward-BBP image when
    above the detection threshold sigma^2 (1 + sqrt(c)) (otherwise it stays in
    the bulk).  Returns (list of sorted spectra, list of (p, c, sigma2, k))."""
    rng = np.random.default_rng(seed)
    spectra, meta = [], []
    for s in range(n):
        p = int(rng.integers(p_range[0], p_range[1] + 1))
        c = float(rng.uniform(*c_range))
        sigma2 = float(np.exp(rng.uniform(np.log(sigma2_range[0]), np.log(sigma2_range[1]))))
        k = int(rng.integers(0, max_spikes + 1))
        ells = rng.uniform(*spike_range, size=k) * sigma2
        outl = [forward_bbp(l, c, sigma2) for l in ells if l / sigma2 > 1.0 + math.sqrt(c)]
        bulk = mp_inverse_cdf(rng.random(p - len(outl)), c, sigma2)
        spectra.append(np.sort(np.concatenate([bulk, np.asarray(outl, np.float64)])))
        meta.append((p, c, sigma2, len(outl)))
    return spectra, meta
