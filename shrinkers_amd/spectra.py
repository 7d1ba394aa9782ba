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
    """Inverse MP CDF by bisection on theta (64 halvings, f64)."""
    u = np.asarray(u, dtype=np.float64)
    lo = np.zeros_like(u)
    hi = np.full_like(u, math.pi)
    for _ in range(64):
        mid = 0.5 * (lo + hi)
        below = _mp_cdf_theta(mid, c) < u
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    theta = 0.5 * (lo + hi)
    m = 1.0 + c
    r = 2.0 * math.sqrt(c)
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
