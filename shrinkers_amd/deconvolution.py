"""El Karoui / Marchenko-Pastur inversion on top of the GPU at-points kernel.

Mirrors ``spectral_deconvolution`` (src/deconvolution/mod.rs:116-259): a
uniform grid z_k = lambda_k + i*eta, the sample Stieltjes transform g(z_k)
from ``compute_stieltjes_at_points`` (GPU), then per grid point
a(z) = 1 - c - c z g(z), w = z / a(z), m_Sigma(w) = a(z) g(z) and the
population density Im(m_Sigma)/pi.  The grid algebra is O(n_points) and runs
on the host, as in the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .stieltjes import compute_stieltjes_at_points, default_eta


@dataclass
class DeconvolutionResult:
    """Field names of the reference's DeconvolutionResult (mod.rs:58-81)."""
    lambda_grid: np.ndarray
    density: np.ndarray
    w_re: np.ndarray
    sample_stieltjes_real: np.ndarray
    sample_stieltjes_imag: np.ndarray
    population_stieltjes_real: np.ndarray
    population_stieltjes_imag: np.ndarray


def deconvolution_grid(eigenvalues: np.ndarray, n_points: int,
                       lambda_min: Optional[float] = None,
                       lambda_max: Optional[float] = None) -> np.ndarray:
    """Grid over [min - 0.2 range, max + 0.2 range] clipped at 0 (mod.rs:141-165)."""
    min_ev = float(np.min(eigenvalues))
    max_ev = float(np.max(eigenvalues))
    rng = (max_ev - min_ev) if (max_ev - min_ev) > 0.0 else 1.0
    margin = 0.2 * rng
    lo = lambda_min if lambda_min is not None else max(min_ev - margin, 0.0)
    hi = lambda_max if lambda_max is not None else max_ev + margin
    step = (hi - lo) / (n_points - 1.0) if n_points > 1 else 0.0
    return np.array([lo + step * float(k) for k in range(n_points)], dtype=np.float64)


def el_karoui_inversion(lambda_grid: np.ndarray, raw_re: np.ndarray, raw_im: np.ndarray,
                        p: int, c: float, eta: float) -> DeconvolutionResult:
    """Per-grid-point MP inversion from RAW at-points sums (mod.rs:206-248)."""
    inv_p = 1.0 / float(p)
    n = lambda_grid.size
    out = {k: np.empty(n) for k in ("density", "w_re", "sr", "si", "pr", "pi")}
    for k in range(n):
        z_real = float(lambda_grid[k])
        z_imag = eta
        g_real = -float(raw_re[k]) * inv_p
        g_imag = float(raw_im[k]) * inv_p
        zg_real = z_real * g_real - z_imag * g_imag
        zg_imag = z_real * g_imag + z_imag * g_real
        a_real = 1.0 - c - c * zg_real
        a_imag = -c * zg_imag
        a_norm_sq = a_real * a_real + a_imag * a_imag
        w_real = (z_real * a_real + z_imag * a_imag) / a_norm_sq if a_norm_sq > 0.0 else z_real
        m_real = a_real * g_real - a_imag * g_imag
        m_imag = a_real * g_imag + a_imag * g_real
        out["density"][k] = m_imag / math.pi
        out["w_re"][k] = w_real
        out["sr"][k] = g_real
        out["si"][k] = g_imag
        out["pr"][k] = m_real
        out["pi"][k] = m_imag
    return DeconvolutionResult(lambda_grid, out["density"], out["w_re"], out["sr"], out["si"],
                               out["pr"], out["pi"])


def spectral_deconvolution(eigenvalues, c: float, n_points: int = 200,
                           eta: Optional[float] = None, lambda_min: Optional[float] = None,
                           lambda_max: Optional[float] = None, method: str = "blocked",
                           parallel: bool = False,
                           cutoff: Optional[float] = None) -> DeconvolutionResult:
    """El Karoui deconvolution with the at-points transform on the GPU.

    ``method`` defaults to ``"blocked"`` as RmtConfig::new does
    (src/config.rs:385-396); the auto presets resolve to ``chebcode_fast``.
    """
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    p = ev.size
    if p == 0:
        e = np.empty(0)
        return DeconvolutionResult(e, e, e, e, e, e, e)
    eta_val = eta if eta is not None else default_eta(p)
    grid = deconvolution_grid(ev, n_points, lambda_min, lambda_max)
    raw_re, raw_im = compute_stieltjes_at_points(grid, ev, eta_val, method, cutoff, parallel)
    return el_karoui_inversion(grid, raw_re, raw_im, p, c, eta_val)
