"""Differentiable Stieltjes transform and El Karoui bulk fit (torch.autograd).

The reference differentiates the transform only with respect to the query
point (``stieltjes_with_deriv_sum``, src/stieltjes/autovec.rs:38-59, exposed
as ``stieltjes_transform_with_deriv``, src/python.rs:758-781).  Here the
exact at-points sums are a ``torch.autograd.Function`` whose backward runs
the library's adjoint kernels, so gradients flow to the eigenvalues, the
query points and eta; and the El Karoui bulk fit -- the residual of the
Marchenko-Pastur equation as written in src/deconvolution/mod.rs:6 for a
discrete population H = sum_k w_k delta_{t_k} -- is a second Function whose
loss and gradients come from one FP64 kernel pass.

All tensors are CUDA float64; kernels are enqueued on torch's current
stream.  There is no CPU implementation.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from ._lib import check
from .stieltjes import _PRECISIONS, _stream_handle, parse_method

_BLOCKED = parse_method("blocked")  # the exact all-points family


def _need_cuda_f64(t, name):
    import torch
    if not (torch.is_tensor(t) and t.is_cuda and t.dtype == torch.float64 and t.dim() == 1):
        raise ValueError(f"{name} must be a 1-D float64 CUDA tensor")


def _eta_value(eta) -> float:
    import torch
    v = float(eta.detach()) if torch.is_tensor(eta) else float(eta)
    if not (math.isfinite(v) and v > 0.0):
        raise ValueError(f"eta must be a positive finite number, got {v}")
    return v


def _prec(precision: str) -> int:
    if precision not in ("f64", "f32c"):
        raise ValueError(f"precision must be 'f64' or 'f32c', got {precision!r}")
    return _PRECISIONS[precision]


def _function_classes():
    import torch

    class StieltjesAtPoints(torch.autograd.Function):
        """Raw exact sums S_m = sum_j 1/((x_m - lam_j) - i eta) -> (Re S, Im S)."""

        @staticmethod
        def forward(ctx, queries, eigenvalues, eta, precision):
            q = queries.detach().contiguous()
            ev = eigenvalues.detach().contiguous()
            e = _eta_value(eta)
            prec = _prec(precision)
            re = torch.empty_like(q)
            im = torch.empty_like(q)
            check(_lib.load().shk_exact_at_points_dev(q.data_ptr(), q.numel(), ev.data_ptr(),
                                                      ev.numel(), e, prec, re.data_ptr(),
                                                      im.data_ptr(), _stream_handle(None)))
            ctx.save_for_backward(q, ev)
            ctx.eta = e
            ctx.prec = prec
            ctx.eta_tensor = torch.is_tensor(eta)
            ctx.eta_shape = eta.shape if torch.is_tensor(eta) else None
            return re, im

        @staticmethod
        def backward(ctx, g_re, g_im):
            q, ev = ctx.saved_tensors
            need_q, need_ev, need_eta = ctx.needs_input_grad[:3]
            need_eta = need_eta and ctx.eta_tensor
            if not (need_q or need_ev or need_eta):
                return None, None, None, None
            g_re = torch.zeros_like(q) if g_re is None else g_re.contiguous()
            g_im = torch.zeros_like(q) if g_im is None else g_im.contiguous()
            gq = torch.empty_like(q) if need_q else None
            gev = torch.empty_like(ev) if need_ev else None
            ge = torch.empty((), dtype=torch.float64, device=q.device) if need_eta else None
            ptr = lambda t: t.data_ptr() if t is not None else None
            check(_lib.load().shk_exact_adjoint_dev(
                q.data_ptr(), q.numel(), ev.data_ptr(), ev.numel(), ctx.eta, g_re.data_ptr(),
                g_im.data_ptr(), ctx.prec, ptr(gev), ptr(gq), ptr(ge), _stream_handle(None)))
            if ge is not None:
                ge = ge.reshape(ctx.eta_shape)
            return gq, gev, ge, None

    class StieltjesAllPoints(torch.autograd.Function):
        """Raw all-points sums S_i = sum_j 1/((lam_i - lam_j) - i eta) ->
        (Re S, Im S); the backward gives the total eigenvalue gradient (query
        and source roles) in one call."""

        @staticmethod
        def forward(ctx, eigenvalues, eta, precision):
            ev = eigenvalues.detach().contiguous()
            e = _eta_value(eta)
            prec = _prec(precision)
            re = torch.empty_like(ev)
            im = torch.empty_like(ev)
            check(_lib.load().shk_all_points_dev(ev.data_ptr(), ev.numel(), e, _BLOCKED, 0.0,
                                                 prec, 0, re.data_ptr(), im.data_ptr(),
                                                 _stream_handle(None)))
            ctx.save_for_backward(ev)
            ctx.eta = e
            ctx.prec = prec
            ctx.eta_tensor = torch.is_tensor(eta)
            ctx.eta_shape = eta.shape if torch.is_tensor(eta) else None
            return re, im

        @staticmethod
        def backward(ctx, g_re, g_im):
            (ev,) = ctx.saved_tensors
            need_ev, need_eta = ctx.needs_input_grad[:2]
            need_eta = need_eta and ctx.eta_tensor
            if not (need_ev or need_eta):
                return None, None, None
            g_re = torch.zeros_like(ev) if g_re is None else g_re.contiguous()
            g_im = torch.zeros_like(ev) if g_im is None else g_im.contiguous()
            gev = torch.empty_like(ev) if need_ev else None
            ge = torch.empty((), dtype=torch.float64, device=ev.device) if need_eta else None
            ptr = lambda t: t.data_ptr() if t is not None else None
            check(_lib.load().shk_exact_adjoint_all_points_dev(
                ev.data_ptr(), ev.numel(), ctx.eta, g_re.data_ptr(), g_im.data_ptr(), ctx.prec,
                ptr(gev), ptr(ge), _stream_handle(None)))
            if ge is not None:
                ge = ge.reshape(ctx.eta_shape)
            return gev, ge, None

    class ElKaroui(torch.autograd.Function):
        """Mean squared residual of the MP equation (FP64 kernel)."""

        @staticmethod
        def forward(ctx, g_re, g_im, x, eta, c, atoms, weights):
            tens = [t.detach().contiguous() for t in (g_re, g_im, x, atoms, weights)]
            gr, gi, xx, t, w = tens
            e = _eta_value(eta)
            m, k = xx.numel(), t.numel()
            if gr.numel() != m or gi.numel() != m or w.numel() != k:
                raise ValueError("g_re, g_im and x must have one value per grid point, "
                                 "weights one per atom")
            dev = xx.device
            loss = torch.empty((), dtype=torch.float64, device=dev)
            ggr, ggi = torch.empty_like(gr), torch.empty_like(gi)
            gw = torch.empty_like(w)
            ge = torch.empty((), dtype=torch.float64, device=dev)
            check(_lib.load().shk_el_karoui_fit_dev(
                xx.data_ptr(), gr.data_ptr(), gi.data_ptr(), m, e, float(c), t.data_ptr(),
                w.data_ptr(), k, loss.data_ptr(), ggr.data_ptr(), ggi.data_ptr(), gw.data_ptr(),
                ge.data_ptr(), _stream_handle(None)))
            ctx.save_for_backward(ggr, ggi, gw, ge)
            ctx.eta_tensor = torch.is_tensor(eta)
            ctx.eta_shape = eta.shape if torch.is_tensor(eta) else None
            return loss

        @staticmethod
        def backward(ctx, gl):
            ggr, ggi, gw, ge = ctx.saved_tensors
            ng = ctx.needs_input_grad
            out_ge = (gl * ge).reshape(ctx.eta_shape) if (ng[3] and ctx.eta_tensor) else None
            return (gl * ggr if ng[0] else None, gl * ggi if ng[1] else None, None, out_ge,
                    None, None, gl * gw if ng[6] else None)

    return StieltjesAtPoints, ElKaroui, StieltjesAllPoints


_CLASSES = None


def _classes():
    global _CLASSES
    if _CLASSES is None:
        _CLASSES = _function_classes()
    return _CLASSES


def stieltjes_at_points(queries, eigenvalues, eta, precision: str = "f32c"):
    """Raw exact sums at ``queries`` as ``(re, im)``, differentiable with
    respect to ``queries``, ``eigenvalues`` and ``eta`` (a float, or a 0-d
    CUDA float64 tensor for its gradient)."""
    _need_cuda_f64(queries, "queries")
    _need_cuda_f64(eigenvalues, "eigenvalues")
    return _classes()[0].apply(queries, eigenvalues, eta, precision)


def stieltjes_all_points(eigenvalues, eta, precision: str = "f32c", scaled: bool = True):
    """All-points transform (queries = eigenvalues, diagonal included),
    differentiable with respect to the eigenvalues and ``eta``; scaled by 1/p
    when ``scaled``.  The backward is one call of the all-points adjoint
    (each unordered pair once where the size allows)."""
    _need_cuda_f64(eigenvalues, "eigenvalues")
    re, im = _classes()[2].apply(eigenvalues, eta, precision)
    if scaled:
        p = eigenvalues.numel()
        return re / p, im / p
    return re, im


def sample_stieltjes(grid, eigenvalues, eta, precision: str = "f32c"):
    """g(z) = (1/p) sum_j 1/(lam_j - z) at z = grid + i eta, in the reference
    inversion's convention (Im g > 0): g = -conj(S)/p for the raw sums S
    (src/deconvolution/mod.rs:210-213)."""
    re, im = stieltjes_at_points(grid, eigenvalues, eta, precision)
    p = eigenvalues.numel()
    return -re / p, im / p


def el_karoui_residual_loss(g_re, g_im, grid, eta, c: float, atoms, weights):
    """(1/M) sum_m |g_m - sum_k w_k / (t_k a_m - z_m)|^2, a_m = 1 - c - c z_m g_m."""
    for t, nm in ((g_re, "g_re"), (g_im, "g_im"), (grid, "grid"), (atoms, "atoms"),
                  (weights, "weights")):
        _need_cuda_f64(t, nm)
    if not (math.isfinite(float(c)) and float(c) > 0.0):
        raise ValueError(f"c must be a positive finite number, got {c}")
    return _classes()[1].apply(g_re, g_im, grid, eta, float(c), atoms, weights)


def el_karoui_fit_loss(eigenvalues, eta, logits, atoms, c: float, grid,
                       precision: str = "f32c"):
    """El Karoui bulk-fit loss for population weights softmax(logits) on
    ``atoms``: the sample transform at ``grid + i eta`` (exact kernel in
    ``precision``) fed to the MP residual.  Differentiable with respect to
    ``eigenvalues``, ``eta`` (0-d tensor) and ``logits``."""
    import torch
    g_re, g_im = sample_stieltjes(grid, eigenvalues, eta, precision)
    w = torch.softmax(logits, dim=0)
    return el_karoui_residual_loss(g_re, g_im, grid, eta, c, atoms, w)


def el_karoui_fit_step(eigenvalues, eta: float, weights, atoms, c: float, grid,
                       precision: str = "f32c", stream=None):
    """One fit step in a single library call (no autograd, no host round
    trips): returns ``(loss, grad_eigenvalues, grad_eta, grad_weights)`` as
    CUDA float64 tensors (loss and grad_eta 0-d) for the loss of
    ``el_karoui_residual_loss`` at g = sample transform of ``eigenvalues`` at
    ``grid + i eta``.  ``eta`` is a Python float.  The weights' gradient
    chains to logits as w * (g_w - <w, g_w>) for softmax weights."""
    import torch
    for t, nm in ((eigenvalues, "eigenvalues"), (weights, "weights"), (atoms, "atoms"),
                  (grid, "grid")):
        _need_cuda_f64(t, nm)
    if weights.numel() != atoms.numel():
        raise ValueError("one weight per atom")
    ev = eigenvalues.detach().contiguous()
    x = grid.detach().contiguous()
    t_ = atoms.detach().contiguous()
    w = weights.detach().contiguous()
    dev = ev.device
    loss = torch.empty((), dtype=torch.float64, device=dev)
    g_ev = torch.empty_like(ev)
    g_eta = torch.empty((), dtype=torch.float64, device=dev)
    g_w = torch.empty_like(w)
    check(_lib.load().shk_el_karoui_fit_step_dev(
        ev.data_ptr(), ev.numel(), x.data_ptr(), x.numel(), _eta_value(eta), float(c),
        t_.data_ptr(), w.data_ptr(), t_.numel(), _prec(precision), loss.data_ptr(),
        g_ev.data_ptr(), g_eta.data_ptr(), g_w.data_ptr(), _stream_handle(stream)))
    return loss, g_ev, g_eta, g_w


def bulk_fit_problem(p: int = 10_000, c: float = 0.25, n_atoms: int = 256,
                     n_grid: int = 512, seed: int = 0):
    """The bulk-fit workload: MP(c, 1) eigenvalues by inverse-CDF sampling
    (no outliers), ``n_atoms`` atoms log-spaced on [0.2, 5], ``n_grid`` grid
    points spanning [lam_min, lam_max], eta = 1/sqrt(p).  Returns numpy
    arrays (eigenvalues, grid, atoms) and eta."""
    from .spectra import mp_inverse_cdf_spectrum
    ev = mp_inverse_cdf_spectrum(p, c=c, seed=seed, spikes=())
    grid = np.linspace(float(ev[0]), float(ev[-1]), n_grid)
    atoms = np.exp(np.linspace(math.log(0.2), math.log(5.0), n_atoms))
    return ev, grid, atoms, 1.0 / math.sqrt(p)
