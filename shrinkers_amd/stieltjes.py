"""Stieltjes-transform API backed by the gfx950 HIP library.

Mirrors the reference's method API (src/stieltjes/mod.rs) and its Python
bindings (src/python.rs:740-837): same function names, argument meaning,
defaults, return layout and error messages.  Every computation runs in
libshrinkers_hip.so on the GPU; there is no CPU fallback.

Conventions (src/stieltjes/term.rs:3-7, src/stieltjes/mod.rs:29-34):
``S_i = sum_j 1/((x_i - lam_j) - i*eta)``; ``compute_all_stieltjes`` and
``stieltjes_transform`` return ``S/p``; ``compute_stieltjes_at_points``,
``compute_stieltjes_scaled_ray``, ``stieltjes_transform_with_deriv`` and the
``ChebCodeBatch`` evaluations return raw sums.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, dptr, fptr

# ---- method vocabulary -------------------------------------------------------

#: Canonical method names (StieltjesMethod::name, src/config.rs:92-118).
METHODS = (
    "naive", "autovec", "blocked", "blocked_autovec", "blocked_tiled", "blocked_windowed",
    "blocked_hybrid", "adaptive", "fft5", "fft3", "fft2", "nufft", "fmm", "chebcode",
    "chebcode_fast", "chebcode_xtreme", "chebcode_balanced", "ewald", "dst", "auto", "hodlr",
    "speed_auto", "accuracy_auto",
)

#: Methods with a GPU implementation here; the others raise NotImplementedError.
GPU_METHODS = frozenset({
    "naive", "autovec", "blocked", "blocked_autovec", "blocked_tiled", "blocked_windowed",
    "blocked_hybrid", "chebcode", "chebcode_fast", "chebcode_xtreme", "chebcode_balanced",
    "auto", "speed_auto", "accuracy_auto",
})

#: ChebPreset constants (theta, n, leaf_cap, f32 far field):
#: src/stieltjes/chebcode.rs:942-973.
CHEB_PRESETS = {
    "chebcode": (0.5, 11, 32, False),
    "chebcode_fast": (1.0, 8, 32, True),
    "chebcode_xtreme": (0.25, 11, 16, False),
    "chebcode_balanced": (0.55, 11, 32, False),
}

#: Additional operating point with no reference counterpart: the DEFAULT
#: opening angle with interpolation order 12 and an f64 far field.
EXTRA_CHEB_PRESETS = {
    "chebcode_n12": (0.5, 12, 32, False),
}

_PRECISIONS = {"f64": _lib.PREC_F64, "f32c": _lib.PREC_F32C, "f32": _lib.PREC_F32}


def parse_method(method: str) -> int:
    """Method name or alias -> id (src/python.rs:321-350); ValueError if unknown."""
    if not isinstance(method, str):
        raise TypeError(f"method must be a str, got {type(method).__name__}")
    out = ctypes.c_int()
    st = _lib.load().shk_method_from_name(method.encode(), ctypes.byref(out))
    if st != _lib.SHK_OK:
        raise ValueError(f"unknown method '{method}'")
    return out.value


def method_name(method_id: int) -> str:
    name = _lib.load().shk_method_name(int(method_id))
    if not name:
        raise ValueError(f"invalid method id {method_id}")
    return name.decode()


def default_eta(p: int) -> float:
    """eta = 0.1/sqrt(p): src/stieltjes/mod.rs:105-108."""
    return 0.1 / math.sqrt(p)


def symmetric_range(precision: str = "f32c"):
    """``(min_p, max_p)``: the eigenvalue counts for which the all-points
    entries take the symmetric kernel (each unordered pair evaluated once)."""
    lo, hi = ctypes.c_size_t(), ctypes.c_size_t()
    check(_lib.load().shk_symmetric_range(_PRECISIONS[precision], ctypes.byref(lo),
                                          ctypes.byref(hi)))
    return lo.value, hi.value


def default_eta_bulk(p: int) -> float:
    """eta = 0.4/sqrt(p): src/stieltjes/mod.rs:149-152."""
    return 0.4 / math.sqrt(p)


# ---- boundary validation (src/python.rs:138-318) -----------------------------

def _owned_f64_vec(array, what: str) -> np.ndarray:
    if not isinstance(array, np.ndarray) or array.dtype != np.float64 or array.ndim != 1:
        raise TypeError(f"'{what}' must be a 1-D numpy.ndarray of float64")
    if not array.flags["C_CONTIGUOUS"]:
        raise ValueError(f"{what} must be contiguous")
    return np.array(array, dtype=np.float64, copy=True)


def _require_finite(v: np.ndarray, what: str) -> None:
    if not np.all(np.isfinite(v)):
        raise ValueError(f"{what} must be finite (found NaN or infinity)")


def _inferred_f64(value, what: str) -> Optional[float]:
    """float, None or 'inferred' (InferredF64, src/python.rs:94-130)."""
    if value is None:
        return None
    if isinstance(value, str):
        if value == "inferred":
            return None
        raise ValueError("expected a float, None, or the string 'inferred'")
    try:
        return float(value)
    except (TypeError, ValueError):
        raise ValueError("expected a float, None, or the string 'inferred'") from None


def _require_positive_finite(value: float, what: str) -> float:
    if not (math.isfinite(value) and value > 0.0):
        raise ValueError(f"{what} must be a positive finite number, got {value}")
    return value


def _owned_spectrum_and_eta(eigenvalues, eta):
    ev = _owned_f64_vec(eigenvalues, "eigenvalues")
    _require_finite(ev, "eigenvalues")
    if ev.size == 0:
        raise ValueError("eigenvalues must be non-empty")
    v = _inferred_f64(eta, "eta")
    if v is None:
        v = default_eta(ev.size)
    return ev, _require_positive_finite(v, "eta")


def _parse_parallel(parallel) -> bool:
    if parallel is None:
        return False  # Parallelism::Auto resolves to sequential (src/config.rs:222-226)
    if isinstance(parallel, (bool, np.bool_)):
        return bool(parallel)
    raise TypeError(f"parallel must be a bool or None, got {type(parallel).__name__}")


def _cutoff_ratio(cutoff) -> float:
    v = _inferred_f64(cutoff, "cutoff")
    if v is None:
        return 0.0  # CutoffConfig::Disabled
    return _require_positive_finite(v, "cutoff ratio")


def _require_gpu_method(mid: int) -> None:
    name = method_name(mid)
    if name not in GPU_METHODS:
        raise NotImplementedError(
            f"method '{name}' has no GPU implementation in shrinkers_amd "
            "(available: the exact family and the chebcode presets)")


# ---- dispatcher mirrors --------------------------------------------------------

def compute_all_stieltjes(eigenvalues, eta: float, method: str = "blocked",
                          cutoff: Optional[float] = None, parallel: bool = False,
                          precision: str = "f64"):
    """Transform at every eigenvalue, scaled by 1/p (src/stieltjes/mod.rs:601-790).

    Returns ``(re, im)`` float64 arrays.  ``precision`` is ``"f64"`` (FP64
    kernel) or ``"f32c"`` (FP32-compensated kernel, ~1e-7 normwise).
    """
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    p = ev.size
    re = np.empty(p)
    im = np.empty(p)
    if p == 0:
        return re, im
    mid = parse_method(method)
    _require_gpu_method(mid)
    prec = _PRECISIONS[precision]
    check(_lib.load().shk_compute_all_stieltjes(dptr(ev), p, float(eta), mid,
                                                float(cutoff or 0.0), int(bool(parallel)), prec,
                                                dptr(re), dptr(im)))
    return re, im


def compute_all_stieltjes_f32(eigenvalues, eta: float, method: str = "blocked",
                              cutoff: Optional[float] = None):
    """Single-precision entry (src/stieltjes/mod.rs:805-851): float32 in/out, scaled."""
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float32)
    p = ev.size
    re = np.empty(p, np.float32)
    im = np.empty(p, np.float32)
    if p == 0:
        return re, im
    mid = parse_method(method)
    _require_gpu_method(mid)
    check(_lib.load().shk_compute_all_stieltjes_f32(fptr(ev), p, float(np.float32(eta)), mid,
                                                    float(cutoff or 0.0), fptr(re), fptr(im)))
    return re, im


def compute_stieltjes_at_points(query_points, eigenvalues, eta: float, method: str = "blocked",
                                cutoff: Optional[float] = None, parallel: bool = False):
    """RAW sums at arbitrary query points (src/stieltjes/mod.rs:394-494)."""
    q = np.ascontiguousarray(query_points, dtype=np.float64)
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    re = np.empty(q.size)
    im = np.empty(q.size)
    if q.size == 0:
        return re, im
    mid = parse_method(method)
    _require_gpu_method(mid)
    check(_lib.load().shk_compute_stieltjes_at_points(
        dptr(q), q.size, dptr(ev), ev.size, float(eta), mid, float(cutoff or 0.0),
        int(bool(parallel)), dptr(re), dptr(im)))
    return re, im


def compute_stieltjes_scaled_ray(eigenvalues, scale: float):
    """RAW sums at z_i = lam_i (1 + i*scale) (src/stieltjes/mod.rs:304-319)."""
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    re = np.empty(ev.size)
    im = np.empty(ev.size)
    if ev.size == 0:
        return re, im
    check(_lib.load().shk_compute_stieltjes_scaled_ray(dptr(ev), ev.size, float(scale), dptr(re),
                                                       dptr(im)))
    return re, im


def compute_all_stieltjes_with_deriv(eigenvalues, eta: float):
    """RAW values and d/dx (src/stieltjes/autovec.rs:63-76)."""
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    outs = [np.empty(ev.size) for _ in range(4)]
    if ev.size == 0:
        return tuple(outs)
    check(_lib.load().shk_compute_all_stieltjes_with_deriv(dptr(ev), ev.size, float(eta),
                                                           *[dptr(o) for o in outs]))
    return tuple(outs)


# ---- Python binding mirrors (src/python.rs:740-837) --------------------------

def stieltjes_transform(eigenvalues, eta="inferred", method: str = "blocked",
                        precision: str = "f64", cutoff="inferred", parallel=False) -> dict:
    """Empirical Stieltjes transform at every eigenvalue, scaled by 1/p.

    Same contract as the reference binding (src/python.rs:783-837): returns
    ``{"real", "imag"}`` (float64, or float32 for ``precision="f32"``) with
    ``imag > 0``.  ``precision="f32c"`` (an addition) selects the
    FP32-compensated kernel and returns float64.
    """
    ev, eta_val = _owned_spectrum_and_eta(eigenvalues, eta)
    mid = parse_method(method)
    par = _parse_parallel(parallel)
    cut = _cutoff_ratio(cutoff)
    if precision not in ("f64", "f32", "f32c"):
        raise ValueError(f"unknown precision '{precision}' (expected 'f64' or 'f32')")
    _require_gpu_method(mid)
    lib = _lib.load()
    p = ev.size
    if precision == "f32":
        ev32 = ev.astype(np.float32)
        re = np.empty(p, np.float32)
        im = np.empty(p, np.float32)
        check(lib.shk_compute_all_stieltjes_f32(fptr(ev32), p, float(np.float32(eta_val)), mid,
                                                cut, fptr(re), fptr(im)))
        return {"real": re, "imag": im}
    re = np.empty(p)
    im = np.empty(p)
    check(lib.shk_compute_all_stieltjes(dptr(ev), p, eta_val, mid, cut, int(par),
                                        _PRECISIONS[precision], dptr(re), dptr(im)))
    return {"real": re, "imag": im}


def stieltjes_transform_with_deriv(eigenvalues, eta="inferred") -> dict:
    """Values and analytic derivative at every eigenvalue (src/python.rs:758-781).

    As in the reference code path (compute_all_stieltjes_with_deriv,
    src/stieltjes/autovec.rs:63-76) the four arrays are raw sums.
    """
    ev, eta_val = _owned_spectrum_and_eta(eigenvalues, eta)
    vr, vi, dr, di = compute_all_stieltjes_with_deriv(ev, eta_val)
    return {"real": vr, "imag": vi, "deriv_real": dr, "deriv_imag": di}


# ---- ChebCodeBatch (src/stieltjes/chebcode.rs:1148-1362) ---------------------

class ChebCodeBatch:
    """A ChebCode tree built once on the GPU and evaluated many times."""

    def __init__(self, handle, p: int, f32_far: bool):
        self._h = handle
        self.p = p
        self.f32_far = f32_far

    @classmethod
    def build_mode(cls, eigenvalues, theta: float, n: int, leaf_cap: int,
                   f32_far: bool = False) -> "ChebCodeBatch":
        ev = np.ascontiguousarray(eigenvalues, dtype=np.float64)
        h = ctypes.c_void_p()
        check(_lib.load().shk_chebcode_build(dptr(ev), ev.size, float(theta), int(n),
                                             int(leaf_cap), int(bool(f32_far)), ctypes.byref(h)))
        return cls(h, ev.size, bool(f32_far))

    @classmethod
    def build(cls, eigenvalues, theta: float, n: int, leaf_cap: int) -> "ChebCodeBatch":
        return cls.build_mode(eigenvalues, theta, n, leaf_cap, False)

    @classmethod
    def build_preset(cls, eigenvalues, preset: str) -> "ChebCodeBatch":
        theta, n, leaf, f32 = {**CHEB_PRESETS, **EXTRA_CHEB_PRESETS}[preset]
        return cls.build_mode(eigenvalues, theta, n, leaf, f32)

    @classmethod
    def build_dev(cls, ev, theta: float, n: int, leaf_cap: int, f32_far: bool = False,
                  stream=None) -> "ChebCodeBatch":
        """Build from a CUDA float64 tensor on ``stream`` (default: torch's current)."""
        import torch
        if not (ev.is_cuda and ev.dtype == torch.float64 and ev.dim() == 1 and ev.is_contiguous()):
            raise ValueError("ev must be a contiguous 1-D float64 CUDA tensor")
        h = ctypes.c_void_p()
        check(_lib.load().shk_chebcode_build_dev(ev.data_ptr(), ev.numel(), float(theta), int(n),
                                                 int(leaf_cap), int(bool(f32_far)),
                                                 _stream_handle(stream), ctypes.byref(h)))
        return cls(h, ev.numel(), bool(f32_far))

    def evaluate_dev(self, eta: float, scale: float = 1.0, out=None, stream=None):
        """Evaluate at the tree's own eigenvalues (build order) into CUDA
        float64 tensors, enqueued on ``stream``; returns ``(re, im)``."""
        import torch
        if out is None:
            out = (torch.empty(self.p, dtype=torch.float64, device="cuda"),
                   torch.empty(self.p, dtype=torch.float64, device="cuda"))
        check(_lib.load().shk_chebcode_evaluate_dev(self._h, float(eta), float(scale),
                                                    out[0].data_ptr(), out[1].data_ptr(),
                                                    _stream_handle(stream)))
        return out

    def evaluate_local_dev(self, eta: float, scale: float = 1.0, out=None, stream=None):
        """Evaluate at the tree's own eigenvalues (build order) with local
        expansions: M2L into the Chebyshev points of target nodes no wider than
        eta/4 and L2P per query, instead of the per-query panel sums.  f64
        trees of order 11 or 12.  Returns CUDA float64 ``(re, im)``."""
        import torch
        if out is None:
            out = (torch.empty(self.p, dtype=torch.float64, device="cuda"),
                   torch.empty(self.p, dtype=torch.float64, device="cuda"))
        check(_lib.load().shk_chebcode_evaluate_local_dev(self._h, float(eta), float(scale),
                                                          out[0].data_ptr(), out[1].data_ptr(),
                                                          _stream_handle(stream)))
        return out

    def evaluate_many_dev(self, etas, out=None, stream=None):
        """evaluate_many into CUDA float64 tensors of shape (len(etas), p),
        enqueued on ``stream``; two etas per traversal.  Returns ``(re, im)``."""
        import torch
        e = np.ascontiguousarray(etas, dtype=np.float64)
        if out is None:
            out = (torch.empty((e.size, self.p), dtype=torch.float64, device="cuda"),
                   torch.empty((e.size, self.p), dtype=torch.float64, device="cuda"))
        if e.size:
            check(_lib.load().shk_chebcode_evaluate_many_dev(
                self._h, _lib.dptr(e), e.size, out[0].data_ptr(), out[1].data_ptr(),
                _stream_handle(stream)))
        return out

    def count_work(self, eta: float, query=None, stream=None):
        """``(near_pairs, far_terms)`` of one traversal pass: the summed sizes
        of the leaves each query visits and n per accepted panel, with the
        evaluation's node decisions.  ``query`` is a CUDA float64 tensor of
        points, or None for the tree's own eigenvalues.  Synchronises."""
        import torch
        near, far = ctypes.c_int64(), ctypes.c_int64()
        if query is not None:
            if not (query.is_cuda and query.dtype == torch.float64 and query.dim() == 1
                    and query.is_contiguous()):
                raise ValueError("query must be a contiguous 1-D float64 CUDA tensor")
            qp, nq = query.data_ptr(), query.numel()
        else:
            qp, nq = None, 0
        check(_lib.load().shk_chebcode_count_work(self._h, qp, nq, float(eta), ctypes.byref(near),
                                                  ctypes.byref(far), _stream_handle(stream)))
        return near.value, far.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().shk_chebcode_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def node_count(self) -> int:
        n = ctypes.c_size_t()
        check(_lib.load().shk_chebcode_info(self._h, ctypes.byref(n), None))
        return n.value

    def export(self, n: int) -> dict:
        """Pre-order tree arrays (for tests)."""
        c = self.node_count
        out = dict(lo=np.empty(c), hi=np.empty(c), lo_idx=np.empty(c, np.int64),
                   hi_idx=np.empty(c, np.int64), right=np.empty(c, np.int32),
                   nodes=np.empty(c * n), w=np.empty(c * n))
        check(_lib.load().shk_chebcode_export(
            self._h, dptr(out["lo"]), dptr(out["hi"]),
            out["lo_idx"].ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            out["hi_idx"].ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            out["right"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            dptr(out["nodes"]), dptr(out["w"])))
        return out

    def evaluate(self, eta: float):
        re = np.empty(self.p)
        im = np.empty(self.p)
        check(_lib.load().shk_chebcode_evaluate(self._h, float(eta), dptr(re), dptr(im)))
        return re, im

    def evaluate_points(self, points, eta: float, parallel: bool = False):
        q = np.ascontiguousarray(points, dtype=np.float64)
        re = np.empty(q.size)
        im = np.empty(q.size)
        if q.size:
            check(_lib.load().shk_chebcode_evaluate_points(self._h, dptr(q), q.size, float(eta),
                                                           dptr(re), dptr(im)))
        return re, im

    def evaluate_many(self, etas):
        e = np.ascontiguousarray(etas, dtype=np.float64)
        re = np.empty((e.size, self.p))
        im = np.empty((e.size, self.p))
        if e.size:
            check(_lib.load().shk_chebcode_evaluate_many(self._h, dptr(e), e.size, dptr(re),
                                                         dptr(im)))
        return re, im


# ---- device-resident entry points (torch tensors) -----------------------------

def _stream_handle(stream) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def all_points_dev(ev, eta: float, method: str = "blocked", precision: str = "f64",
                   scaled: bool = True, cutoff: Optional[float] = None, out=None, stream=None):
    """Transform of a CUDA float64 tensor, enqueued on ``stream`` (default:
    torch's current stream).  Returns ``(re, im)`` float64 tensors."""
    import torch
    if not (ev.is_cuda and ev.dtype == torch.float64 and ev.dim() == 1 and ev.is_contiguous()):
        raise ValueError("ev must be a contiguous 1-D float64 CUDA tensor")
    p = ev.numel()
    re, im = out if out is not None else (torch.empty_like(ev), torch.empty_like(ev))
    mid = parse_method(method)
    _require_gpu_method(mid)
    check(_lib.load().shk_all_points_dev(ev.data_ptr(), p, float(eta), mid,
                                         float(cutoff or 0.0), _PRECISIONS[precision],
                                         int(bool(scaled)), re.data_ptr(), im.data_ptr(),
                                         _stream_handle(stream)))
    return re, im


def at_points_dev(query, ev, eta: float, method: str = "blocked",
                  cutoff: Optional[float] = None, out=None, stream=None):
    """Raw sums at CUDA float64 query points (see compute_stieltjes_at_points)."""
    import torch
    for t, nm in ((query, "query"), (ev, "ev")):
        if not (t.is_cuda and t.dtype == torch.float64 and t.dim() == 1 and t.is_contiguous()):
            raise ValueError(f"{nm} must be a contiguous 1-D float64 CUDA tensor")
    re, im = out if out is not None else (torch.empty_like(query), torch.empty_like(query))
    mid = parse_method(method)
    _require_gpu_method(mid)
    check(_lib.load().shk_at_points_dev(query.data_ptr(), query.numel(), ev.data_ptr(),
                                        ev.numel(), float(eta), mid, float(cutoff or 0.0),
                                        re.data_ptr(), im.data_ptr(), _stream_handle(stream)))
    return re, im


# ---- ChebCode one-shot transforms in a caller's workspace -----------------------

def chebcode_workspace_bytes(p: int, method: str = "chebcode_fast") -> int:
    """Device bytes the workspace entries need for p eigenvalues."""
    n = ctypes.c_size_t()
    check(_lib.load().shk_chebcode_workspace_bytes(int(p), parse_method(method), ctypes.byref(n)))
    return n.value


def _workspace(ws):
    import torch
    if not (ws.is_cuda and ws.is_contiguous()):
        raise ValueError("the workspace must be a contiguous CUDA tensor")
    return ws.data_ptr(), ws.numel() * ws.element_size()


def chebcode_all_points_ws_dev(ev, eta: float, method: str, workspace, scaled: bool = True,
                               out=None, stream=None):
    """ChebCode transform of ascending CUDA float64 eigenvalues, built and
    evaluated in ``workspace`` (a CUDA tensor of at least
    chebcode_workspace_bytes bytes) with no host wait and no allocation, so a
    HIP graph can capture it.  Equal bit for bit to all_points_dev with the
    same method; NaN outputs on input that is not ascending."""
    import torch
    if not (ev.is_cuda and ev.dtype == torch.float64 and ev.dim() == 1 and ev.is_contiguous()):
        raise ValueError("ev must be a contiguous 1-D float64 CUDA tensor")
    wp, wb = _workspace(workspace)
    re, im = out if out is not None else (torch.empty_like(ev), torch.empty_like(ev))
    check(_lib.load().shk_chebcode_all_points_ws_dev(
        ev.data_ptr(), ev.numel(), float(eta), parse_method(method), int(bool(scaled)), wp, wb,
        re.data_ptr(), im.data_ptr(), _stream_handle(stream)))
    return re, im


def chebcode_at_points_ws_dev(query, ev, eta: float, method: str, workspace, out=None,
                              stream=None):
    """Raw ChebCode sums at CUDA float64 query points (the at_points_dev tree),
    built in ``workspace`` like chebcode_all_points_ws_dev."""
    import torch
    for t, nm in ((query, "query"), (ev, "ev")):
        if not (t.is_cuda and t.dtype == torch.float64 and t.dim() == 1 and t.is_contiguous()):
            raise ValueError(f"{nm} must be a contiguous 1-D float64 CUDA tensor")
    wp, wb = _workspace(workspace)
    re, im = out if out is not None else (torch.empty_like(query), torch.empty_like(query))
    check(_lib.load().shk_chebcode_at_points_ws_dev(
        query.data_ptr(), query.numel(), ev.data_ptr(), ev.numel(), float(eta),
        parse_method(method), wp, wb, re.data_ptr(), im.data_ptr(), _stream_handle(stream)))
    return re, im


def chebcode_workspace_status(workspace, stream=None) -> bool:
    """Whether the last workspace call found its eigenvalues ascending.
    Waits for the stream."""
    wp, _ = _workspace(workspace)
    ok = ctypes.c_int()
    check(_lib.load().shk_chebcode_workspace_status(wp, _stream_handle(stream), ctypes.byref(ok)))
    return bool(ok.value)


# ---- measurement hooks ----------------------------------------------------------

def profile_enable(on: bool = True) -> None:
    check(_lib.load().shk_profile_enable(int(bool(on))))


def profile_reset() -> None:
    check(_lib.load().shk_profile_reset())


def profile_read():
    """(total kernel milliseconds, launches) of the dominant kernels since reset."""
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(_lib.load().shk_profile_read(ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value
