"""Many independent spectra of uneven size.

``SpectrumBatch`` / ``all_points_many_dev``: one library call transforms
every spectrum -- for the exact family as one segmented kernel launch over
all spectra (``n_streams=0``, the default), or as the single-spectrum kernels
spread over a few library streams (``n_streams`` 1..8, which can also return
per-spectrum device start/end times).  Each result equals the split-tile
single-spectrum kernel bit for bit (the all-points entry's own result outside
the range in which it takes the symmetric kernel, ``symmetric_range()``).

``BatchedStieltjes``: a fixed set of spectra kept on the device with its
spectrum table; the segmented launch is captured in a hipGraph and replayed
by each ``run``, with per-spectrum device times and a tuner that picks the
split-tile kernel's tile height or the symmetric kernel.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import check
from .stieltjes import _PRECISIONS, _require_gpu_method, _stream_handle, parse_method


def all_points_many_dev(ev, offsets, etas, method: str = "blocked", precision: str = "f32c",
                        scaled: bool = True, n_streams: int = 0, out=None, times: bool = False,
                        stream=None):
    """All-points transforms of the spectra packed back to back in the CUDA
    float64 tensor ``ev`` (``offsets``: n+1 int64 starts, ``etas``: n etas).
    Returns ``(re, im)`` packed the same way, plus an (n, 2) array of
    per-spectrum start/end milliseconds from the batch start when ``times``
    (which synchronises)."""
    import torch
    if not (ev.is_cuda and ev.dtype == torch.float64 and ev.dim() == 1 and ev.is_contiguous()):
        raise ValueError("ev must be a contiguous 1-D float64 CUDA tensor")
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    e = np.ascontiguousarray(etas, dtype=np.float64)
    if off.size != e.size + 1 or (off.size and off[-1] != ev.numel()):
        raise ValueError("offsets must have one more entry than etas and end at ev.numel()")
    mid = parse_method(method)
    _require_gpu_method(mid)
    re, im = out if out is not None else (torch.empty_like(ev), torch.empty_like(ev))
    t = np.zeros((e.size, 2), dtype=np.float32) if times else None
    check(_lib.load().shk_all_points_many_dev(
        ev.data_ptr(), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), e.size, _lib.dptr(e),
        mid, _PRECISIONS[precision], int(bool(scaled)), re.data_ptr(), im.data_ptr(),
        int(n_streams),
        t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if times else None,
        _stream_handle(stream)))
    return (re, im, t.astype(np.float64)) if times else (re, im)


class SpectrumBatch:
    """Spectra (a list of 1-D arrays) copied to the GPU once, back to back,
    with one eta each; ``run`` transforms them all in one library call."""

    def __init__(self, spectra: Sequence, etas: Sequence[float]):
        import torch
        arrays = [np.ascontiguousarray(s, dtype=np.float64) for s in spectra]
        if len(arrays) != len(etas):
            raise ValueError("one eta per spectrum")
        self.sizes = np.array([a.size for a in arrays], dtype=np.int64)
        self.offsets = np.zeros(len(arrays) + 1, dtype=np.int64)
        np.cumsum(self.sizes, out=self.offsets[1:])
        self.etas = np.ascontiguousarray(etas, dtype=np.float64)
        self.ev = torch.from_numpy(np.concatenate(arrays) if arrays else np.zeros(0)).cuda()
        self.re = torch.empty_like(self.ev)
        self.im = torch.empty_like(self.ev)

    def run(self, method: str = "blocked", precision: str = "f32c", scaled: bool = True,
            n_streams: int = 0, times: bool = False, stream=None):
        """Enqueue every transform; returns the per-spectrum (start, end) ms
        array when ``times``, else None."""
        res = all_points_many_dev(self.ev, self.offsets, self.etas, method, precision, scaled,
                                  n_streams, out=(self.re, self.im), times=times, stream=stream)
        return res[2] if times else None

    def results(self):
        """Per-spectrum (re, im) numpy arrays of the last run (synchronises)."""
        import torch
        torch.cuda.synchronize()
        re = self.re.cpu().numpy()
        im = self.im.cpu().numpy()
        return [(re[a:b], im[a:b]) for a, b in zip(self.offsets[:-1], self.offsets[1:])]


class BatchedStieltjes:
    """All-points transforms of ``spectra`` (a list of non-empty 1-D arrays)
    with one eta per spectrum, FP32-compensated.  Eigenvalues are copied to
    the GPU once; each ``run`` replays the captured launch."""

    def __init__(self, spectra: Sequence, etas: Sequence[float], scaled: bool = True):
        import torch
        arrays = [np.ascontiguousarray(s, dtype=np.float64) for s in spectra]
        if len(arrays) != len(etas):
            raise ValueError("one eta per spectrum")
        self.sizes = np.array([a.size for a in arrays], dtype=np.int64)
        self.offsets = np.zeros(len(arrays) + 1, dtype=np.int64)
        np.cumsum(self.sizes, out=self.offsets[1:])
        self.etas = np.ascontiguousarray(etas, dtype=np.float64)
        self.ev = torch.from_numpy(np.concatenate(arrays) if arrays else np.zeros(0)).cuda()
        self.re = torch.empty_like(self.ev)
        self.im = torch.empty_like(self.ev)
        h = ctypes.c_void_p()
        check(_lib.load().shk_batch_create(
            self.ev.data_ptr(), self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            len(arrays), _lib.dptr(self.etas), int(bool(scaled)), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().shk_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def tune(self, stream=None):
        """Time the configurations on this batch (1, 2, 4 queries per lane of
        the split-tile kernel; 8, the symmetric kernel, where accepted) and
        keep the fastest; returns ``(queries_per_lane, ms)``."""
        r, ms = ctypes.c_int(), ctypes.c_double()
        check(_lib.load().shk_batch_tune(self._h, self.re.data_ptr(), self.im.data_ptr(),
                                         _stream_handle(stream), ctypes.byref(r),
                                         ctypes.byref(ms)))
        return r.value, ms.value

    def set_config(self, queries_per_lane: int) -> None:
        """Configuration: 1, 2 (default: results equal the split-tile
        single-spectrum kernel's bit for bit) or 4 queries per lane of the
        split-tile kernel, or 8 for the symmetric kernel (each unordered pair
        once; results equal the single-spectrum symmetric transform's bit for
        bit)."""
        check(_lib.load().shk_batch_set_config(self._h, int(queries_per_lane)))

    def run(self, stream=None):
        """Enqueue the batch; returns the (re, im) CUDA tensors (concatenated)."""
        check(_lib.load().shk_batch_run(self._h, self.re.data_ptr(), self.im.data_ptr(),
                                        _stream_handle(stream)))
        return self.re, self.im

    def results(self):
        """Per-spectrum (re, im) numpy arrays of the last run (synchronises)."""
        import torch
        torch.cuda.synchronize()
        re = self.re.cpu().numpy()
        im = self.im.cpu().numpy()
        return [(re[a:b], im[a:b]) for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def spectrum_times_us(self):
        """Per-spectrum (start, end) in microseconds of the last run."""
        import torch
        torch.cuda.synchronize()
        n = self.sizes.size
        s0 = np.empty(n)
        s1 = np.empty(n)
        check(_lib.load().shk_batch_spectrum_times(self._h, _lib.dptr(s0), _lib.dptr(s1)))
        return s0, s1
