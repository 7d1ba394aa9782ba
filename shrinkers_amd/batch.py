"""Many spectra in one persistent GPU launch (hipGraph), FP32-compensated.

Each spectrum's result is its own all-points transform, as
``compute_all_stieltjes(method="blocked", precision="f32c")`` would give it;
the batch runs them all in one captured launch with a work queue, and keeps
per-spectrum device timestamps for latency statistics.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import check
from .stieltjes import _stream_handle


class BatchedStieltjes:
    """All-points transforms of ``spectra`` (a list of 1-D arrays) with one
    eta per spectrum.  Eigenvalues are copied to the GPU once."""

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
        """Time the launch candidates on this batch and keep the fastest;
        returns ``(queries_per_lane, workgroups_per_cu, ms)``."""
        r, w, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(_lib.load().shk_batch_tune(self._h, self.re.data_ptr(), self.im.data_ptr(),
                                         _stream_handle(stream), ctypes.byref(r), ctypes.byref(w),
                                         ctypes.byref(ms)))
        return r.value, w.value, ms.value

    def set_config(self, queries_per_lane: int, workgroups_per_cu: int) -> None:
        check(_lib.load().shk_batch_set_config(self._h, int(queries_per_lane),
                                               int(workgroups_per_cu)))

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
