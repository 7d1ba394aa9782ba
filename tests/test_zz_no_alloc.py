"""The all-points entry that allocates nothing (shk_all_points_noalloc_dev),
through the C ABI, against the allocating entry (shk_all_points_dev) on the
same inputs: two calls in a row, the second on a smaller spectrum, in both
precisions.  Where both run the symmetric kernel (8 192 <= p for FP32-
compensated, 18 432 <= p for FP64) the results are the same bits (the band
height differs, the order of the additions does not); below, the allocating
entry runs the split-tile kernel and the bars are the exact family's: 1e-13
(FP64) and 1e-6 (FP32-compensated) normwise.  Sizes above the entry's
maximum (262 144) are refused before any work."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu



def _noalloc():
    from shrinkers_amd import _lib
    return _lib.load().shk_all_points_noalloc_dev


def _call(fn, d, eta, prec, scaled):
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd.stieltjes import _stream_handle
    re, im = torch.full_like(d, np.nan), torch.full_like(d, np.nan)
    st = fn(d.data_ptr(), d.numel(), float(eta), prec, int(scaled), re.data_ptr(), im.data_ptr(),
            _stream_handle(None))
    _lib.check(st)
    return re, im


@pytest.mark.parametrize("precision", ["f32c", "f64"])
@pytest.mark.parametrize("sizes", [(50_000, 20_000), (262_144, 3000), (9000, 700)])
def test_noalloc_matches_allocating_entry(shk, precision, sizes):
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    fn = _noalloc()
    prec = _lib.PREC_F32C if precision == "f32c" else _lib.PREC_F64
    sym_lo, _ = shk.symmetric_range(precision)
    for k, p in enumerate(sizes):
        print(f"no-alloc {precision} call {k}: p={p}", flush=True)
        ev = S.mp_inverse_cdf_spectrum(p, c=0.3, seed=k)
        eta = 1.0 / np.sqrt(p)
        d = torch.from_numpy(ev).cuda()
        got = _call(fn, d, eta, prec, scaled=True)
        want = shk.all_points_dev(d, eta, "blocked", precision=precision, scaled=True)
        torch.cuda.synchronize()
        if p >= sym_lo:
            assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1]), p
        else:
            tol = 1e-6 if precision == "f32c" else 1e-13
            g = [t.cpu().numpy() for t in got]
            w = [t.cpu().numpy() for t in want]
            assert rel_l2(g[0], g[1], w[0], w[1]) <= tol, p


def test_noalloc_raw_and_small_eta(shk, orc):
    """Raw sums at p = 5 000 against the oracle; an eta below the FP32-
    compensated range (1e-20) runs FP64 and still meets the FP64 bar."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    fn = _noalloc()
    ev = S.mp_inverse_cdf_spectrum(5000)
    d = torch.from_numpy(ev).cuda()
    for eta, prec, tol in ((1.0 / np.sqrt(5000), _lib.PREC_F32C, 1e-6),
                           (1.0 / np.sqrt(5000), _lib.PREC_F64, 1e-13),
                           (1e-20, _lib.PREC_F32C, 1e-13)):
        re, im = _call(fn, d, eta, prec, scaled=False)
        torch.cuda.synchronize()
        rr, ri = orc.tiled_parallel(ev, eta)
        assert rel_l2(re.cpu().numpy(), im.cpu().numpy(), rr, ri) <= tol, (eta, prec)


def test_noalloc_refuses_above_maximum(shk):
    import torch
    from shrinkers_amd import _lib
    fn = _noalloc()
    d = torch.zeros(262_145, dtype=torch.float64, device="cuda")
    out = torch.empty_like(d)
    st = fn(d.data_ptr(), d.numel(), 0.01, _lib.PREC_F64, 1, out.data_ptr(), out.data_ptr(), None)
    assert st == _lib.SHK_ERR_BAD_ARG
    assert "maximum" in _lib.last_error()
