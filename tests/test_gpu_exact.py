"""GPU parity of the exact O(p^2) family against the CPU oracle.

Bars: FP64 kernel within 1e-13 (normwise and per element against ||S||_inf)
of the oracle's restatement of the reference tiled kernel; FP32-compensated
kernel within 1e-6 normwise of the FP64 oracle; golden master at 1e-14.
"""
import json
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


def spectra(p):
    from shrinkers_amd import spectra as S
    out = {"log": S.log_spectrum(p), "harness": S.harness_spectrum(p)}
    if p >= 8:
        out["mp"] = S.mp_inverse_cdf_spectrum(p)
    return out


def assert_close(got_re, got_im, ref_re, ref_im, tol, what=""):
    err = rel_l2(got_re, got_im, ref_re, ref_im)
    scale = max(np.max(np.abs(ref_re)), np.max(np.abs(ref_im)))
    emax = max(np.max(np.abs(got_re - ref_re)), np.max(np.abs(got_im - ref_im))) / scale
    print(f"{what}: rel-L2 {err:.3e}  max/||S||inf {emax:.3e}", flush=True)
    assert err <= tol, f"{what}: rel-L2 {err:.3e} > {tol:.1e}"
    assert emax <= tol, f"{what}: max err {emax:.3e} > {tol:.1e}"


@pytest.mark.parametrize("p", [1, 2, 3, 7, 50, 64, 65, 300, 1000, 1024, 2001, 2002, 2003, 4001])
def test_f64_blocked_matches_oracle(shk, orc, p):
    for name, ev in spectra(p).items():
        eta = 1.0 / np.sqrt(p)
        r, i = shk.compute_all_stieltjes(ev, eta, "blocked")
        rr, ri = orc.tiled_parallel(ev, eta)
        assert_close(r, i, rr / p, ri / p, 1e-13, f"blocked f64 {name} p={p}")


def test_p1_is_pure_diagonal(shk):
    # cacheblock.rs:1604-1607: only the diagonal 1/(0 - i eta) = i/eta survives.
    r, i = shk.compute_all_stieltjes(np.array([2.0]), 0.5, "blocked")
    assert r[0] == 0.0
    assert abs(i[0] - 2.0) <= 4e-16


def test_all_exact_methods_agree(shk, orc):
    # mod.rs:859-913: every exact method vs naive/p, 1e-13 absolute.
    evals = np.log1p((np.arange(50) + 0.5) * 0.2)
    eta = 0.05
    ref = [orc.naive_sum(x, evals, eta) for x in evals]
    ref_r = np.array([a for a, _ in ref]) / 50
    ref_i = np.array([b for _, b in ref]) / 50
    for m in ("naive", "autovec", "blocked", "blocked_autovec", "blocked_tiled",
              "blocked_windowed", "blocked_hybrid", "accuracy_auto"):
        r, i = shk.compute_all_stieltjes(evals, eta, m)
        assert np.max(np.abs(r - ref_r)) < 1e-13, m
        assert np.max(np.abs(i - ref_i)) < 1e-13, m


def test_unsorted_input(shk, orc):
    # cacheblock.rs:1580-1601
    evals = np.array([3.0, -1.5, 2.25, 0.7, 4.2, -0.3, 1.1])
    eta = 0.37
    r, i = shk.compute_all_stieltjes(evals, eta, "blocked")
    for k, x in enumerate(evals):
        a, b = orc.autovec_sum(x, evals, eta)
        assert abs(r[k] * 7 - a) < 1e-12 and abs(i[k] * 7 - b) < 1e-12


def test_stieltjes_transform_binding(shk, orc):
    rng = np.random.default_rng(0)
    ev = np.sort(0.5 + rng.uniform(0, 1.5, 200))
    res = shk.stieltjes_transform(ev)
    assert set(res) == {"real", "imag"}
    assert np.all(res["imag"] > 0)
    eta = 0.1 / np.sqrt(200)
    rr, ri = orc.tiled_parallel(ev, eta)
    assert_close(res["real"], res["imag"], rr / 200, ri / 200, 1e-13, "binding")
    a = shk.stieltjes_transform(ev[:80].copy(), method="naive")
    b = shk.stieltjes_transform(ev[:80].copy(), method="blocked_tiled")
    np.testing.assert_allclose(a["real"], b["real"], atol=1e-11)
    f32 = shk.stieltjes_transform(ev[:100].copy(), precision="f32")
    f64 = shk.stieltjes_transform(ev[:100].copy(), precision="f64")
    assert f32["real"].dtype == np.float32
    np.testing.assert_allclose(f32["real"], f64["real"], rtol=5e-2)
    serial = shk.stieltjes_transform(ev, parallel=False)
    multi = shk.stieltjes_transform(ev, parallel=True)
    np.testing.assert_allclose(serial["real"], multi["real"], atol=1e-12)


@pytest.mark.parametrize("eta_scale", [1.0, 0.1])
@pytest.mark.parametrize("p", [1000, 4001, 20000])
def test_f32c_within_1e6(shk, orc, p, eta_scale):
    for name, ev in spectra(p).items():
        eta = eta_scale / np.sqrt(p)
        r, i = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
        rr, ri = orc.tiled_parallel(ev, eta)
        err = rel_l2(r, i, rr / p, ri / p)
        print(f"f32c {name} p={p} eta={eta_scale}/sqrt(p): rel-L2 {err:.3e}", flush=True)
        assert err <= 1e-6


def test_f32c_unsorted(shk, orc):
    p = 3000
    from shrinkers_amd import spectra as S
    ev = orc.shuffled(S.mp_inverse_cdf_spectrum(p), 11)
    eta = 1.0 / np.sqrt(p)
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
    rr, ri = orc.tiled_parallel(ev, eta)
    assert rel_l2(r, i, rr / p, ri / p) <= 1e-6


def test_f32_entry_point(shk, orc):
    # mod.rs:805-851: float in, float out, scaled.  The kernel is the
    # compensated one, so it is checked against f64 on the rounded input.
    p = 5000
    ev32 = np.sort(np.random.default_rng(3).uniform(0.3, 2.5, p)).astype(np.float32)
    eta = np.float32(1.0 / np.sqrt(p))
    r, i = shk.compute_all_stieltjes_f32(ev32, eta, "blocked")
    assert r.dtype == np.float32
    rr, ri = orc.tiled_parallel(ev32.astype(np.float64), float(eta))
    assert rel_l2(r, i, rr / p, ri / p) <= 1e-6


def test_cutoff_and_windowed(shk, orc):
    # cacheblock.rs:1779-1884: windowed == cutoff, hybrid = exact re + windowed im.
    p = 2000
    ev = np.sort(orc.log_spectrum(p))
    eta = 0.1 / np.sqrt(p)
    cut = 10.0
    ref = np.array([orc.sum_cutoff(x, ev, eta, cut) for x in ev])
    for m in ("blocked", "blocked_tiled", "blocked_windowed"):
        r, i = shk.compute_all_stieltjes(ev, eta, m, cutoff=cut)
        assert_close(r, i, ref[:, 0] / p, ref[:, 1] / p, 1e-12, f"{m} cutoff")
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked_hybrid", cutoff=cut)
    er, ei = orc.tiled_parallel(ev, eta)
    assert np.max(np.abs(r - er / p)) <= 1e-12 * np.max(np.abs(er / p))
    assert np.max(np.abs(i - ref[:, 1] / p)) <= 1e-12 * np.max(np.abs(ref[:, 1] / p))


def test_at_points_matches_single_point(shk, orc):
    # cacheblock.rs:1651-1688
    p = 300
    ev = np.sort(orc.log_spectrum(p))
    eta = 0.1 / np.sqrt(p)
    q = np.linspace(ev[0] - 0.5, ev[-1] + 0.5, 200)
    for cut in (None, 10.0):
        r, i = shk.compute_stieltjes_at_points(q, ev, eta, "blocked", cutoff=cut)
        for k, x in enumerate(q):
            a, b = orc.sum_cutoff(x, ev, eta, cut)
            assert abs(r[k] - a) < 1e-12 and abs(i[k] - b) < 1e-12


def test_golden_master_through_gpu(shk):
    from shrinkers_amd.deconvolution import spectral_deconvolution
    g = json.load(open(GOLDEN))["golden_master"]
    res = spectral_deconvolution(np.array(g["evals"]), g["c"], g["n_points"], eta=g["eta"])
    for key in ("density", "w_re", "sample_stieltjes_imag", "population_stieltjes_imag"):
        got = getattr(res, key)
        exp = np.array(g[key])
        # approx::assert_relative_eq!(epsilon = 1e-14): absolute or relative
        ok = np.abs(got - exp) <= np.maximum(1e-14, 1e-14 * np.abs(exp))
        assert ok.all(), (key, np.max(np.abs(got - exp)))


def test_with_deriv(shk, orc):
    p = 300
    ev = orc.log_spectrum(p)
    eta = 0.05 / np.sqrt(p)
    got = shk.compute_all_stieltjes_with_deriv(ev, eta)
    ref = orc.with_deriv(ev, eta)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g - r)) <= 1e-12 * np.max(np.abs(r))
    res = shk.stieltjes_transform_with_deriv(ev, eta=eta)
    assert set(res) == {"real", "imag", "deriv_real", "deriv_imag"}


def test_scaled_ray(shk, orc):
    ev = 0.4 + (np.arange(61) + 1.0) * 0.05  # inverse.rs:301-303
    h = 0.13
    r, i = shk.compute_stieltjes_scaled_ray(ev, h)
    rr, ri = orc.scaled_ray(ev, h)
    assert np.max(np.abs(r - rr)) <= 1e-12 * np.max(np.abs(rr))
    assert np.max(np.abs(i - ri)) <= 1e-12 * np.max(np.abs(ri))


def test_large_p_properties(shk, orc):
    """p = 200 000: rows spot-checked against the oracle's single-point sum,
    antisymmetry sum_i Re S_i = 0, Im S_i >= 1/eta, FP32C vs FP64."""
    from shrinkers_amd import spectra as S
    p = 200_000
    ev = S.mp_inverse_cdf_spectrum(p, c=0.1)
    eta = 1.0 / np.sqrt(p)
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked")
    rows = np.linspace(0, p - 1, 64).astype(int)
    for k in rows:
        a, b = orc.autovec_sum(ev[k], ev, eta)
        assert abs(r[k] * p - a) <= 1e-12 * abs(b) and abs(i[k] * p - b) <= 1e-12 * abs(b)
    assert abs(np.sum(r)) <= 1e-12 * np.sum(np.abs(r)) * np.sqrt(p)
    assert np.all(i * p >= (1.0 / eta) * (1 - 1e-12))
    r2, i2 = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
    err = rel_l2(r2, i2, r, i)
    print(f"f32c vs f64 at p={p}: {err:.3e}", flush=True)
    assert err <= 1e-6


def test_errors_are_raised(shk):
    with pytest.raises(NotImplementedError):
        shk.compute_all_stieltjes(np.array([1.0, 2.0]), 0.1, "fft5")
    with pytest.raises(shk.ShkError):
        shk.compute_all_stieltjes(np.array([1.0, np.nan]), 0.1, "blocked")
    with pytest.raises(shk.ShkError):
        shk.compute_all_stieltjes(np.array([1.0, 2.0]), -0.1, "blocked")


def test_f64_paired_reciprocal_fallback(shk, orc):
    """The FP64 kernels take two reciprocals from one v_rcp_f64 of the
    product of two pair denominators and fall back to one reciprocal each
    where that product is not a normal number.  Scaling eigenvalues and eta
    by 2^300 makes every product overflow (all pairs take the fallback) and
    changes the sums by exactly 2^-300; a bulk near 1 with outliers near
    2^400 mixes both paths; a tiny eta with duplicated eigenvalues makes the
    products of coinciding pairs underflow."""
    from shrinkers_amd import spectra as S
    ev = S.mp_inverse_cdf_spectrum(3000, c=0.3)
    eta = 1.0 / np.sqrt(ev.size)
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked")
    big = np.ldexp(1.0, 300)
    rs, is_ = shk.compute_all_stieltjes(ev * big, eta * big, "blocked")
    assert_close(rs * big, is_ * big, r, i, 1e-13, "all pairs on the fallback")

    rng = np.random.default_rng(3)
    mixed = np.sort(np.concatenate([ev[:2000], np.ldexp(rng.uniform(1, 2, 40), 400)]))
    eta_m = 0.01
    r, i = shk.compute_all_stieltjes(mixed, eta_m, "blocked")
    rr, ri = orc.tiled_parallel(mixed, eta_m)
    assert_close(r, i, rr / mixed.size, ri / mixed.size, 1e-13, "paired + fallback mix")

    dup = np.sort(np.concatenate([ev[:500], ev[:500], ev[200:300]]))
    eta_t = np.ldexp(1.0, -300)
    r, i = shk.compute_all_stieltjes(dup, eta_t, "blocked")
    rr, ri = orc.tiled_parallel(dup, eta_t)
    assert_close(r, i, rr / dup.size, ri / dup.size, 1e-13, "underflowing products")


@pytest.mark.parametrize("p", [8192, 8193, 10000, 24576, 24577, 25087, 50000, 65536 + 511])
def test_symmetric_f32c_kernel(shk, orc, p):
    """The all-points FP32-compensated entry takes the symmetric kernel for
    8 192 <= p <= 4 194 304 (each unordered pair once, credited to both
    ends; tiles of 512).  Against the FP64 oracle within 1e-6 normwise, and
    per element within 1e-6 of max|S|; against the split-tile kernel (the
    spectrum as its own query points) within 1e-7.  Sizes: whole tiles, one
    element over, a partial last tile, the headline size, a partial tile
    past 65 536."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    ev = S.mp_inverse_cdf_spectrum(p, c=0.25)
    eta = 1.0 / np.sqrt(p)
    d = torch.from_numpy(ev).cuda()
    re, im = shk.all_points_dev(d, eta, "blocked", precision="f32c", scaled=False)
    sr, si = torch.empty_like(d), torch.empty_like(d)
    assert _lib.load().shk_exact_at_points_dev(d.data_ptr(), p, d.data_ptr(), p, eta, 1,
                                               sr.data_ptr(), si.data_ptr(), None) == _lib.SHK_OK
    torch.cuda.synchronize()
    re, im, sr, si = (t.cpu().numpy() for t in (re, im, sr, si))
    rr, ri = orc.tiled_parallel(ev, eta)
    assert_close(re, im, rr, ri, 1e-6, f"symmetric f32c p={p}")
    assert rel_l2(re, im, sr, si) <= 1e-7


@pytest.mark.parametrize("p", [10000, 30000])
def test_symmetric_f32c_hard_inputs(shk, orc, p):
    """Unsorted input (wide per-lane query spans: the compensated forms),
    heavy duplicates with a small eta (1e-4 and 1e-12: d = 0 exactly for the
    tied pairs, terms of 1/eta^2 = 1e24 in f32), and a log-spaced spectrum,
    at sizes inside the symmetric kernel's range (10 000 lies in the part
    below 24 576 that it took over from the split-tile kernel)."""
    from shrinkers_amd import spectra as S
    rng = np.random.default_rng(5)
    lo, hi = shk.symmetric_range("f32c")
    assert lo <= p <= hi
    cases = {
        "unsorted": (orc.shuffled(S.mp_inverse_cdf_spectrum(p), 3), 1.0 / np.sqrt(p)),
        "duplicates": (np.sort(np.repeat(rng.uniform(0.5, 2.0, p // 8), 8)), 1e-4),
        "duplicates_tiny_eta": (np.sort(np.repeat(rng.uniform(0.5, 2.0, p // 8), 8)), 1e-12),
        "log": (S.log_spectrum(p), 1.0 / np.sqrt(p)),
    }
    for name, (ev, eta) in cases.items():
        r, i = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
        rr, ri = orc.tiled_parallel(ev, eta)
        n = ev.size
        err = rel_l2(r, i, rr / n, ri / n)
        print(f"symmetric f32c {name}: rel-L2 {err:.3e}", flush=True)
        assert err <= 1e-6, name


@pytest.mark.parametrize("p", [18432, 20000, 24576, 24577, 50000])
def test_symmetric_f64_kernel(shk, orc, p):
    """The FP64 all-points entry takes the symmetric FP64 kernel for
    18 432 <= p <= 4 194 304: within 1e-13 of the FP64 oracle (normwise and per
    element against max|S|), and within 1e-13 of the split-tile FP64 kernel
    (the spectrum as its own query points)."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    ev = S.mp_inverse_cdf_spectrum(p, c=0.25)
    eta = 1.0 / np.sqrt(p)
    d = torch.from_numpy(ev).cuda()
    re, im = shk.all_points_dev(d, eta, "blocked", precision="f64", scaled=False)
    sr, si = torch.empty_like(d), torch.empty_like(d)
    assert _lib.load().shk_exact_at_points_dev(d.data_ptr(), p, d.data_ptr(), p, eta, 0,
                                               sr.data_ptr(), si.data_ptr(), None) == _lib.SHK_OK
    torch.cuda.synchronize()
    re, im, sr, si = (t.cpu().numpy() for t in (re, im, sr, si))
    rr, ri = orc.tiled_parallel(ev, eta)
    assert_close(re, im, rr, ri, 1e-13, f"symmetric f64 p={p}")
    assert_close(re, im, sr, si, 1e-13, f"symmetric vs split f64 p={p}")


def test_symmetric_f64_paired_fallback(shk):
    """The symmetric FP64 kernel's paired reciprocals fall back where the
    denominators' product is not normal: eigenvalues and eta scaled by 2^300
    (every pair on the fallback) change the sums by exactly 2^-300."""
    from shrinkers_amd import spectra as S
    ev = S.mp_inverse_cdf_spectrum(30000, c=0.3)
    eta = 1.0 / np.sqrt(ev.size)
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked")
    big = np.ldexp(1.0, 300)
    rs, is_ = shk.compute_all_stieltjes(ev * big, eta * big, "blocked")
    assert_close(rs * big, is_ * big, r, i, 1e-13, "symmetric f64, all pairs on the fallback")


def test_host_calls_from_several_threads(shk):
    """Host-array entry points called from four threads at once give the
    one-thread results bit for bit. Each calling thread has its own stream,
    cached device buffers, pinned staging buffer and (ChebCode) host mailbox;
    ctypes releases the GIL, so the calls overlap on the device. Sizes span the
    split-tile and symmetric exact kernels, a grid, and a one-shot ChebCode
    build; f32 covers the float32 entry."""
    from concurrent.futures import ThreadPoolExecutor
    from shrinkers_amd import spectra as S
    cases = []
    for k, p in enumerate((3000, 9000, 30000, 50000)):
        ev = S.mp_inverse_cdf_spectrum(p)
        cases.append((ev, 1.0 / np.sqrt(p), np.linspace(ev.min(), ev.max(), 200 + k)))

    def run(case):
        ev, eta, grid = case
        outs = [shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c"),
                shk.compute_all_stieltjes(ev, eta, "blocked"),
                shk.compute_stieltjes_at_points(grid, ev, eta, "blocked"),
                shk.compute_all_stieltjes(ev, eta, "chebcode_fast"),
                shk.compute_all_stieltjes_f32(ev, eta, "blocked")]
        return [np.asarray(o) for o in outs]

    want = [run(c) for c in cases]
    with ThreadPoolExecutor(4) as pool:
        for rep in range(3):
            got = list(pool.map(run, cases[rep % 2:] + cases[:rep % 2]))
            order = cases[rep % 2:] + cases[:rep % 2]
            for case, g in zip(order, got):
                w = want[next(i for i, c in enumerate(cases) if c is case)]
                for a, b in zip(w, g):
                    assert a.dtype == b.dtype and np.array_equal(a, b), f"rep {rep}, p={case[0].size}"


@pytest.mark.parametrize("p", [3000, 10000])
def test_f32c_outside_eta_range_runs_fp64(shk, orc, p):
    """The FP32-compensated entries form eta^2 and 1/eta^2 in f32; outside
    1e-18 <= eta <= 1e18 they run the FP64 kernels instead.  At eta = 1e-20
    (1/eta^2 = 1e40 overflows f32) and eta = 1e20 (eta^2 overflows), with
    duplicated eigenvalues, every f32c entry returns the FP64 result: the
    all-points entry (split-tile size 3000, symmetric size 10000), the
    at-points entry and the float32 entry, all within the FP64 bar of the
    oracle."""
    import torch
    from shrinkers_amd import spectra as S
    from shrinkers_amd import _lib
    ev = np.sort(np.concatenate([S.mp_inverse_cdf_spectrum(p - 100), S.mp_inverse_cdf_spectrum(p)[:100]]))
    for eta in (1e-20, 1e20):
        rr, ri = orc.tiled_parallel(ev, eta)
        r, i = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
        assert np.all(np.isfinite(r)) and np.all(np.isfinite(i))
        assert_close(r, i, rr / p, ri / p, 1e-13, f"f32c routed to FP64, eta={eta:g}")
        d = torch.from_numpy(ev).cuda()
        r2, i2 = shk.all_points_dev(d, eta, "blocked", precision="f32c", scaled=False)
        sr, si = torch.empty_like(d), torch.empty_like(d)
        assert _lib.load().shk_exact_at_points_dev(d.data_ptr(), p, d.data_ptr(), p, eta, 1,
                                                   sr.data_ptr(), si.data_ptr(),
                                                   None) == _lib.SHK_OK
        torch.cuda.synchronize()
        assert_close(r2.cpu().numpy(), i2.cpu().numpy(), rr, ri, 1e-13, "all_points_dev")
        assert_close(sr.cpu().numpy(), si.cpu().numpy(), rr, ri, 1e-13, "exact_at_points_dev")
    # the float32 entry at eta = 1e-20: the FP64 path on the widened input
    ev32 = ev.astype(np.float32)
    r3, i3 = shk.compute_all_stieltjes_f32(ev32, 1e-20, "blocked")
    rr, ri = orc.tiled_parallel(ev32.astype(np.float64), float(np.float32(1e-20)))
    assert np.all(np.isfinite(r3)) and np.all(np.isfinite(i3))
    assert rel_l2(r3, i3, rr / p, ri / p) <= 1e-6


def test_f32c_inside_eta_range_small_eta(shk, orc):
    """Just inside the range (eta = 1e-15, duplicates: diagonal terms of
    1e30 in f32) the FP32-compensated kernels still meet their 1e-6 bar."""
    from shrinkers_amd import spectra as S
    base = S.mp_inverse_cdf_spectrum(4000)
    for ev in (np.sort(np.repeat(base[:500], 6)), np.sort(np.repeat(base[:1500], 6))):
        eta = 1e-15
        r, i = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
        rr, ri = orc.tiled_parallel(ev, eta)
        assert rel_l2(r, i, rr / ev.size, ri / ev.size) <= 1e-6, ev.size


def test_f32c_batch_outside_eta_range(shk):
    """The prepared batch runs FP32-compensated only: it refuses an eta
    outside [1e-18, 1e18] with SHK_ERR_BAD_ETA.  The one-call batch runs
    such a set in FP64 (bit for bit the FP64 call)."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    from shrinkers_amd.batch import BatchedStieltjes, all_points_many_dev
    sp = [S.mp_inverse_cdf_spectrum(p) for p in (700, 1300)]
    etas = [1e-20, 0.03]
    with pytest.raises(_lib.ShkError) as e:
        BatchedStieltjes(sp, etas)
    assert e.value.status == _lib.SHK_ERR_BAD_ETA
    ev = torch.from_numpy(np.concatenate(sp)).cuda()
    off = np.array([0, 700, 2000], dtype=np.int64)
    r32, i32 = all_points_many_dev(ev, off, etas, precision="f32c")
    r64, i64 = all_points_many_dev(ev, off, etas, precision="f64")
    torch.cuda.synchronize()
    assert torch.equal(r32, r64) and torch.equal(i32, i64)


def test_config1_spiked_wishart(shk, orc):
    """Config 1: eigenvalues of X^T X / n for X of shape 4096 x 1024 with
    i.i.d. N(0, 1) entries scaled by sqrt(Sigma), Sigma = diag(12, 7, 4, 1,
    ..., 1) (seed 0), eta = 1/sqrt(p), z_i = lam_i + i eta for all i.  The
    FP64 GPU transform within 1e-13 of the oracle's tiled and symmetric
    restatements, the FP32-compensated one within 1e-6, and the ChebCode
    presets within 1e-13 (f64) / 1e-6 (FAST) of the oracle's ChebCode."""
    from shrinkers_amd import spectra as S
    ev = S.spiked_wishart_spectrum(4096, 1024)
    p = ev.size
    eta = 1.0 / np.sqrt(p)
    rr, ri = orc.tiled_parallel(ev, eta)
    sr, si = orc.symmetric_sweep(ev, eta)
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked")
    assert_close(r, i, rr / p, ri / p, 1e-13, "config 1 f64 vs tiled oracle")
    assert_close(r, i, sr / p, si / p, 1e-13, "config 1 f64 vs symmetric oracle")
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
    assert rel_l2(r, i, rr / p, ri / p) <= 1e-6
    for preset in ("chebcode", "chebcode_fast", "chebcode_xtreme", "chebcode_balanced"):
        r, i = shk.compute_all_stieltjes(ev, eta, preset)
        orr, ori = orc.chebcode_preset(ev, eta, preset)
        tol = 1e-6 if preset == "chebcode_fast" else 1e-13
        assert rel_l2(r, i, orr / p, ori / p) <= tol, preset


def _with_scratch_limit(nbytes, fn):
    from shrinkers_amd import _lib
    lib = _lib.load()
    _lib.check(lib.shk_set_symmetric_scratch_limit(int(nbytes)))
    try:
        return fn()
    finally:
        _lib.check(lib.shk_set_symmetric_scratch_limit(0))


@pytest.mark.parametrize("precision", ["f32c", "f64"])
@pytest.mark.parametrize("p,limit_mb", [(30000, 3), (30000, 12), (20481, 1)])
def test_symmetric_banded_equals_single_band(shk, orc, precision, p, limit_mb):
    """With the partial scratch capped below the whole triangle's (16 nt p
    bytes: 56 MB at 3e4, 13 MB at 20 481), the symmetric kernel runs the
    triangle in bands of rows and folds each band in slot order: the result
    equals the one-band run bit for bit (3 MB at 3e4: 30 bands of 2 rows;
    12 MB: 5 bands of 12 rows; 1 MB at 20 481, a partial last tile: 41 bands
    of 1 row)."""
    import torch
    from shrinkers_amd import spectra as S
    ev = S.mp_inverse_cdf_spectrum(p, c=0.3)
    eta = 1.0 / np.sqrt(p)
    d = torch.from_numpy(ev).cuda()
    want = shk.all_points_dev(d, eta, "blocked", precision=precision, scaled=False)
    torch.cuda.synchronize()
    got = _with_scratch_limit(limit_mb << 20, lambda: shk.all_points_dev(
        d, eta, "blocked", precision=precision, scaled=False))
    torch.cuda.synchronize()
    assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1])
    rr, ri = orc.tiled_parallel(ev, eta)
    tol = 1e-6 if precision == "f32c" else 1e-13
    assert rel_l2(got[0].cpu().numpy(), got[1].cpu().numpy(), rr, ri) <= tol


def test_symmetric_above_old_maximum(shk):
    """p = 300 000 (above the 262 144 the unbanded partials allowed: 2.8 GB of
    them) runs the symmetric kernels in bands within the 512 MiB budget:
    FP64 within 1e-13 of the split-tile FP64 kernel (the spectrum as its own
    query points), FP32-compensated within 1e-6 of FP64."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    p = 300_000
    ev = S.mp_inverse_cdf_spectrum(p, c=0.2)
    eta = 1.0 / np.sqrt(p)
    d = torch.from_numpy(ev).cuda()
    r64, i64 = shk.all_points_dev(d, eta, "blocked", precision="f64", scaled=False)
    r32, i32 = shk.all_points_dev(d, eta, "blocked", precision="f32c", scaled=False)
    sr, si = torch.empty_like(d), torch.empty_like(d)
    assert _lib.load().shk_exact_at_points_dev(d.data_ptr(), p, d.data_ptr(), p, eta, 0,
                                               sr.data_ptr(), si.data_ptr(), None) == _lib.SHK_OK
    torch.cuda.synchronize()
    r64, i64, r32, i32, sr, si = (t.cpu().numpy() for t in (r64, i64, r32, i32, sr, si))
    assert_close(r64, i64, sr, si, 1e-13, "banded symmetric f64 vs split f64 at 3e5")
    assert rel_l2(r32, i32, r64, i64) <= 1e-6



def test_exact_entries_under_capture(shk):
    """Under HIP graph capture (torch.cuda.graph on a side stream) the entry
    points that need scratch memory (the all-points transform, the at-points
    transform at many points, the all-points adjoint) are refused with
    SHK_ERR_BAD_ARG before any work is enqueued, since a captured call
    cannot share the thread's cached buffers; the capture stays valid.  The
    no-allocation entry (f32c and FP64, one band and several) and the
    at-points transform at a grid of points (no scratch), captured beside
    them, replay bit-equal to their eager calls, three times, with eager
    calls of every entry and synchronisations between the replays; the eager
    calls keep their results bit for bit."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    p1, p2 = 20_000, 200_000
    ev1 = torch.from_numpy(np.ascontiguousarray(S.mp_inverse_cdf_spectrum(p1, c=0.25))).cuda()
    ev2 = torch.from_numpy(np.ascontiguousarray(S.mp_inverse_cdf_spectrum(p2, c=0.3))).cuda()
    e1, e2 = 1.0 / np.sqrt(p1), 1.0 / np.sqrt(p2)
    q = ev1[::97].contiguous()          # 207 points: the grid kernel, no scratch
    q_many = ev1[:5000].contiguous()    # 5 000 points: the split plan and its partials
    rng = np.random.default_rng(5)
    g_re = torch.from_numpy(rng.standard_normal(p1)).cuda()
    g_im = torch.from_numpy(rng.standard_normal(p1)).cuda()
    noalloc = _lib.load().shk_all_points_noalloc_dev

    def na(ev, eta, prec):
        def call(out, st):
            _lib.check(noalloc(ev.data_ptr(), ev.numel(), eta, prec, 1, out[0].data_ptr(),
                               out[1].data_ptr(), st.cuda_stream))
        return call, (ev.numel(), ev.numel())

    def adjoint(out, st):
        _lib.check(_lib.load().shk_exact_adjoint_all_points_dev(
            ev1.data_ptr(), p1, e1, g_re.data_ptr(), g_im.data_ptr(), _lib.PREC_F32C,
            out[0].data_ptr(), out[1].data_ptr(), st.cuda_stream))

    def grid(prec):
        def call(out, st):
            _lib.check(_lib.load().shk_exact_at_points_dev(
                q.data_ptr(), q.numel(), ev1.data_ptr(), p1, e1, prec, out[0].data_ptr(),
                out[1].data_ptr(), st.cuda_stream))
        return call, (q.numel(), q.numel())

    captured = [na(ev1, e1, _lib.PREC_F32C), na(ev1, e1, _lib.PREC_F64),
                na(ev2, e2, _lib.PREC_F32C), grid(_lib.PREC_F64), grid(_lib.PREC_F32C),
                (lambda out, st: shk.at_points_dev(q, ev1, e1, "blocked", out=out, stream=st),
                 (q.numel(), q.numel()))]
    refused = [
        (lambda out, st: shk.all_points_dev(ev1, e1, "blocked", precision="f32c", out=out,
                                            stream=st), (p1, p1)),
        (lambda out, st: shk.all_points_dev(ev1, e1, "blocked", precision="f64", out=out,
                                            stream=st), (p1, p1)),
        (lambda out, st: shk.at_points_dev(q_many, ev1, e1, "blocked", out=out, stream=st),
         (q_many.numel(), q_many.numel())),
        (adjoint, (p1, 1)),
    ]
    main = torch.cuda.current_stream()

    def run_all(calls):
        res = []
        for fn, (n0, n1) in calls:
            out = (torch.empty(n0, dtype=torch.float64, device="cuda"),
                   torch.empty(n1, dtype=torch.float64, device="cuda"))
            fn(out, main)
            res.append(out)
        torch.cuda.synchronize()
        return res

    want_cap = run_all(captured)
    want_ref = run_all(refused)
    outs = [(torch.empty_like(a), torch.empty_like(b)) for a, b in want_cap]
    s = torch.cuda.Stream()
    s.wait_stream(main)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for fn, (n0, n1) in refused:
            out = (torch.empty(n0, dtype=torch.float64, device="cuda"),
                   torch.empty(n1, dtype=torch.float64, device="cuda"))
            with pytest.raises(_lib.ShkError) as err:
                fn(out, s)
            assert err.value.status == _lib.SHK_ERR_BAD_ARG
            assert "graph capture" in err.value.message
        for (fn, _), out in zip(captured, outs):
            fn(out, s)

    def same(got, want):
        return all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
                   for a, b in zip(got, want))

    for rep in range(3):
        for a, b in outs:
            a.fill_(float("nan"))
            b.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert same(outs, want_cap), rep
        assert same(run_all(captured), want_cap), rep
        assert same(run_all(refused), want_ref), rep


@pytest.mark.parametrize("nq,p", [(1, 8), (1, 50_000), (20, 160), (200, 50_000), (513, 10_000),
                                  (4096, 32_768), (4096, 40_001)])
def test_grid_query_path(shk, orc, nq, p):
    """A few query points against many sources (at most 4 096 points, at
    least 8 sources per point) take the one-workgroup-per-point kernel
    instead of the split plan: FP64 within 1e-13 of the oracle's at-points
    sums, FP32-compensated within 1e-6, through shk_exact_at_points_dev and
    (FP64) the method API.  Half of the points are eigenvalues themselves (a
    zero pair difference), the rest spread beyond both ends of the
    spectrum."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    from shrinkers_amd.stieltjes import _PRECISIONS
    rng = np.random.default_rng(nq + p)
    ev = (S.mp_inverse_cdf_spectrum(p, c=0.3, seed=1) if p >= 64
          else np.sort(rng.uniform(0.2, 3.0, p)))
    q = np.sort(np.concatenate([rng.uniform(ev[0] - 0.5, ev[-1] + 0.5, nq - nq // 2),
                                rng.choice(ev, nq // 2)]))
    eta = 1.0 / np.sqrt(p)
    rr, ri = orc.at_points(q, ev, eta)
    dq, de = torch.from_numpy(q).cuda(), torch.from_numpy(ev).cuda()
    for prec, tol in (("f64", 1e-13), ("f32c", 1e-6)):
        re, im = torch.full_like(dq, np.nan), torch.full_like(dq, np.nan)
        assert _lib.load().shk_exact_at_points_dev(
            dq.data_ptr(), nq, de.data_ptr(), p, eta, _PRECISIONS[prec], re.data_ptr(),
            im.data_ptr(), None) == _lib.SHK_OK
        torch.cuda.synchronize()
        assert_close(re.cpu().numpy(), im.cpu().numpy(), rr, ri, tol, f"grid {prec} {nq}x{p}")
    r, i = shk.compute_stieltjes_at_points(q, ev, eta, "blocked")
    assert_close(r, i, rr, ri, 1e-13, f"grid method API {nq}x{p}")


def test_speed_auto_follows_the_measured_table(shk, orc):
    """SpeedAuto resolves by size: the exact transform up to p = 20 000 (FP64
    within 1e-13 of the oracle), ChebCode FAST above (the same bits as asking
    for chebcode_fast); at points by the pair count (exact up to 2e8 pairs)."""
    from shrinkers_amd import spectra as S
    ev = S.mp_inverse_cdf_spectrum(1000, c=0.25)
    eta = 1.0 / np.sqrt(ev.size)
    r, i = shk.compute_all_stieltjes(ev, eta, "speed_auto")
    rr, ri = orc.tiled_parallel(ev, eta)
    assert_close(r, i, rr / ev.size, ri / ev.size, 1e-13, "speed_auto p=1000")
    big = S.mp_inverse_cdf_spectrum(30_000, c=0.25)
    eta = 1.0 / np.sqrt(big.size)
    a = shk.compute_all_stieltjes(big, eta, "speed_auto")
    b = shk.compute_all_stieltjes(big, eta, "chebcode_fast")
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    q = np.linspace(big[0] - 0.2, big[-1] + 0.2, 200)
    r, i = shk.compute_stieltjes_at_points(q, big, eta, "speed_auto")
    rr, ri = orc.at_points(q, big, eta)
    assert_close(r, i, rr, ri, 1e-13, "speed_auto at 200 points")
