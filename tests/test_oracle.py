"""Pins the CPU oracle (oracle/) against checks that do not restate it:
closed forms, exact rational arithmetic, the reference's golden master and
recorded error tables, and the reference's own test bounds."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import rel_l2

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


def exact_sum(x, ev, eta):
    """sum_j 1/((x - lam_j) - i eta) in exact rational arithmetic."""
    X, E = Fraction(x), Fraction(eta)
    re = im = Fraction(0)
    for lj in ev:
        d = X - Fraction(lj)
        den = d * d + E * E
        re += d / den
        im += E / den
    return float(re), float(im)


def exact_deriv(x, ev, eta):
    """S'(x) = -sum 1/(x - lam_j - i eta)^2 exactly."""
    X, E = Fraction(x), Fraction(eta)
    dr = di = Fraction(0)
    for lj in ev:
        d = X - Fraction(lj)
        den = d * d + E * E
        dr -= (d * d - E * E) / (den * den)
        di -= 2 * d * E / (den * den)
    return float(dr), float(di)


def small_cases():
    rng = np.random.default_rng(5)
    yield np.array([1.0, 2.0, 3.0]), 0.1
    yield np.array([3.0, -1.5, 2.25, 0.7, 4.2, -0.3, 1.1]), 0.37
    yield np.log1p((np.arange(24) + 0.5) * 0.2), 0.05
    yield rng.uniform(0.2, 3.0, 17), 0.013
    yield np.array([2.0]), 0.5
    yield np.array([1.0, 1.0, 1.0, 2.5]), 0.2  # duplicates


def test_term_closed_form(orc):
    # term.rs:51-59: 1/((2-1) - 0.1i) = (1 + 0.1i)/1.01
    r, i = orc.term(2.0, 1.0, 0.1)
    assert abs(r - 1 / 1.01) < 1e-15 and abs(i - 0.1 / 1.01) < 1e-15


def test_three_point_closed_form(orc):
    # deconvolution/mod.rs:646-658: evals [1,2,3], z = 2 + 0.1i: S = 10.198...i
    ev = np.array([1.0, 2.0, 3.0])
    for fn in (orc.naive_sum, orc.autovec_sum):
        r, i = fn(2.0, ev, 0.1)
        assert abs(r) < 1e-12 and abs(i - 10.198019801980198) < 1e-12


@pytest.mark.parametrize("case", list(range(6)))
def test_exact_kernels_against_rationals(orc, case):
    ev, eta = list(small_cases())[case]
    p = ev.size
    ref = np.array([exact_sum(x, ev, eta) for x in ev])
    scale = np.max(np.abs(ref))
    tol = 8 * p * np.finfo(float).eps * scale
    kernels = {
        "exact": orc.exact_stieltjes(ev, eta),
        "symmetric": orc.symmetric_sweep(ev, eta),
        "tiled": orc.tiled_parallel(ev, eta),
        "at_points": orc.at_points(ev, ev, eta),
        "naive": tuple(np.array(v) for v in zip(*[orc.naive_sum(x, ev, eta) for x in ev])),
        "autovec": tuple(np.array(v) for v in zip(*[orc.autovec_sum(x, ev, eta) for x in ev])),
    }
    for name, (r, i) in kernels.items():
        assert np.max(np.abs(r - ref[:, 0])) <= tol, name
        assert np.max(np.abs(i - ref[:, 1])) <= tol, name


def test_diagonal_only(orc):
    # cacheblock.rs:1604-1607
    r, i = orc.symmetric_sweep(np.array([2.0]), 0.5)
    assert r[0] == 0.0 and abs(i[0] - 2.0) < 1e-15


def test_symmetric_vs_full_square(orc):
    # cacheblock.rs:1609-1623
    big = np.sort(((np.arange(1000) * 7919) % 9973) / 9973.0 * 4.0 - 2.0)
    sr, si = orc.symmetric_sweep(big, 0.11)
    pr, pi = orc.tiled_parallel(big, 0.11)
    scale = np.maximum(np.abs(sr), np.abs(si))
    assert np.max((np.abs(pr - sr) + np.abs(pi - si)) / scale) < 1e-12


def test_tiled_remainder_sizes(orc):
    # mod.rs:1036-1085, the p % 4 remainder regression
    for p in (300, 2001, 2002, 2003, 4001):
        ev = np.sort(orc.log_spectrum(p))
        eta = 0.1 / np.sqrt(p)
        sr, si = orc.symmetric_sweep(ev, eta)
        pr, pi = orc.tiled_parallel(ev, eta)
        scale = np.maximum(np.abs(sr), np.abs(si))
        assert np.max((np.abs(pr - sr) + np.abs(pi - si)) / scale) < 1e-12


def test_recorded_exact_family_errors(orc):
    """The recorded blocked-parallel vs blocked-seq rel-L2 on harness_spectrum
    (docs/pareto/bench_after.json) is reproduced by tiled vs symmetric
    restatements to 3 significant digits: the restatements carry the
    reference's rounding."""
    rows = json.load(open(GOLDEN))["bench_after"]["rows"]
    rec = {(r["method"], r["p"], r["par"]): r["err"] for r in rows}
    for p in (2000, 5000):
        ev = orc.harness_spectrum(p)
        eta = 1.0 / np.sqrt(p)
        ip = 1.0 / p
        sr, si = orc.symmetric_sweep(ev, eta)
        pr, pi = orc.tiled_parallel(ev, eta)
        err = rel_l2(pr * ip, pi * ip, sr * ip, si * ip)
        want = rec[("blocked", p, "parallel")]
        assert abs(err - want) <= 0.002 * want, (p, err, want)


def test_recorded_chebcode_errors(orc):
    """Every f64 ChebCode preset reproduces its recorded rel-L2 on
    harness_spectrum to 3 significant digits; the FAST preset does too with the
    opening multiplier of panels outside the query set to 1.0 (the recorded
    run predates OUTSIDE_THETA2)."""
    rows = json.load(open(GOLDEN))["bench_after"]["rows"]
    rec = {(r["method"], r["p"], r["par"]): r["err"] for r in rows}
    for p in (1000, 2000, 5000):
        ev = orc.harness_spectrum(p)
        eta = 1.0 / np.sqrt(p)
        sr, si = orc.symmetric_sweep(ev, eta)
        for preset in ("chebcode", "chebcode_xtreme", "chebcode_balanced"):
            r, i = orc.chebcode_preset(ev, eta, preset)
            err = rel_l2(r, i, sr, si)
            want = rec[(preset, p, "seq")]
            assert abs(err - want) <= 0.002 * want, (preset, p, err, want)
        orc.set_outside_theta2(1.0)
        try:
            r, i = orc.chebcode_preset(ev, eta, "chebcode_fast")
        finally:
            orc.set_outside_theta2(1.69)
        err = rel_l2(r, i, sr, si)
        want = rec[("chebcode_fast", p, "seq")]
        assert abs(err - want) <= 0.002 * want, ("fast", p, err, want)


def test_chebcode_reference_bounds(orc):
    # chebcode.rs:1402-1419 (vs autovec, max-rel < 1e-2) at p = 512, 2000
    for p in (512, 2000):
        ev = orc.log_spectrum(p)
        eta = 0.1 / np.sqrt(p)
        r, i = orc.chebcode_preset(ev, eta, "chebcode")
        ref = np.array([orc.autovec_sum(x, ev, eta) for x in ev])
        assert np.max(np.abs(r - ref[:, 0]) / np.maximum(np.abs(ref[:, 0]), 1e-12)) < 1e-2
        assert np.max(np.abs(i - ref[:, 1]) / np.maximum(np.abs(ref[:, 1]), 1e-12)) < 1e-2
    # chebcode.rs:1422-1441 (n = 64, 65, 96)
    p = 500
    ev = orc.log_spectrum(p)
    eta = 0.1 / np.sqrt(p)
    er, ei = orc.exact_stieltjes(ev, eta)
    for n in (64, 65, 96):
        r, i = orc.chebcode_all(ev, eta, 0.5, n, 32)
        assert rel_l2(r, i, er, ei) < 1e-6
    # chebcode.rs:1464-1487 (f32 far field tracks f64)
    p = 4000
    ev = orc.log_spectrum(p)
    eta = 0.1 / np.sqrt(p)
    a = orc.chebcode_all(ev, eta, 0.5, 9, 32, False)
    b = orc.chebcode_all(ev, eta, 0.5, 9, 32, True)
    assert rel_l2(b[0], b[1], a[0], a[1]) < 1e-4


def test_two_eta_traversal(orc):
    # chebcode.rs:1543-1562
    p = 256
    ev = orc.log_spectrum(p)
    t = orc.ChebTree(ev, 9, 0.3, 16)
    for x in ev[::17]:
        r, i, _, _ = t.eval_point(x, 0.05)
        (r0, i0), _ = t.eval_x2(x, 0.05, 0.05)
        assert abs(r - r0) < 1e-9 and abs(i - i0) < 1e-9


def test_tree_invariants(orc):
    """Pre-order layout: left child = node+1, leaves partition the sources in
    order, leaf weights sum to the source count, parents keep their mass."""
    ev = np.sort(orc.harness_spectrum(5000))
    for n, theta, leaf in ((11, 0.5, 32), (8, 1.0, 32), (11, 0.25, 16)):
        t = orc.ChebTree(ev, n, theta, leaf).export()
        right, lo_idx, hi_idx = t["right"], t["lo_idx"], t["hi_idx"]
        w = t["w"].reshape(-1, n)
        leaves = np.where(right < 0)[0]
        assert lo_idx[leaves[0]] == 0 and hi_idx[leaves[-1]] == ev.size
        np.testing.assert_array_equal(hi_idx[leaves[:-1]], lo_idx[leaves[1:]])
        assert np.all(hi_idx[leaves] - lo_idx[leaves] <= leaf)
        for k in np.where(right >= 0)[0]:
            assert lo_idx[k + 1] == lo_idx[k] and hi_idx[right[k]] == hi_idx[k]
            assert hi_idx[k + 1] == lo_idx[right[k]]
        mass = w.sum(axis=1)
        np.testing.assert_allclose(mass, hi_idx - lo_idx, rtol=1e-12)


def test_derivative(orc):
    ev = list(small_cases())[2][0]
    eta = 0.05
    vr, vi, dr, di = orc.with_deriv(ev, eta)
    for k, x in enumerate(ev):
        er, ei = exact_deriv(x, ev, eta)
        s = max(abs(er), abs(ei))
        assert abs(dr[k] - er) <= 1e-11 * s and abs(di[k] - ei) <= 1e-11 * s
    # central differences, autovec.rs:115-148
    p = 300
    ev = orc.log_spectrum(p)
    eta = 0.05 / np.sqrt(p)
    vr, vi, dr, di = orc.with_deriv(ev, eta)
    for i in (0, 1, p // 3, p - 1):
        h = min(1e-5, max(abs(ev[i]), 1.0) * 1e-5)
        sp = orc.autovec_sum(ev[i] + h, ev, eta)
        sm = orc.autovec_sum(ev[i] - h, ev, eta)
        fd = ((sp[0] - sm[0]) / (2 * h), (sp[1] - sm[1]) / (2 * h))
        tol = 8e-2 if i == 0 else 5e-4
        assert abs(dr[i] - fd[0]) < tol * (abs(fd[0]) + 1.0)
        assert abs(di[i] - fd[1]) < tol * (abs(fd[1]) + 1.0)


def test_scaled_ray(orc):
    ev = 0.4 + (np.arange(12) + 1.0) * 0.05
    h = 0.13
    r, i = orc.scaled_ray(ev, h)
    for k, x in enumerate(ev):
        er, ei = exact_sum(x, ev, h * x)
        assert abs(r[k] - er) < 1e-12 * abs(ei) and abs(i[k] - ei) < 1e-12 * abs(ei)


def test_cutoff_window(orc):
    ev = np.sort(orc.log_spectrum(400))
    eta = 0.1 / 20.0
    x = ev[123]
    r, i = orc.sum_cutoff(x, ev, eta, 10.0)
    keep = ev[np.abs(x - ev) <= 10.0 * eta]
    er, ei = exact_sum(x, keep, eta)
    assert abs(r - er) < 1e-10 * abs(ei) and abs(i - ei) < 1e-10 * abs(ei)


def test_golden_master_cpu(orc):
    """Oracle at-points kernel + the El Karoui algebra reproduce the
    reference golden master (deconvolution/mod.rs:281-397) at 1e-14."""
    from shrinkers_amd.deconvolution import deconvolution_grid, el_karoui_inversion
    g = json.load(open(GOLDEN))["golden_master"]
    ev = np.array(g["evals"])
    grid = deconvolution_grid(ev, g["n_points"])
    r, i = orc.at_points(grid, ev, g["eta"])
    res = el_karoui_inversion(grid, r, i, ev.size, g["c"], g["eta"])
    for key in ("density", "w_re", "sample_stieltjes_imag", "population_stieltjes_imag"):
        got, exp = getattr(res, key), np.array(g[key])
        assert np.all(np.abs(got - exp) <= np.maximum(1e-14, 1e-14 * np.abs(exp))), key


def test_f32_reference_bound(orc):
    # cacheblock.rs:1767-1793
    p = 500
    ev = np.sort(orc.log_spectrum(p))
    eta = 0.1 / np.sqrt(p)
    r64, i64 = orc.symmetric_sweep(ev, eta)
    r32, i32 = orc.tiled_f32(ev.astype(np.float32), np.float32(eta))
    err = max(np.max(np.abs(r64 - r32) / np.maximum(np.abs(r64), 1e-12)),
              np.max(np.abs(i64 - i32) / np.maximum(np.abs(i64), 1e-12)))
    assert err < 5e-2


def test_chebcode_error_at_config3_is_the_algorithms(orc):
    """At p = 2e5, MP c = 0.5 with the three outliers, eta = 1/sqrt(p), the
    reference algorithm as restated here (build and traversals) lands above
    the documented classes: FAST 2.81e-5 (class ~2e-5) and DEFAULT 3.26e-9
    (class ~5e-10) against the exact sums.  The GPU measured the same two
    figures against its FP64 exact kernel (FAST 2.811e-5, DEFAULT 3.259e-9),
    so the excess is the algorithm's on this input, not the GPU's: pinned
    here to 2 digits."""
    from shrinkers_amd import spectra as S
    p = 200_000
    ev = S.mp_inverse_cdf_spectrum(p, c=0.5)
    eta = 1.0 / np.sqrt(p)
    er, ei = orc.tiled_parallel(ev, eta)
    for preset, want in (("chebcode_fast", 2.811e-5), ("chebcode", 3.259e-9)):
        r, i = orc.chebcode_preset(ev, eta, preset)
        err = rel_l2(r, i, er, ei)
        print(f"oracle {preset} at p=2e5 c=0.5: {err:.4e}", flush=True)
        assert abs(err / want - 1.0) < 0.01, (preset, err, want)


def test_config1_spiked_wishart_generator():
    """Config 1's spectrum (4096 x 1024 Wishart, Sigma = diag(12, 7, 4, 1, ...))
    carries three outliers whose inverse-BBP images (spiked/estimation.rs:
    10-11 inverted) recover the spikes within 5 %, and a bulk inside the
    Marchenko-Pastur support [(1 - sqrt c)^2, (1 + sqrt c)^2] up to edge
    fluctuations (the reference's sanity figure for this setup,
    README.md:42-49, is a 3-spike detection within ~2 %)."""
    from shrinkers_amd import spectra as S
    ev = S.spiked_wishart_spectrum(4096, 1024)
    c = 1024 / 4096
    top = ev[-3:][::-1]
    b = top + 1.0 - c
    ell = 0.5 * (b + np.sqrt(b * b - 4.0 * top))
    assert np.all(np.abs(ell / np.array([12.0, 7.0, 4.0]) - 1.0) < 0.05), ell
    lo, hi = (1 - np.sqrt(c)) ** 2, (1 + np.sqrt(c)) ** 2
    assert ev[0] > 0.9 * lo and ev[-4] < 1.1 * hi


def test_nufft_grid_size_rule(orc):
    """The coarse-grid NUFFT's sizing rule against the reference's own test
    (nufft.rs:239-263): the cap for a wide spectrum at small eta, 32 768 at
    range 2 / eta 0.01, the floor without padding, the cap (not an overflow)
    at eta = 1e-300, and an explicit override."""
    assert orc.nufft_grid_size(4.0, 1e-4) == 65536
    assert orc.nufft_grid_size(2.0, 0.01) == 32768
    assert orc.nufft_grid_size(0.1, 1.0, pad_eta_mult=0.0, pad_range_frac=0.0) == 8192
    assert orc.nufft_grid_size(2.0, 1e-300) == 65536
    assert orc.nufft_grid_size(4.0, 1e-4, m_override=12345) == 12345


def test_nufft_reference_bounds(orc):
    """nufft.rs:265-318: on the log spectrum at p = 8192, eta = 1/sqrt(p), the
    NUFFT is within 1e-4 normwise of the exact sums at all points and on a
    200-point grid, and the parallel splat agrees with the sequential one to
    1e-9."""
    p = 8192
    ev = orc.log_spectrum(p)
    eta = 1.0 / np.sqrt(p)
    er, ei = orc.tiled_parallel(ev, eta)
    r, i = orc.nufft_all(ev, eta)
    assert rel_l2(r, i, er, ei) < 1e-4
    rp, ip = orc.nufft_all(ev, eta, threads=4)
    assert max(np.max(np.abs(r - rp)), np.max(np.abs(i - ip))) < 1e-9
    grid = np.array([ev[0] + (ev[-1] - ev[0]) * k / 199.0 for k in range(200)])
    gr, gi = orc.at_points(grid, ev, eta)
    nr, ni = orc.nufft_at_points(ev, grid, eta)
    assert rel_l2(nr, ni, gr, gi) < 1e-4


def test_nufft_recorded_error_class(orc):
    """The reference's measured NUFFT error against the exact sums
    (docs/nufft_algorithms.md:141-147: rel-L2 2.1e-5 at p = 5e4, 2.7e-5 at
    1e5, eta = 1/sqrt(p)) on its benchmark spectrum mp_spectrum(p, 0.25, 7)
    (benches/support/mod.rs:42-54: uniform on the MP support plus a uniform
    jitter of 0.1, drawn from rand's StdRng).  The error is the kernel's
    periodization floor, a property of the distribution, so a spectrum of
    the same distribution from numpy's generator reproduces it to 2 digits."""
    for p, want in ((50_000, 2.1e-5), (100_000, 2.7e-5)):
        c = 0.25
        lo, hi = max(1 - np.sqrt(c), 0.01) ** 2, (1 + np.sqrt(c)) ** 2
        rng = np.random.default_rng(7)
        t = lo + rng.random(p) * (hi - lo)
        ev = np.sort(t + rng.random(p) * 0.1)
        eta = 1.0 / np.sqrt(p)
        er, ei = orc.tiled_parallel(ev, eta)
        r, i = orc.nufft_all(ev, eta, threads=4)
        err = rel_l2(r, i, er, ei)
        print(f"nufft p={p}: {err:.3e} (recorded {want:.1e})", flush=True)
        assert abs(err / want - 1.0) < 0.05, (p, err)
