"""Boundary validation of stieltjes_transform, mirroring the reference's
Python tests (tests/test_python_api.py:469-517) and validation helpers
(src/python.rs:138-318).  Every case raises before any GPU work."""
import numpy as np
import pytest


def sample_spectrum(p=200, seed=0):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(np.sort(0.5 + rng.uniform(0.0, 1.5, size=p)))


def test_empty_raises(shk):
    with pytest.raises(ValueError, match="non-empty"):
        shk.stieltjes_transform(np.array([], dtype=np.float64))


def test_bad_eta_raises(shk):
    with pytest.raises(ValueError, match="eta"):
        shk.stieltjes_transform(sample_spectrum(50), eta=-0.1)
    with pytest.raises(ValueError, match="eta must be a positive finite number"):
        shk.stieltjes_transform(sample_spectrum(50), eta=float("inf"))
    with pytest.raises(ValueError, match="expected a float, None, or the string 'inferred'"):
        shk.stieltjes_transform(sample_spectrum(50), eta="auto")


def test_bad_parallel_type_raises(shk):
    with pytest.raises(TypeError):
        shk.stieltjes_transform(sample_spectrum(50), parallel="yes")


def test_non_contiguous_raises(shk):
    with pytest.raises(ValueError, match="contiguous"):
        shk.stieltjes_transform(sample_spectrum(20)[::2])


def test_non_finite_raises(shk):
    ev = sample_spectrum(20)
    ev[3] = np.nan
    with pytest.raises(ValueError, match=r"eigenvalues must be finite \(found NaN or infinity\)"):
        shk.stieltjes_transform(ev)


def test_unknown_method_and_precision(shk):
    with pytest.raises(ValueError, match="unknown method 'nope'"):
        shk.stieltjes_transform(sample_spectrum(20), method="nope")
    with pytest.raises(ValueError, match="unknown precision 'f16'"):
        shk.stieltjes_transform(sample_spectrum(20), precision="f16")


def test_bad_cutoff(shk):
    with pytest.raises(ValueError, match="cutoff ratio must be a positive finite number"):
        shk.stieltjes_transform(sample_spectrum(20), cutoff=-1.0)


def test_out_of_scope_methods_are_refused(shk):
    for m in ("fft5", "nufft", "hodlr", "fmm", "ewald", "adaptive", "dst"):
        with pytest.raises(NotImplementedError, match=m):
            shk.stieltjes_transform(sample_spectrum(20), method=m)


def test_wrong_dtype_is_type_error(shk):
    with pytest.raises(TypeError):
        shk.stieltjes_transform(sample_spectrum(20).astype(np.float32))
    with pytest.raises(TypeError):
        shk.stieltjes_transform([1.0, 2.0, 3.0])


def test_default_etas(shk):
    from shrinkers_amd import stieltjes as st
    # mod.rs:1257-1283
    for p in (120, 400, 1000, 2000, 10_000):
        assert abs(st.default_eta_bulk(p) * np.sqrt(p) - 0.4) < 1e-12
    assert abs(st.default_eta(400) * 20.0 - 0.1) < 1e-12


def test_generators_match_oracle(orc):
    from shrinkers_amd import spectra as S
    for p in (1, 2, 3, 1000):
        np.testing.assert_array_equal(S.harness_spectrum(p), orc.harness_spectrum(p))
    np.testing.assert_array_equal(S.lcg_uniforms(42, 100), orc.lcg(42, 100))
    np.testing.assert_array_equal(S.log_spectrum(50), orc.log_spectrum(50))


def test_mp_generator(shk):
    """Inverse-CDF draws follow MP(c=0.25): CDF closed form vs quadrature,
    KS distance of 50k draws, forward-BBP outliers (estimation.rs:8-11)."""
    import math

    from scipy import integrate, stats

    from shrinkers_amd import spectra as S
    c = 0.25
    a, b = (1 - c ** 0.5) ** 2, (1 + c ** 0.5) ** 2
    dens = lambda x: math.sqrt(max((b - x) * (x - a), 0.0)) / (2 * math.pi * c * x)  # noqa: E731
    for th in (0.3, 1.0, 2.0, 2.8):
        x = 1 + c - 2 * c ** 0.5 * math.cos(th)
        q, _ = integrate.quad(dens, a, x, limit=200)
        assert abs(q - S._mp_cdf_theta(np.array([th]), c)[0]) < 1e-8
    ev = S.mp_inverse_cdf_spectrum(50_000)
    assert ev.size == 50_000 and np.all(np.diff(ev) >= 0)
    np.testing.assert_allclose(ev[-3:], [4.0 * (1 + c / 3.0), 7.0 * (1 + c / 6.0),
                                          12.0 * (1 + c / 11.0)], rtol=1e-15)
    bulk = ev[:-3]
    cdf = lambda x: S._mp_cdf_theta(np.arccos(np.clip((1 + c - x) / (2 * c ** 0.5), -1, 1)), c)  # noqa: E731
    assert stats.kstest(bulk, cdf).statistic < 0.01
