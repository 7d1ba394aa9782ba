"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/shrinkers_stieltjes.h, and rejects bad arguments before
touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shrinkers_stieltjes.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shk_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("shk_compute_all_stieltjes", "shk_compute_stieltjes_at_points",
                 "shk_compute_all_stieltjes_f32", "shk_chebcode_build",
                 "shk_chebcode_evaluate_points", "shk_chebcode_evaluate_many",
                 "shk_all_points_dev"):
        assert must in names


def test_library_exports_every_declared_symbol(shk):
    from shrinkers_amd import _lib
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_method_names_and_aliases(shk):
    from shrinkers_amd import stieltjes as st
    # src/python.rs:321-350
    aliases = {"fft_fast": "nufft", "treecode": "fmm", "chebyshev": "chebcode",
               "chebf": "chebcode_fast", "chebx": "chebcode_xtreme", "chebb": "chebcode_balanced",
               "speed": "speed_auto", "accuracy": "accuracy_auto"}
    for k, name in enumerate(st.METHODS):
        assert st.parse_method(name) == k
        assert st.method_name(k) == name
    for a, name in aliases.items():
        assert st.method_name(st.parse_method(a)) == name
    with pytest.raises(ValueError, match="unknown method 'bogus'"):
        st.parse_method("bogus")


def test_status_codes_before_device(shk):
    from shrinkers_amd import _lib
    lib = _lib.load()
    ev = np.array([1.0, np.nan, 2.0])
    out = np.empty(3)
    st = lib.shk_compute_all_stieltjes(_lib.dptr(ev), 3, 0.1, 2, 0.0, 0, 0, _lib.dptr(out),
                                       _lib.dptr(out))
    assert st == _lib.SHK_ERR_NONFINITE
    assert "finite" in _lib.last_error()
    ev = np.array([1.0, 2.0, 3.0])
    st = lib.shk_compute_all_stieltjes(_lib.dptr(ev), 3, -1.0, 2, 0.0, 0, 0, _lib.dptr(out),
                                       _lib.dptr(out))
    assert st == _lib.SHK_ERR_BAD_ETA
    st = lib.shk_compute_all_stieltjes(_lib.dptr(ev), 3, 0.1, 8, 0.0, 0, 0, _lib.dptr(out),
                                       _lib.dptr(out))
    assert st == _lib.SHK_ERR_UNSUPPORTED and "fft5" in _lib.last_error()
    st = lib.shk_compute_all_stieltjes(_lib.dptr(ev), 3, 0.1, 99, 0.0, 0, 0, _lib.dptr(out),
                                       _lib.dptr(out))
    assert st == _lib.SHK_ERR_BAD_METHOD
    # p = 0 returns OK and writes nothing (src/stieltjes/mod.rs:611-613)
    assert lib.shk_compute_all_stieltjes(None, 0, 0.1, 2, 0.0, 0, 0, None, None) == _lib.SHK_OK
    h = ctypes.c_void_p()
    assert lib.shk_chebcode_build(None, 0, 0.5, 11, 32, 0, ctypes.byref(h)) == _lib.SHK_ERR_EMPTY


def test_auto_resolution(shk):
    from shrinkers_amd import _lib
    lib = _lib.load()
    assert lib.shk_resolve_auto_method(19, 0, 50000) == 14   # auto -> chebcode_fast
    assert lib.shk_resolve_auto_method(21, 1, 20001) == 14   # speed_auto above 20 000
    assert lib.shk_resolve_auto_method(21, 1, 20000) == 2    # speed_auto: exact up to 20 000
    assert lib.shk_resolve_auto_method(19, 0, 10) == 2
    assert lib.shk_resolve_auto_method(22, 0, 10) == 2       # accuracy_auto -> blocked
    assert lib.shk_resolve_auto_method(4, 0, 10) == 4
    assert abs(lib.shk_default_eta(400) - 0.005) < 1e-15


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU may be present")
def test_no_device_is_an_error_not_a_fallback(shk):
    from shrinkers_amd import _lib
    lib = _lib.load()
    ev = np.array([1.0, 2.0, 3.0])
    out = np.empty(3)
    st = lib.shk_compute_all_stieltjes(_lib.dptr(ev), 3, 0.1, 2, 0.0, 0, 0, _lib.dptr(out),
                                       _lib.dptr(out))
    assert st == _lib.SHK_ERR_NO_DEVICE


def test_batch_argument_checks(shk):
    from shrinkers_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    off = np.array([0, 3, 5], dtype=np.int64)
    eta = np.array([0.1, 0.1])
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    dummy = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    assert lib.shk_batch_create(dummy, offp, 0, _lib.dptr(eta), 1, ctypes.byref(h)) == _lib.SHK_ERR_EMPTY
    bad = np.array([1, 3, 5], dtype=np.int64)
    assert lib.shk_batch_create(dummy, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 2,
                                _lib.dptr(eta), 1, ctypes.byref(h)) == _lib.SHK_ERR_BAD_ARG
    empty = np.array([0, 3, 3], dtype=np.int64)
    assert lib.shk_batch_create(dummy, empty.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 2,
                                _lib.dptr(eta), 1, ctypes.byref(h)) == _lib.SHK_ERR_EMPTY
    neg = np.array([0.1, -1.0])
    assert lib.shk_batch_create(dummy, offp, 2, _lib.dptr(neg), 1, ctypes.byref(h)) == _lib.SHK_ERR_BAD_ETA
    assert lib.shk_batch_set_config(None, 2) == _lib.SHK_ERR_BAD_ARG
    assert lib.shk_batch_run(None, dummy, dummy, None) == _lib.SHK_ERR_BAD_ARG
    assert lib.shk_batch_destroy(None) == _lib.SHK_OK


def test_many_argument_checks(shk):
    """shk_all_points_many_dev rejects bad arguments before any device work."""
    from shrinkers_amd import _lib
    lib = _lib.load()
    eta = np.array([0.1, 0.1])
    dummy = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    i64 = lambda a: np.asarray(a, dtype=np.int64).ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    off = np.array([0, 3, 5], dtype=np.int64)
    assert lib.shk_all_points_many_dev(dummy, i64([1, 3, 5]), 2, _lib.dptr(eta), 2, 1, 1, dummy,
                                       dummy, 2, None, None) == _lib.SHK_ERR_BAD_ARG
    assert lib.shk_all_points_many_dev(dummy, i64([0, 3, 2]), 2, _lib.dptr(eta), 2, 1, 1, dummy,
                                       dummy, 2, None, None) == _lib.SHK_ERR_BAD_ARG
    for ns in (-1, 9):
        assert lib.shk_all_points_many_dev(dummy, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                           2, _lib.dptr(eta), 2, 1, 1, dummy, dummy, ns, None,
                                           None) == _lib.SHK_ERR_BAD_ARG
    bad = np.array([0.1, -2.0])
    assert lib.shk_all_points_many_dev(dummy, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       2, _lib.dptr(bad), 2, 1, 1, dummy, dummy, 2, None,
                                       None) == _lib.SHK_ERR_BAD_ETA
    assert lib.shk_all_points_many_dev(dummy, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       2, _lib.dptr(eta), 8, 1, 1, dummy, dummy, 2, None,
                                       None) == _lib.SHK_ERR_UNSUPPORTED   # fft5
    assert lib.shk_all_points_many_dev(None, None, 0, None, 2, 1, 1, None, None, 2, None,
                                       None) == _lib.SHK_OK


def test_all_points_adjoint_argument_checks(shk):
    """shk_exact_adjoint_all_points_dev rejects bad arguments before any
    device work; with no output requested it does nothing."""
    from shrinkers_amd import _lib
    lib = _lib.load()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    f = lib.shk_exact_adjoint_all_points_dev
    assert f(dummy, 5, 0.1, dummy, dummy, _lib.PREC_F64, None, None, None) == _lib.SHK_OK
    assert f(None, 5, 0.1, dummy, dummy, _lib.PREC_F64, dummy, None, None) == _lib.SHK_ERR_BAD_ARG
    assert f(dummy, 5, 0.1, None, dummy, _lib.PREC_F64, dummy, None, None) == _lib.SHK_ERR_BAD_ARG
    assert f(dummy, 5, -0.1, dummy, dummy, _lib.PREC_F64, dummy, None, None) == _lib.SHK_ERR_BAD_ETA
    assert f(dummy, 5, 0.1, dummy, dummy, _lib.PREC_F32, dummy, None, None) == _lib.SHK_ERR_BAD_ARG
    assert "precision" in _lib.last_error()


def test_chebcode_workspace_sizes(shk):
    """The workspace bound (every node slot: 2p plus the top walk's slots)
    grows with p and with the interpolation order, covers the tree arrays of
    the largest tree a build can make, and the workspace entries take the
    ChebCode presets only.  Sizes need no device."""
    from shrinkers_amd import _lib
    lib = _lib.load()
    n = ctypes.c_size_t()

    def size(p, m):
        st = lib.shk_chebcode_workspace_bytes(p, m, ctypes.byref(n))
        return st, n.value

    assert size(0, 14) == (_lib.SHK_OK, 0)
    prev = 0
    for p in (1, 2, 33, 1000, 50_000, 1_000_000):
        st, fast = size(p, 14)       # chebcode_fast, n = 8
        st2, deflt = size(p, 13)     # chebcode, n = 11
        assert st == st2 == _lib.SHK_OK
        assert deflt > fast > prev and fast % 256 == 0
        # node arrays alone at 2p - 1 nodes: 3 f64, 4 i32, 2n f64, float4, int4, 2n f32
        assert fast >= (2 * p - 1) * (3 * 8 + 4 * 4 + 16 * 8 + 16 + 16 + 8 * 8)
        prev = fast
    for m in (2, 19, 21, 22, 999):   # blocked, auto, speed_auto, accuracy_auto, unknown
        assert size(1000, m)[0] == _lib.SHK_ERR_BAD_METHOD, m
    assert lib.shk_chebcode_workspace_bytes(1000, 14, None) == _lib.SHK_ERR_BAD_ARG
