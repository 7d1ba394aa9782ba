"""Reverse mode on the GPU: the exact at-points pair and its adjoint, the El
Karoui fit kernel, and the differentiated bulk fit (p = 10 000, c = 0.25,
256 log-spaced atoms, 512 grid points, eta = 1/sqrt(p)).

Every gradient is compared with the oracle's FP64 restatement (itself pinned
by finite differences and exact invariances in test_oracle_adjoint.py), and
the full fit's gradients additionally with central finite differences of the
GPU FP64 forward.  Tolerances: FP64 path 1e-12 (kernels) / 1e-10 (composed
fit); FP32-compensated path 1e-6 for the forward sums and 5e-6 normwise for
gradients (measured on MI355X: 1.6e-7 to 9.7e-7).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _t(a, grad=False):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    return t.requires_grad_(grad)


def _rel(a, b):
    a, b = np.atleast_1d(np.asarray(a, dtype=np.float64)), np.atleast_1d(np.asarray(b))
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _spectrum(p, c=0.25, seed=0):
    from shrinkers_amd import spectra
    return spectra.mp_inverse_cdf_spectrum(p, c=c, seed=seed, spikes=())


@pytest.mark.parametrize("precision,tol", [("f64", 1e-13), ("f32c", 1e-6)])
def test_exact_at_points_dev_matches_oracle(shk, orc, precision, tol):
    import torch
    from shrinkers_amd import autograd as ag
    ev = _spectrum(5000)
    grid = np.linspace(ev[0] - 0.2, ev[-1] + 0.2, 700)
    eta = 1.0 / np.sqrt(ev.size)
    with torch.no_grad():
        re, im = ag.stieltjes_at_points(_t(grid), _t(ev), eta, precision)
    torch.cuda.synchronize()
    rr, ri = orc.at_points(grid, ev, eta)
    err = rel_l2(re.cpu().numpy(), im.cpu().numpy(), rr, ri)
    print(f"at-points {precision}: rel-L2 {err:.3e}", flush=True)
    assert err <= tol


def _adjoint_case(precision, q, ev, eta, seed):
    import torch
    from shrinkers_amd import autograd as ag
    rng = np.random.default_rng(seed)
    gre, gim = rng.standard_normal(q.size), rng.standard_normal(q.size)
    tq, tev, te = _t(q, True), _t(ev, True), _t(np.float64(eta).reshape(()), True)
    re, im = ag.stieltjes_at_points(tq, tev, te, precision)
    (re * _t(gre)).sum().add((im * _t(gim)).sum()).backward()
    torch.cuda.synchronize()
    return (tq.grad.cpu().numpy(), tev.grad.cpu().numpy(), float(te.grad.cpu()), gre, gim)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32c", 5e-6)])
def test_adjoint_matches_oracle(shk, orc, precision, tol):
    ev = _spectrum(3000)
    rng = np.random.default_rng(7)
    grid = np.concatenate([np.linspace(ev[0] - 0.1, ev[-1] + 0.1, 400),
                           ev[rng.choice(ev.size, 50, replace=False)]])  # some x_m = lam_j
    eta = 1.0 / np.sqrt(ev.size)
    gq, gev, ge, gre, gim = _adjoint_case(precision, grid, ev, eta, seed=1)
    rev, rq, re_ = orc.adjoint_at_points(grid, ev, eta, gre, gim)
    e_ev, e_q, e_eta = _rel(gev, rev), _rel(gq, rq), abs(ge - re_) / abs(re_)
    print(f"adjoint {precision}: lam {e_ev:.3e} x {e_q:.3e} eta {e_eta:.3e}", flush=True)
    assert e_ev <= tol and e_q <= tol and e_eta <= tol


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32c", 5e-6)])
@pytest.mark.parametrize("p,nq", [(5000, 200), (10_000, 512), (40_000, 4096), (9, 1)])
def test_adjoint_at_grid_points(shk, orc, precision, tol, p, nq):
    """A grid of query points (at most 4 096, at least 8 eigenvalues per
    point): the forward sums and the query sweep of the adjoint (dL/dx,
    dL/deta) run one workgroup per point.  Against the oracle's FP64 adjoint,
    with a quarter of the points on eigenvalues."""
    ev = _spectrum(p) if p >= 64 else np.sort(np.random.default_rng(p).uniform(0.2, 3.0, p))
    rng = np.random.default_rng(p + nq)
    grid = np.sort(np.concatenate([rng.uniform(ev[0] - 0.3, ev[-1] + 0.3, nq - nq // 4),
                                   rng.choice(ev, nq // 4, replace=False)]))
    eta = 1.0 / np.sqrt(p)
    gq, gev, ge, gre, gim = _adjoint_case(precision, grid, ev, eta, seed=p)
    rev, rq, re_ = orc.adjoint_at_points(grid, ev, eta, gre, gim)
    e_ev, e_q, e_eta = _rel(gev, rev), _rel(gq, rq), abs(ge - re_) / abs(re_)
    print(f"grid adjoint {precision} {nq}x{p}: lam {e_ev:.3e} x {e_q:.3e} eta {e_eta:.3e}",
          flush=True)
    assert e_ev <= tol and e_q <= tol and e_eta <= tol


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32c", 5e-6)])
def test_adjoint_all_points(shk, orc, precision, tol):
    """Queries = eigenvalues: autograd adds the query and eigenvalue parts."""
    import torch
    from shrinkers_amd import autograd as ag
    ev = _spectrum(2000, c=0.5, seed=3)
    eta = 1.0 / np.sqrt(ev.size)
    rng = np.random.default_rng(2)
    gre, gim = rng.standard_normal(ev.size), rng.standard_normal(ev.size)
    tev = _t(ev, True)
    re, im = ag.stieltjes_all_points(tev, eta, precision, scaled=False)
    ((re * _t(gre)).sum() + (im * _t(gim)).sum()).backward()
    torch.cuda.synchronize()
    rev, rq, _ = orc.adjoint_at_points(ev, ev, eta, gre, gim)
    err = _rel(tev.grad.cpu().numpy(), rev + rq)
    print(f"all-points adjoint {precision}: {err:.3e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("p,nq,eta", [(1, 1, 0.1), (130, 1, 0.05), (1, 129, 0.05),
                                      (257, 300, 1e-6), (64, 64, 3.0)])
def test_adjoint_edge_shapes(shk, orc, precision, p, nq, eta):
    rng = np.random.default_rng(p * 1000 + nq)
    ev = np.sort(rng.uniform(0.2, 3.0, p))
    q = np.sort(rng.uniform(0.1, 3.1, nq))
    if nq > 1 and p > 1:
        q[: min(nq, p) // 2] = ev[: min(nq, p) // 2]
    gq, gev, ge, gre, gim = _adjoint_case(precision, q, ev, eta, seed=p + nq)
    rev, rq, re_ = orc.adjoint_at_points(q, ev, eta, gre, gim)
    tol = 1e-12 if precision == "f64" else 5e-6
    assert _rel(gev, rev) <= tol
    assert _rel(gq, rq) <= tol
    assert abs(ge - re_) <= tol * abs(re_)


def _all_points_adjoint(ev, eta, gre, gim, precision, want_grad=True, want_eta=True):
    """shk_exact_adjoint_all_points_dev through the C ABI."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd.stieltjes import _PRECISIONS, _stream_handle
    tev, tre, tim = _t(ev), _t(gre), _t(gim)
    gev = torch.full_like(tev, np.nan) if want_grad else None
    ge = torch.full((), np.nan, dtype=torch.float64, device="cuda") if want_eta else None
    _lib.check(_lib.load().shk_exact_adjoint_all_points_dev(
        tev.data_ptr(), ev.size, float(eta), tre.data_ptr(), tim.data_ptr(),
        _PRECISIONS[precision], gev.data_ptr() if want_grad else None,
        ge.data_ptr() if want_eta else None, _stream_handle(None)))
    torch.cuda.synchronize()
    return (gev.cpu().numpy() if want_grad else None,
            float(ge.cpu()) if want_eta else None)


def _two_sweeps(ev, eta, gre, gim, precision):
    """shk_exact_adjoint_dev at queries = eigenvalues: (dL/dlam as source,
    dL/dx as query, dL/deta)."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd.stieltjes import _PRECISIONS, _stream_handle
    tev, tre, tim = _t(ev), _t(gre), _t(gim)
    gev, gq = torch.empty_like(tev), torch.empty_like(tev)
    ge = torch.empty((), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().shk_exact_adjoint_dev(
        tev.data_ptr(), ev.size, tev.data_ptr(), ev.size, float(eta), tre.data_ptr(),
        tim.data_ptr(), _PRECISIONS[precision], gev.data_ptr(), gq.data_ptr(), ge.data_ptr(),
        _stream_handle(None)))
    torch.cuda.synchronize()
    return gev.cpu().numpy(), gq.cpu().numpy(), float(ge.cpu())


def _hard_spectrum(case, p):
    from shrinkers_amd import spectra
    rng = np.random.default_rng(p)
    if case == "mp_spikes":
        return spectra.mp_inverse_cdf_spectrum(p, c=0.3, seed=4)  # outliers 12, 7, 4
    if case == "duplicates":
        return np.repeat(np.linspace(0.5, 2.0, p // 9), 9)
    if case == "unsorted":
        return rng.permutation(spectra.mp_inverse_cdf_spectrum(p, c=0.5, seed=5))
    raise ValueError(case)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32c", 5e-6)])
@pytest.mark.parametrize("p", [8192, 8193, 9001])
def test_symmetric_all_points_adjoint_matches_oracle(shk, orc, precision, tol, p):
    """Each unordered pair once (p >= 8192): the total eigenvalue gradient
    (query and source roles) and the eta gradient against the oracle's two
    FP64 sweeps."""
    ev = _hard_spectrum("mp_spikes", p)
    eta = 1.0 / np.sqrt(p)
    rng = np.random.default_rng(p + 11)
    gre, gim = rng.standard_normal(p), rng.standard_normal(p)
    gev, ge = _all_points_adjoint(ev, eta, gre, gim, precision)
    rev, rq, re_ = orc.adjoint_at_points(ev, ev, eta, gre, gim)
    e_ev, e_eta = _rel(gev, rev + rq), abs(ge - re_) / abs(re_)
    print(f"symmetric adjoint {precision} p={p}: lam {e_ev:.3e} eta {e_eta:.3e}", flush=True)
    assert e_ev <= tol and e_eta <= tol


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32c", 5e-6)])
@pytest.mark.parametrize("case", ["duplicates", "unsorted"])
def test_symmetric_all_points_adjoint_hard_inputs(shk, orc, precision, tol, case):
    """Ties (pairs at d = 0 off the diagonal) and unsorted input."""
    ev = _hard_spectrum(case, 9216)
    eta = 0.5 / np.sqrt(ev.size)
    rng = np.random.default_rng(3)
    gre, gim = rng.standard_normal(ev.size), rng.standard_normal(ev.size)
    gev, ge = _all_points_adjoint(ev, eta, gre, gim, precision)
    rev, rq, re_ = orc.adjoint_at_points(ev, ev, eta, gre, gim)
    e_ev, e_eta = _rel(gev, rev + rq), abs(ge - re_) / abs(re_)
    print(f"symmetric adjoint {precision} {case}: lam {e_ev:.3e} eta {e_eta:.3e}", flush=True)
    assert e_ev <= tol and e_eta <= tol


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_symmetric_all_points_adjoint_large(shk, precision):
    """p = 50 000 (the transform's headline size): against the two GPU FP64
    sweeps; deterministic; each output alone equals the joint call."""
    ev = _hard_spectrum("mp_spikes", 50_000)
    eta = 1.0 / np.sqrt(ev.size)
    rng = np.random.default_rng(9)
    gre, gim = rng.standard_normal(ev.size), rng.standard_normal(ev.size)
    gev, ge = _all_points_adjoint(ev, eta, gre, gim, precision)
    rev, rq, re_ = _two_sweeps(ev, eta, gre, gim, "f64")
    tol = 1e-12 if precision == "f64" else 5e-6
    e_ev, e_eta = _rel(gev, rev + rq), abs(ge - re_) / abs(re_)
    print(f"symmetric adjoint {precision} p=50000: lam {e_ev:.3e} eta {e_eta:.3e}", flush=True)
    assert e_ev <= tol and e_eta <= tol
    gev2, ge2 = _all_points_adjoint(ev, eta, gre, gim, precision)
    assert np.array_equal(gev, gev2) and ge == ge2
    only_grad, _ = _all_points_adjoint(ev, eta, gre, gim, precision, want_eta=False)
    assert np.array_equal(only_grad, gev)
    _, only_eta = _all_points_adjoint(ev, eta, gre, gim, precision, want_grad=False)
    assert abs(only_eta - ge) <= tol * abs(ge)   # the eta-only call runs the query sweep


@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("p,eta", [(3000, 0.02), (3000, 1e-7), (8192, 1e-10), (700, 1e-12)])
def test_all_points_adjoint_small_eta(shk, orc, precision, p, eta):
    """The all-points adjoint runs the symmetric kernels at every size up to
    their maximum, with the pair (i, i) left out of the gradient sums (its
    terms +-G_re/eta^2 cancel exactly).  Against the oracle's all-points
    gradient, which leaves it out too, at sizes below the old two-sweep
    threshold and at eta down to 1e-12, where 1/eta^2 = 1e24 would swamp the
    result of any cancellation.  FP32-compensated arithmetic runs for
    eta >= 1e-9 only; below, the call runs FP64 and equals the FP64 call
    bit for bit."""
    ev = _hard_spectrum("mp_spikes", p)
    rng = np.random.default_rng(p)
    gre, gim = rng.standard_normal(p), rng.standard_normal(p)
    gev, ge = _all_points_adjoint(ev, eta, gre, gim, precision)
    rg, re_ = orc.adjoint_all_points(ev, eta, gre, gim)
    f32 = precision == "f32c" and eta >= 1e-9
    tol = 5e-6 if f32 else 1e-12
    e_ev, e_eta = _rel(gev, rg), abs(ge - re_) / abs(re_)
    print(f"all-points adjoint {precision} p={p} eta={eta:g}: lam {e_ev:.3e} eta {e_eta:.3e}",
          flush=True)
    assert e_ev <= tol and e_eta <= tol
    if precision == "f32c" and not f32:
        fev, fe = _all_points_adjoint(ev, eta, gre, gim, "f64")
        assert np.array_equal(gev, fev) and ge == fe


def test_all_points_autograd_symmetric(shk, orc):
    """stieltjes_all_points (forward and backward through the all-points
    entries) at p = 8200 in FP64, scaled: eigenvalue and eta gradients
    against the oracle."""
    import torch
    from shrinkers_amd import autograd as ag
    p = 8200
    ev = _hard_spectrum("mp_spikes", p)
    eta = 1.0 / np.sqrt(p)
    rng = np.random.default_rng(21)
    gre, gim = rng.standard_normal(p), rng.standard_normal(p)
    tev, te = _t(ev, True), _t(np.float64(eta).reshape(()), True)
    re, im = ag.stieltjes_all_points(tev, te, "f64", scaled=True)
    ((re * _t(gre)).sum() + (im * _t(gim)).sum()).backward()
    torch.cuda.synchronize()
    rr, ri = orc.at_points(ev, ev, eta)
    assert rel_l2(re.detach().cpu().numpy(), im.detach().cpu().numpy(), rr / p, ri / p) <= 1e-13
    rev, rq, re_ = orc.adjoint_at_points(ev, ev, eta, gre / p, gim / p)
    assert _rel(tev.grad.cpu().numpy(), rev + rq) <= 1e-12
    assert abs(float(te.grad.cpu()) - re_) <= 1e-12 * abs(re_)


def _fit_inputs(seed=0):
    from shrinkers_amd import autograd as ag
    ev, grid, atoms, eta = ag.bulk_fit_problem()
    rng = np.random.default_rng(seed)
    logits = 0.5 * rng.standard_normal(atoms.size)
    return ev, grid, atoms, eta, logits


def test_el_karoui_kernel_matches_oracle(shk, orc):
    import torch
    from shrinkers_amd import autograd as ag
    ev, grid, atoms, eta, logits = _fit_inputs()
    sr, si = orc.at_points(grid, ev, eta)
    g_re, g_im = -sr / ev.size, si / ev.size
    w = np.exp(logits) / np.exp(logits).sum()
    c = 0.25
    tgr, tgi, tw, te = _t(g_re, True), _t(g_im, True), _t(w, True), _t(np.float64(eta).reshape(()), True)
    loss = ag.el_karoui_residual_loss(tgr, tgi, _t(grid), te, c, _t(atoms), tw)
    loss.backward()
    torch.cuda.synchronize()
    l_ref, ggr, ggi, gw, ge = orc.ek_fit(grid, g_re, g_im, eta, c, atoms, w)
    assert abs(float(loss.detach()) - l_ref) <= 1e-13 * l_ref
    assert _rel(tgr.grad.cpu().numpy(), ggr) <= 1e-12
    assert _rel(tgi.grad.cpu().numpy(), ggi) <= 1e-12
    assert _rel(tw.grad.cpu().numpy(), gw) <= 1e-12
    assert abs(float(te.grad) - ge) <= 1e-12 * abs(ge)


def _reference_fit_gradient(orc, ev, grid, atoms, eta, logits, c):
    """FP64 reference gradient composed from the oracle's pinned pieces."""
    p = ev.size
    w = np.exp(logits - logits.max())
    w /= w.sum()
    sr, si = orc.at_points(grid, ev, eta)
    g_re, g_im = -sr / p, si / p
    loss, ggr, ggi, gw, ge_direct = orc.ek_fit(grid, g_re, g_im, eta, c, atoms, w)
    # chain to the raw sums: g_re = -Re S/p, g_im = Im S/p
    gev, _, ge_s = orc.adjoint_at_points(grid, ev, eta, -ggr / p, ggi / p)
    glog = w * (gw - np.dot(w, gw))
    return loss, gev, ge_direct + ge_s, glog


def _gpu_fit(ev, grid, atoms, eta, logits, c, precision):
    import torch
    from shrinkers_amd import autograd as ag
    tev, te, tl = _t(ev, True), _t(np.float64(eta).reshape(()), True), _t(logits, True)
    loss = ag.el_karoui_fit_loss(tev, te, tl, _t(atoms), c, _t(grid), precision)
    loss.backward()
    torch.cuda.synchronize()
    return (float(loss.detach()), tev.grad.cpu().numpy(), float(te.grad), tl.grad.cpu().numpy())


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32c", 5e-6)])
def test_bulk_fit_gradients_match_fp64_reference(shk, orc, precision, tol):
    ev, grid, atoms, eta, logits = _fit_inputs()
    c = 0.25
    l_ref, gev_ref, ge_ref, gl_ref = _reference_fit_gradient(orc, ev, grid, atoms, eta, logits, c)
    loss, gev, ge, gl = _gpu_fit(ev, grid, atoms, eta, logits, c, precision)
    errs = (abs(loss - l_ref) / l_ref, _rel(gev, gev_ref), abs(ge - ge_ref) / abs(ge_ref),
            _rel(gl, gl_ref))
    print(f"bulk fit {precision}: loss {errs[0]:.2e} lam {errs[1]:.2e} eta {errs[2]:.2e} "
          f"logits {errs[3]:.2e}", flush=True)
    assert max(errs) <= tol


def test_bulk_fit_gradients_match_central_differences(shk):
    import torch
    from shrinkers_amd import autograd as ag
    ev, grid, atoms, eta, logits = _fit_inputs()
    c = 0.25
    loss, gev, ge, gl = _gpu_fit(ev, grid, atoms, eta, logits, c, "f64")

    def L(ev_=ev, eta_=eta, lg=logits):
        with torch.no_grad():
            return float(ag.el_karoui_fit_loss(_t(ev_), eta_, _t(lg), _t(atoms), c, _t(grid), "f64"))

    h = 1e-6
    for j in (0, 1234, ev.size // 2, ev.size - 1):
        e1, e0 = ev.copy(), ev.copy()
        e1[j] += h
        e0[j] -= h
        fd = (L(ev_=e1) - L(ev_=e0)) / (2 * h)
        assert abs(gev[j] - fd) <= 1e-6 * np.max(np.abs(gev)), (j, gev[j], fd)
    he = 1e-7 * eta
    fd = (L(eta_=eta + he) - L(eta_=eta - he)) / (2 * he)
    assert abs(ge - fd) <= 1e-6 * abs(fd), (ge, fd)
    for k in (0, 77, 255):
        l1, l0 = logits.copy(), logits.copy()
        l1[k] += h
        l0[k] -= h
        fd = (L(lg=l1) - L(lg=l0)) / (2 * h)
        assert abs(gl[k] - fd) <= 1e-6 * np.max(np.abs(gl)), (k, gl[k], fd)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32c", 5e-6)])
def test_fused_fit_step_matches_reference_and_autograd(shk, orc, precision, tol):
    """shk_el_karoui_fit_step_dev: the same loss and gradients as the FP64
    reference gradient (to the path's tolerance) and as the autograd
    composition of the same kernels (to rounding: 1e-12)."""
    import torch
    from shrinkers_amd import autograd as ag
    ev, grid, atoms, eta, logits = _fit_inputs()
    c = 0.25
    w = np.exp(logits - logits.max())
    w /= w.sum()
    loss, gev, ge, gw = ag.el_karoui_fit_step(_t(ev), eta, _t(w), _t(atoms), c, _t(grid), precision)
    torch.cuda.synchronize()
    loss, gev, ge, gw = float(loss), gev.cpu().numpy(), float(ge), gw.cpu().numpy()
    l_ref, gev_ref, ge_ref, gl_ref = _reference_fit_gradient(orc, ev, grid, atoms, eta, logits, c)
    glog = w * (gw - np.dot(w, gw))
    errs = (abs(loss - l_ref) / l_ref, _rel(gev, gev_ref), abs(ge - ge_ref) / abs(ge_ref),
            _rel(glog, gl_ref))
    print(f"fused step {precision}: loss {errs[0]:.2e} lam {errs[1]:.2e} eta {errs[2]:.2e} "
          f"logits {errs[3]:.2e}", flush=True)
    assert max(errs) <= tol
    a_loss, a_gev, a_ge, a_gl = _gpu_fit(ev, grid, atoms, eta, logits, c, precision)
    assert abs(loss - a_loss) <= 1e-12 * a_loss
    assert _rel(gev, a_gev) <= 1e-12
    assert abs(ge - a_ge) <= 1e-12 * abs(a_ge)
    assert _rel(glog, a_gl) <= 1e-12


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_all_points_adjoint_tiny(shk, orc, precision):
    """p = 0 writes a zero eta gradient; p = 1 and 2 (two-sweep sizes) match
    the oracle, the single eigenvalue's own term cancelling in the total."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd.stieltjes import _PRECISIONS, _stream_handle
    ge = torch.full((), np.nan, dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().shk_exact_adjoint_all_points_dev(
        None, 0, 0.1, None, None, _PRECISIONS[precision], None, ge.data_ptr(),
        _stream_handle(None)))
    torch.cuda.synchronize()
    assert float(ge.cpu()) == 0.0
    for ev in (np.array([1.3]), np.array([0.7, 1.9])):
        gre, gim = np.linspace(0.5, 1.5, ev.size), np.linspace(-1.0, 2.0, ev.size)
        gev, g_eta = _all_points_adjoint(ev, 0.05, gre, gim, precision)
        rev, rq, re_ = orc.adjoint_at_points(ev, ev, 0.05, gre, gim)
        tol = 1e-12 if precision == "f64" else 5e-6
        assert np.allclose(gev, rev + rq, rtol=tol, atol=tol * np.max(np.abs(rev)))
        assert abs(g_eta - re_) <= tol * abs(re_)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_symmetric_adjoint_banded_equals_single_band(shk, precision):
    """The all-points adjoint with its partial scratch capped at 2 MB (bands of
    a few rows at p = 20 000) equals the one-band run bit for bit, gradient and
    eta gradient."""
    from shrinkers_amd import _lib
    p = 20_000
    ev = _hard_spectrum("mp_spikes", p)
    eta = 1.0 / np.sqrt(p)
    rng = np.random.default_rng(4)
    gre, gim = rng.standard_normal(p), rng.standard_normal(p)
    want = _all_points_adjoint(ev, eta, gre, gim, precision)
    lib = _lib.load()
    _lib.check(lib.shk_set_symmetric_scratch_limit(2 << 20))
    try:
        got = _all_points_adjoint(ev, eta, gre, gim, precision)
    finally:
        _lib.check(lib.shk_set_symmetric_scratch_limit(0))
    assert np.array_equal(got[0], want[0]) and got[1] == want[1]
