"""Seeded random sweep of the GPU paths against the oracle.

Each case draws a spectrum (Marchenko-Pastur, uniform, heavy duplicates,
tight clusters or log-spaced values), a scale and shift, an order (sorted or
shuffled), a size and an eta, then checks, at the bars of the per-path tests:
- the all-points transform, FP64 (1e-13 normwise and per element against
  ||S||_inf) and FP32-compensated (1e-6), against the oracle's tiled sums;
- the at-points sums at random points (some on eigenvalues, some outside the
  spectrum), FP64 through the host entry (1e-13) and FP32-compensated through
  the device entry (1e-6);
- the ChebCode DEFAULT and FAST transforms against the oracle's ChebCode
  (1e-13 / 1e-6) up to 20 000 eigenvalues, and the workspace build of the
  sorted spectrum against the one-shot entry bit for bit;
- the at-points adjoint (dL/dlambda, dL/dx, dL/deta) against the oracle's FP64
  adjoint (1e-12 FP64, 5e-6 FP32-compensated) up to 8 000 eigenvalues.
The draws cover sizes that are not multiples of the kernels' tiles, both
sides of the grid-point and symmetric-kernel switches, and both sides of the
FP32 pair forms (tile widths small and large against eta).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

N_CASES = 36
KINDS = ("mp", "uniform", "duplicates", "clusters", "log")


def _draw(seed):
    from shrinkers_amd import spectra as S
    rng = np.random.default_rng(9000 + seed)
    kind = KINDS[seed % len(KINDS)]
    p = int(np.exp(rng.uniform(0.0, np.log(30_000))))
    if kind == "mp" and p >= 8:
        base = S.mp_inverse_cdf_spectrum(p, c=float(rng.uniform(0.05, 0.9)), seed=seed)
    elif kind == "duplicates":
        vals = rng.uniform(0.0, 1.0, max(1, p // 9 + 1))
        base = np.repeat(vals, 9)[:p]
    elif kind == "clusters":
        centres = rng.uniform(-3.0, 3.0, max(1, p // 50 + 1))
        base = (centres[:, None] + 1e-7 * rng.random((centres.size, 50))).ravel()[:p]
    elif kind == "log":
        base = np.exp(rng.uniform(-9.0, 3.0, p))
    else:
        base = rng.uniform(-1.0, 1.0, p)
    scale = 10.0 ** rng.uniform(-3.0, 3.0)
    shift = 0.0 if rng.random() < 0.5 else float(rng.normal()) * 5.0 * scale
    ev = base * scale + shift
    ev = np.sort(ev) if rng.random() < 0.5 else rng.permutation(ev)
    spread = float(np.ptp(ev)) if ev.size > 1 else 0.0
    unit = spread if spread > 0 else scale
    eta = unit / max(p, 1) * 10.0 ** rng.uniform(-1.5, 2.5)
    nq = int(np.exp(rng.uniform(0.0, np.log(3000))))
    lo, hi = float(ev.min()), float(ev.max())
    pad = max(spread, unit) * 0.3
    q = rng.uniform(lo - pad, hi + pad, nq)
    on = rng.random(nq) < 0.25
    q[on] = rng.choice(ev, int(on.sum()))
    return kind, ev, eta, q, rng


def _tag(seed, kind, ev, eta, q):
    return f"case {seed} {kind} p={ev.size} eta={eta:.3e} nq={q.size}"


def _close(got_re, got_im, ref_re, ref_im, tol, what):
    err = rel_l2(got_re, got_im, ref_re, ref_im)
    scale = max(np.max(np.abs(ref_re)), np.max(np.abs(ref_im)))
    emax = max(np.max(np.abs(got_re - ref_re)), np.max(np.abs(got_im - ref_im))) / scale
    print(f"{what}: rel-L2 {err:.3e} max/||S||inf {emax:.3e}", flush=True)
    assert err <= tol, f"{what}: rel-L2 {err:.3e} > {tol:.1e}"
    return emax


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_exact_cases(shk, orc, seed):
    import torch
    from shrinkers_amd import autograd as ag
    kind, ev, eta, q, _ = _draw(seed)
    tag = _tag(seed, kind, ev, eta, q)
    print(tag, flush=True)
    p = ev.size
    rr, ri = orc.tiled_parallel(ev, eta)
    rr, ri = rr / p, ri / p
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked")
    emax = _close(r, i, rr, ri, 1e-13, f"{tag} all-points f64")
    assert emax <= 1e-13, f"{tag}: all-points f64 max err {emax:.3e}"
    r, i = shk.compute_all_stieltjes(ev, eta, "blocked", precision="f32c")
    _close(r, i, rr, ri, 1e-6, f"{tag} all-points f32c")
    qr, qi = orc.at_points(q, ev, eta)
    r, i = shk.compute_stieltjes_at_points(q, ev, eta, "blocked")
    emax = _close(r, i, qr, qi, 1e-13, f"{tag} at-points f64")
    assert emax <= 1e-13, f"{tag}: at-points f64 max err {emax:.3e}"
    tq = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    tev = torch.from_numpy(np.ascontiguousarray(ev)).cuda()
    with torch.no_grad():
        gr, gi = ag.stieltjes_at_points(tq, tev, eta, "f32c")
    torch.cuda.synchronize()
    _close(gr.cpu().numpy(), gi.cpu().numpy(), qr, qi, 1e-6, f"{tag} at-points f32c")


@pytest.mark.parametrize("seed", range(0, N_CASES, 2))
def test_random_chebcode_cases(shk, orc, seed):
    kind, ev, eta, q, _ = _draw(seed)
    if ev.size > 20_000:
        ev = ev[:20_000]
    tag = _tag(seed, kind, ev, eta, q)
    print(tag, flush=True)
    p = ev.size
    import torch
    d = torch.from_numpy(np.ascontiguousarray(np.sort(ev))).cuda()
    for preset, tol in (("chebcode", 1e-13), ("chebcode_fast", 1e-6)):
        rr, ri = orc.chebcode_preset(ev, eta, preset)
        r, i = shk.compute_all_stieltjes(ev, eta, preset)
        _close(r, i, rr / p, ri / p, tol, f"{tag} {preset}")
        # the workspace build (ascending input) against the one-shot entry
        ws = torch.empty(shk.chebcode_workspace_bytes(p, preset), dtype=torch.uint8,
                         device="cuda")
        got = shk.chebcode_all_points_ws_dev(d, eta, preset, ws)
        want = shk.all_points_dev(d, eta, preset)
        torch.cuda.synchronize()
        assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1]), tag


def _rel(a, b):
    a, b = np.atleast_1d(np.asarray(a, np.float64)), np.atleast_1d(np.asarray(b, np.float64))
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


@pytest.mark.parametrize("seed", range(1, N_CASES, 3))
def test_random_adjoint_cases(shk, orc, seed):
    import torch
    from shrinkers_amd import autograd as ag
    kind, ev, eta, q, rng = _draw(seed)
    if ev.size > 8000:
        ev = ev[:8000]
    tag = _tag(seed, kind, ev, eta, q)
    print(tag, flush=True)
    gre, gim = rng.standard_normal(q.size), rng.standard_normal(q.size)
    rev, rq, re_ = orc.adjoint_at_points(q, ev, eta, gre, gim)
    for precision, tol in (("f64", 1e-12), ("f32c", 5e-6)):
        if precision == "f32c" and eta < 1e-9:
            continue  # FP32-compensated adjoints run FP64 below 1e-9
        tq = torch.from_numpy(np.ascontiguousarray(q)).cuda().requires_grad_(True)
        tev = torch.from_numpy(np.ascontiguousarray(ev)).cuda().requires_grad_(True)
        te = torch.tensor(float(eta), dtype=torch.float64, device="cuda", requires_grad=True)
        r, i = ag.stieltjes_at_points(tq, tev, te, precision)
        g_r = torch.from_numpy(gre).cuda()
        g_i = torch.from_numpy(gim).cuda()
        ((r * g_r).sum() + (i * g_i).sum()).backward()
        torch.cuda.synchronize()
        e_ev = _rel(tev.grad.cpu().numpy(), rev)
        e_q = _rel(tq.grad.cpu().numpy(), rq)
        e_eta = abs(float(te.grad.cpu()) - re_) / max(abs(re_), 1e-300)
        print(f"{tag} adjoint {precision}: lam {e_ev:.3e} x {e_q:.3e} eta {e_eta:.3e}",
              flush=True)
        assert e_ev <= tol and e_q <= tol and e_eta <= tol, f"{tag} adjoint {precision}"
