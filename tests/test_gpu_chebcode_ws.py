"""ChebCode one-shot transforms in a caller's workspace (no host wait, no
allocation): bit for bit the one-shot entries' results on ascending input,
captured whole in a HIP graph and replayed on new spectra of the same size,
NaN outputs and a status word on input that is not ascending, and the
argument checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRESETS = ("chebcode", "chebcode_fast", "chebcode_xtreme", "chebcode_balanced")


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _ws(shk, p, method):
    import torch
    return torch.empty(shk.chebcode_workspace_bytes(p, method), dtype=torch.uint8, device="cuda")


def _spectrum(kind, p):
    from shrinkers_amd import spectra as S
    if kind == "mp":
        return S.mp_inverse_cdf_spectrum(p, c=0.25) if p >= 8 else np.sort(
            np.random.default_rng(p).uniform(0.2, 3.0, p))
    if kind == "geometric":  # midpoint splits peel one value per level to the depth cap
        return np.sort(1.3 ** (-np.arange(float(p))))
    if kind == "duplicates":
        return np.sort(np.repeat(np.linspace(0.5, 2.0, p // 37 + 1), 37)[:p])
    if kind == "clusters":
        centres = np.linspace(1.0, 50.0, p // 100 + 1)
        return np.sort(np.concatenate([c + 0.01 * 1.3 ** (-np.arange(100.0)) for c in centres])[:p])
    raise ValueError(kind)


def _equal(got, want, what):
    import torch
    torch.cuda.synchronize()
    for g, w in zip(got, want):
        assert torch.equal(g, w), what


@pytest.mark.parametrize("preset", PRESETS)
@pytest.mark.parametrize("p", [1, 2, 3, 33, 1000, 20_000, 50_000, 200_000])
def test_ws_equals_one_shot(shk, preset, p):
    ev = _cuda(_spectrum("mp", p))
    eta = 1.0 / np.sqrt(p)
    ws = _ws(shk, p, preset)
    print(f"{preset} p={p} workspace {ws.numel()} bytes", flush=True)
    for scaled in (True, False):
        want = shk.all_points_dev(ev, eta, preset, scaled=scaled)
        got = shk.chebcode_all_points_ws_dev(ev, eta, preset, ws, scaled=scaled)
        _equal(got, want, f"{preset} p={p} scaled={scaled}")
    assert shk.chebcode_workspace_status(ws)


@pytest.mark.parametrize("preset", ["chebcode", "chebcode_fast"])
@pytest.mark.parametrize("kind,p", [("geometric", 2000), ("duplicates", 25_900),
                                    ("clusters", 20_000), ("mp", 1_000_000)])
def test_ws_equals_one_shot_deep_and_large(shk, preset, kind, p):
    """Trees that end inside the top walk at the depth cap, heavy duplicates
    (index-midpoint splits), long gaps between clusters, and p = 1e6."""
    ev = _cuda(_spectrum(kind, p))
    eta = 1.0 / np.sqrt(p)
    ws = _ws(shk, p, preset)
    print(f"{preset} {kind} p={p}", flush=True)
    want = shk.all_points_dev(ev, eta, preset)
    got = shk.chebcode_all_points_ws_dev(ev, eta, preset, ws)
    _equal(got, want, f"{preset} {kind} p={p}")


@pytest.mark.parametrize("preset", ["chebcode", "chebcode_fast"])
@pytest.mark.parametrize("p,nq", [(50_000, 1), (50_000, 200), (50_000, 5000), (3000, 4000),
                                  (20_000, 30_000)])
def test_ws_at_points_equals_one_shot(shk, preset, p, nq):
    """At-points trees (grid_leaf_cap's leaf size, wide leaves for few
    points), sparse and dense walks."""
    ev = _spectrum("mp", p)
    rng = np.random.default_rng(p + nq)
    q = _cuda(np.sort(rng.uniform(ev[0] - 0.3, ev[-1] + 0.3, nq)))
    ev = _cuda(ev)
    eta = 1.0 / np.sqrt(p)
    ws = _ws(shk, p, preset)
    print(f"{preset} p={p} nq={nq}", flush=True)
    want = shk.at_points_dev(q, ev, eta, preset)
    got = shk.chebcode_at_points_ws_dev(q, ev, eta, preset, ws)
    _equal(got, want, f"{preset} at points p={p} nq={nq}")


@pytest.mark.parametrize("preset", ["chebcode", "chebcode_fast"])
def test_ws_captured_and_replayed_on_new_spectra(shk, preset):
    """Both workspace entries captured in one HIP graph (torch.cuda.graph,
    global capture mode) on fixed buffers, then replayed on new eigenvalues
    copied into the same buffer: spectra whose trees differ in node count,
    frontier and depth.  Eager one-shot calls run between the replays, and
    each replay equals the eager one-shot results on the same data bit for
    bit."""
    import torch
    p, nq = 20_000, 300
    specs = [_spectrum("mp", p), _spectrum("clusters", p), _spectrum("duplicates", p),
             np.sort(np.concatenate([_spectrum("geometric", 1500),
                                     np.linspace(2.0, 9.0, p - 1500)]))]
    ev = _cuda(specs[0])
    q = _cuda(np.linspace(0.0, 10.0, nq))
    eta = 1.0 / np.sqrt(p)
    ws = _ws(shk, p, preset)
    ws2 = _ws(shk, p, preset)
    out = (torch.empty_like(ev), torch.empty_like(ev))
    qout = (torch.empty_like(q), torch.empty_like(q))
    # warm the library (device set-up) outside the capture
    shk.chebcode_all_points_ws_dev(ev, eta, preset, ws)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        shk.chebcode_all_points_ws_dev(ev, eta, preset, ws, out=out, stream=s)
        shk.chebcode_at_points_ws_dev(q, ev, eta, preset, ws2, out=qout, stream=s)
    for rep, spec in enumerate(specs + specs[:1]):
        ev.copy_(_cuda(spec))
        want = shk.all_points_dev(ev, eta, preset)
        qwant = shk.at_points_dev(q, ev, eta, preset)
        for t in out + qout:
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"{preset} replay {rep}", flush=True)
        _equal(out, want, f"{preset} replay {rep} all points")
        _equal(qout, qwant, f"{preset} replay {rep} at points")


def test_ws_unsorted_input_gives_nan_and_status(shk):
    import torch
    p = 5000
    ev = _spectrum("mp", p)
    sh = np.random.default_rng(3).permutation(ev)
    ws = _ws(shk, p, "chebcode_fast")
    eta = 1.0 / np.sqrt(p)
    r, i = shk.chebcode_all_points_ws_dev(_cuda(sh), eta, "chebcode_fast", ws)
    torch.cuda.synchronize()
    assert torch.isnan(r).all() and torch.isnan(i).all()
    assert not shk.chebcode_workspace_status(ws)
    q = _cuda(np.linspace(0.0, 3.0, 50))
    r, i = shk.chebcode_at_points_ws_dev(q, _cuda(sh), eta, "chebcode", _ws(shk, p, "chebcode"))
    torch.cuda.synchronize()
    assert torch.isnan(r).all() and torch.isnan(i).all()
    # the same workspace on ascending input: the transform again
    r, i = shk.chebcode_all_points_ws_dev(_cuda(ev), eta, "chebcode_fast", ws)
    want = shk.all_points_dev(_cuda(ev), eta, "chebcode_fast")
    _equal((r, i), want, "after unsorted")
    assert shk.chebcode_workspace_status(ws)


def test_ws_argument_checks(shk):
    import torch
    from shrinkers_amd import _lib
    p = 1000
    ev = _cuda(_spectrum("mp", p))
    ws = _ws(shk, p, "chebcode")
    cases = [
        (lambda: shk.chebcode_all_points_ws_dev(ev, 0.03, "chebcode", ws[:-256]),
         _lib.SHK_ERR_BAD_ARG),
        (lambda: shk.chebcode_all_points_ws_dev(ev, 0.03, "chebcode", ws[8:]),
         _lib.SHK_ERR_BAD_ARG),
        (lambda: shk.chebcode_all_points_ws_dev(ev, 0.03, "blocked", ws), _lib.SHK_ERR_BAD_METHOD),
        (lambda: shk.chebcode_all_points_ws_dev(ev, 0.03, "speed_auto", ws),
         _lib.SHK_ERR_BAD_METHOD),
        (lambda: shk.chebcode_all_points_ws_dev(ev, -1.0, "chebcode", ws), _lib.SHK_ERR_BAD_ETA),
        (lambda: shk.chebcode_workspace_bytes(p, "auto"), _lib.SHK_ERR_BAD_METHOD),
    ]
    for k, (call, status) in enumerate(cases):
        with pytest.raises(_lib.ShkError) as err:
            call()
        assert err.value.status == status, (k, err.value.message)
    # n = 11 needs more than the FAST order's workspace
    assert shk.chebcode_workspace_bytes(p, "chebcode") > shk.chebcode_workspace_bytes(
        p, "chebcode_fast")
    r, i = shk.chebcode_all_points_ws_dev(ev, 0.03, "chebcode", ws)
    torch.cuda.synchronize()
    assert torch.isfinite(r).all() and torch.isfinite(i).all()


def test_ws_two_streams_at_once(shk):
    """Two workspaces on two streams with no synchronisation between the
    calls (different spectra, sizes and presets, several calls each): every
    result equals the one-shot entry's on the same input bit for bit."""
    import torch
    cases = [("chebcode_fast", _spectrum("mp", 50_000)), ("chebcode", _spectrum("clusters", 30_000))]
    evs = [_cuda(ev) for _, ev in cases]
    wss = [_ws(shk, ev.size, m) for (m, ev) in cases]
    wants = [shk.all_points_dev(d, 1.0 / np.sqrt(d.numel()), m) for (m, _), d in zip(cases, evs)]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [[], []]
    for rep in range(3):
        for k, ((m, _), d) in enumerate(zip(cases, evs)):
            with torch.cuda.stream(streams[k]):
                outs[k].append(shk.chebcode_all_points_ws_dev(
                    d, 1.0 / np.sqrt(d.numel()), m, wss[k], stream=streams[k]))
    torch.cuda.synchronize()
    for k in range(2):
        for rep, got in enumerate(outs[k]):
            _equal(got, wants[k], f"stream {k} call {rep}")
