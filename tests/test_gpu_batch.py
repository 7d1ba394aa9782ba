"""Batched launch (shk_batch_*): the segmented all-points kernel over a fixed
set of spectra, captured in a hipGraph and replayed.  The graph is replayed
several times with eager work and synchronisations between the replays, and
every replay is checked against the eager single-spectrum results: bit for
bit at the default tile height (r = 2), within the FP32-compensated bound at
r = 1 and 4, and against the FP64 oracle.  r = 8 (the symmetric kernel, in
groups that fit the symmetric scratch limit) equals the single-spectrum
symmetric transform bit for bit."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def small_batch(seed=0):
    rng = np.random.default_rng(seed)
    from shrinkers_amd import spectra as S
    sizes = [1, 2, 63, 64, 65, 127, 129, 255, 257, 500, 1000, 2047, 3000] + \
        list(rng.integers(3, 3000, 11))
    spectra, etas = [], []
    for k, p in enumerate(sizes):
        c = float(rng.uniform(0.05, 0.5))
        ev = S.mp_inverse_cdf_spectrum(int(p), c=c, seed=k, spikes=()) if p >= 8 \
            else np.sort(rng.uniform(0.3, 3.0, int(p)))
        spectra.append(ev)
        etas.append(float(rng.uniform(0.1, 1.0)) / np.sqrt(p))
    return spectra, etas


def singles(shk, spectra, etas, scaled=True):
    import torch
    out = []
    for ev, eta in zip(spectra, etas):
        re, im = shk.all_points_dev(torch.from_numpy(ev).cuda(), eta, "blocked",
                                    precision="f32c", scaled=scaled)
        torch.cuda.synchronize()
        out.append((re.cpu().numpy(), im.cpu().numpy()))
    return out


@pytest.mark.parametrize("scaled", [True, False])
def test_replays_equal_single_calls(shk, orc, scaled):
    import torch
    from shrinkers_amd.batch import BatchedStieltjes
    spectra, etas = small_batch()
    want = singles(shk, spectra, etas, scaled)
    b = BatchedStieltjes(spectra, etas, scaled=scaled)
    for rep in range(4):
        b.run()
        # eager work on the same stream between replays, then a synchronisation
        d = torch.from_numpy(spectra[9]).cuda()
        er, ei = shk.all_points_dev(d, etas[9], "blocked", precision="f32c", scaled=scaled)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(er.cpu().numpy(), want[9][0])
        for k, (gr, gi) in enumerate(b.results()):
            np.testing.assert_array_equal(gr, want[k][0], err_msg=f"replay {rep}, spectrum {k}")
            np.testing.assert_array_equal(gi, want[k][1], err_msg=f"replay {rep}, spectrum {k}")
    for k, (ev, eta) in enumerate(zip(spectra, etas)):
        rr, ri = orc.tiled_parallel(ev, eta)
        if scaled:
            rr, ri = rr / ev.size, ri / ev.size
        assert rel_l2(want[k][0], want[k][1], rr, ri) <= 1e-6
    s0, s1 = b.spectrum_times_us()
    assert np.all(s0 >= 0) and np.all(s1 >= s0) and s0.min() == 0.0


def test_tile_heights(shk, orc):
    """r = 1 and 4 use other tiles (their own centres and pair forms): within
    1e-6 of the FP64 oracle; switching back to r = 2 restores the bits."""
    import torch
    from shrinkers_amd.batch import BatchedStieltjes
    spectra, etas = small_batch(seed=3)
    want = singles(shk, spectra, etas)
    b = BatchedStieltjes(spectra, etas)
    for r in (1, 4, 2):
        b.set_config(r)
        for rep in range(2):
            b.run()
            torch.cuda.synchronize()
            out = b.results()
            for k, (gr, gi) in enumerate(out):
                rr, ri = orc.tiled_parallel(spectra[k], etas[k])
                n = spectra[k].size
                assert rel_l2(gr, gi, rr / n, ri / n) <= 1e-6, (r, k, n)
                if r == 2:
                    np.testing.assert_array_equal(gr, want[k][0])
                    np.testing.assert_array_equal(gi, want[k][1])
        s0, s1 = b.spectrum_times_us()
        assert np.all(s1 >= s0)
    with pytest.raises(Exception):
        b.set_config(3)


def test_config4_workload_tuned(shk):
    """The 512-spectrum workload: tuned launch, replayed; 24 spectra checked
    against the GPU FP64 exact kernel at the FP32-compensated bar."""
    import torch
    from shrinkers_amd import spectra as S
    from shrinkers_amd.batch import BatchedStieltjes
    spectra, meta = S.batched_workload(n=512, seed=0)
    etas = [1.0 / np.sqrt(s.size) for s in spectra]
    b = BatchedStieltjes(spectra, etas)
    r, ms = b.tune()
    print(f"tuned: {r} queries per lane, {ms:.3f} ms per batch", flush=True)
    assert r in (1, 2, 4, 8) and 0.0 < ms < 1000.0
    for _ in range(3):
        b.run()
    out = b.results()
    worst = 0.0
    for k in np.random.default_rng(1).choice(len(spectra), 24, replace=False):
        d = torch.from_numpy(spectra[k]).cuda()
        er, ei = shk.all_points_dev(d, etas[k], "blocked", precision="f64")
        torch.cuda.synchronize()
        worst = max(worst, rel_l2(out[k][0], out[k][1], er.cpu().numpy(), ei.cpu().numpy()))
    print(f"512-spectrum batch: worst rel-L2 vs FP64 {worst:.3e}", flush=True)
    assert worst <= 1e-6
    s0, s1 = b.spectrum_times_us()
    assert np.all(s1 > s0) and s1.max() < 1e6


def sym_batch(seed=5):
    """Spectra from 1 to 20 000 eigenvalues, several inside the all-points
    entry's symmetric range."""
    rng = np.random.default_rng(seed)
    from shrinkers_amd import spectra as S
    sizes = [1, 7, 511, 512, 513, 1500, 8192, 9000, 12289, 20000] + \
        list(rng.integers(2, 6000, 6))
    spectra, etas = [], []
    for k, p in enumerate(sizes):
        c = float(rng.uniform(0.05, 0.5))
        ev = S.mp_inverse_cdf_spectrum(int(p), c=c, seed=100 + k, spikes=(6.0,)) if p >= 8 \
            else np.sort(rng.uniform(0.3, 3.0, int(p)))
        spectra.append(ev)
        etas.append(float(rng.uniform(0.2, 1.0)) / np.sqrt(p))
    return spectra, etas


def test_symmetric_config_replays(shk, orc):
    """r = 8: the symmetric kernel over every spectrum's tile pairs.  Inside
    the all-points entry's symmetric range each result equals that entry's
    bit for bit; every result is within 1e-6 of FP64 (the oracle up to 3000
    eigenvalues, the GPU FP64 kernel above).  Replays with eager work and
    synchronisations between them give the same bits."""
    import torch
    from shrinkers_amd.batch import BatchedStieltjes
    spectra, etas = sym_batch()
    lo, hi = shk.symmetric_range("f32c")
    want = singles(shk, spectra, etas)
    b = BatchedStieltjes(spectra, etas)
    b.set_config(8)
    first = None
    for rep in range(3):
        b.run()
        d = torch.from_numpy(spectra[8]).cuda()
        er, ei = shk.all_points_dev(d, etas[8], "blocked", precision="f32c")
        torch.cuda.synchronize()
        np.testing.assert_array_equal(er.cpu().numpy(), want[8][0])
        out = b.results()
        if first is None:
            first = out
        for k, (gr, gi) in enumerate(out):
            np.testing.assert_array_equal(gr, first[k][0], err_msg=f"replay {rep}, spectrum {k}")
            np.testing.assert_array_equal(gi, first[k][1], err_msg=f"replay {rep}, spectrum {k}")
    for k, (ev, eta) in enumerate(zip(spectra, etas)):
        gr, gi = first[k]
        n = ev.size
        if lo <= n <= hi:
            np.testing.assert_array_equal(gr, want[k][0], err_msg=f"spectrum {k} (p = {n})")
            np.testing.assert_array_equal(gi, want[k][1], err_msg=f"spectrum {k} (p = {n})")
        if n <= 3000:
            rr, ri = orc.tiled_parallel(ev, eta)
            rr, ri = rr / n, ri / n
        else:
            d = torch.from_numpy(ev).cuda()
            er, ei = shk.all_points_dev(d, eta, "blocked", precision="f64")
            torch.cuda.synchronize()
            rr, ri = er.cpu().numpy(), ei.cpu().numpy()
        assert rel_l2(gr, gi, rr, ri) <= 1e-6, (k, n)
    s0, s1 = b.spectrum_times_us()
    assert np.all(s1 >= s0) and s0.min() == 0.0
    # back to the split-tile kernel's default: its bits again
    b.set_config(2)
    b.run()
    for k, (gr, gi) in enumerate(b.results()):
        if not lo <= spectra[k].size <= hi:
            np.testing.assert_array_equal(gr, want[k][0])
            np.testing.assert_array_equal(gi, want[k][1])


def test_symmetric_config_groups(shk):
    """A small symmetric scratch limit cuts the spectra into several groups
    (one launch and fold each): the same bits as one group, raw and scaled.
    A spectrum whose partials alone exceed the limit makes r = 8 refused, and
    the tuner then keeps to the split-tile heights."""
    from shrinkers_amd import _lib
    from shrinkers_amd.batch import BatchedStieltjes
    lib = _lib.load()
    spectra, etas = sym_batch(seed=9)
    for scaled in (True, False):
        b = BatchedStieltjes(spectra, etas, scaled=scaled)
        b.set_config(8)
        b.run()
        want = b.results()
        # the largest spectrum (20 000: 40 tiles) needs 12.8 MB of partials
        _lib.check(lib.shk_set_symmetric_scratch_limit(14 << 20))
        try:
            g = BatchedStieltjes(spectra, etas, scaled=scaled)
            g.set_config(8)
            for _ in range(2):
                g.run()
                for k, (gr, gi) in enumerate(g.results()):
                    np.testing.assert_array_equal(gr, want[k][0], err_msg=f"spectrum {k}")
                    np.testing.assert_array_equal(gi, want[k][1], err_msg=f"spectrum {k}")
            s0, s1 = g.spectrum_times_us()
            assert np.all(s1 >= s0)
        finally:
            _lib.check(lib.shk_set_symmetric_scratch_limit(0))
    _lib.check(lib.shk_set_symmetric_scratch_limit(4 << 20))
    try:
        t = BatchedStieltjes(spectra, etas)
        with pytest.raises(Exception, match="scratch limit"):
            t.set_config(8)
        r, ms = t.tune()
        assert r in (1, 2, 4) and ms > 0.0
    finally:
        _lib.check(lib.shk_set_symmetric_scratch_limit(0))
