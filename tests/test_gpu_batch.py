"""Batched persistent launch: every spectrum equals its own all-points
transform.  The captured graph is replayed several times with eager work
and synchronisations between replays; every replay is checked against the
FP64 reference and the replays against each other (bitwise)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="batched launch switched off: its first GPU run (the tuning pass "
                            "of test_batch_matches_oracle_across_replays) did not finish and no "
                            "cause has been shown; shk_batch_create returns SHK_ERR_UNSUPPORTED"),
]


def small_batch(seed=0):
    rng = np.random.default_rng(seed)
    from shrinkers_amd import spectra as S
    sizes = [1, 2, 63, 64, 65, 127, 129, 500, 1000, 2047, 3000] + list(rng.integers(3, 3000, 13))
    spectra, etas = [], []
    for k, p in enumerate(sizes):
        c = float(rng.uniform(0.05, 0.5))
        ev = S.mp_inverse_cdf_spectrum(int(p), c=c, seed=k, spikes=()) if p >= 8 \
            else np.sort(rng.uniform(0.3, 3.0, int(p)))
        spectra.append(ev)
        etas.append(float(rng.uniform(0.1, 1.0)) / np.sqrt(p))
    return spectra, etas


def test_batch_matches_oracle_across_replays(shk, orc):
    import torch
    from shrinkers_amd.batch import BatchedStieltjes
    spectra, etas = small_batch()
    b = BatchedStieltjes(spectra, etas)
    r, w, ms = b.tune()
    print(f"tuned: {r} queries/lane, {w} workgroups/CU, {ms:.3f} ms", flush=True)
    refs = []
    for ev, eta in zip(spectra, etas):
        rr, ri = orc.tiled_parallel(ev, eta)
        refs.append((rr / ev.size, ri / ev.size))
    first = None
    for rep in range(3):
        b.run()
        # eager work between replays, on the same stream
        d = torch.from_numpy(spectra[8]).cuda()
        er, ei = shk.all_points_dev(d, etas[8], "blocked", precision="f32c")
        torch.cuda.synchronize()
        out = b.results()
        for k, ((gr, gi), (rr, ri)) in enumerate(zip(out, refs)):
            err = rel_l2(gr, gi, rr, ri)
            assert err <= 1e-6, (rep, k, spectra[k].size, err)
        e8 = rel_l2(er.cpu().numpy(), ei.cpu().numpy(), *refs[8])
        assert e8 <= 1e-6
        flat = np.concatenate([np.concatenate(o) for o in out])
        if first is None:
            first = flat
        else:
            np.testing.assert_array_equal(flat, first)
    s0, s1 = b.spectrum_times_us()
    assert np.all(s0 >= 0) and np.all(s1 >= s0)


def test_batch_unscaled_and_configs(shk, orc):
    from shrinkers_amd.batch import BatchedStieltjes
    spectra, etas = small_batch(seed=3)
    b = BatchedStieltjes(spectra, etas, scaled=False)
    for r, w in ((1, 8), (2, 16), (4, 32)):
        b.set_config(r, w)
        b.run()
        for k, (gr, gi) in enumerate(b.results()):
            rr, ri = orc.tiled_parallel(spectra[k], etas[k])
            assert rel_l2(gr, gi, rr, ri) <= 1e-6, (r, w, k)


def test_config4_workload_sample(shk):
    """The 512-spectrum workload: tuned launch, 24 spectra checked against
    the GPU FP64 exact kernel."""
    import torch
    from shrinkers_amd import spectra as S
    from shrinkers_amd.batch import BatchedStieltjes
    spectra, meta = S.batched_workload(n=512, seed=0)
    etas = [1.0 / np.sqrt(s.size) for s in spectra]
    b = BatchedStieltjes(spectra, etas)
    b.tune()
    b.run()
    out = b.results()
    rng = np.random.default_rng(1)
    worst = 0.0
    for k in rng.choice(len(spectra), 24, replace=False):
        d = torch.from_numpy(spectra[k]).cuda()
        er, ei = shk.all_points_dev(d, etas[k], "blocked", precision="f64")
        torch.cuda.synchronize()
        worst = max(worst, rel_l2(out[k][0], out[k][1], er.cpu().numpy(), ei.cpu().numpy()))
    print(f"config 4 sample: worst rel-L2 vs FP64 {worst:.3e}", flush=True)
    assert worst <= 1e-6
