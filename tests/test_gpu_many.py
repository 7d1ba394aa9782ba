"""Many spectra in one library call (shk_all_points_many_dev): every
spectrum's result equals its own single-spectrum transform bit for bit --
on the stream-pool path (the same kernels run on each) and on the segmented
launch (one grid over all spectra that reproduces the single-spectrum
summation order) -- and the FP64 oracle to the path's bar."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def small_batch(seed=0):
    rng = np.random.default_rng(seed)
    from shrinkers_amd import spectra as S
    sizes = [1, 2, 63, 64, 65, 127, 129, 500, 1000, 2047, 3000] + list(rng.integers(3, 3000, 13))
    sizes.insert(4, 0)   # an empty spectrum in the middle
    spectra, etas = [], []
    for k, p in enumerate(sizes):
        c = float(rng.uniform(0.05, 0.5))
        ev = S.mp_inverse_cdf_spectrum(int(p), c=c, seed=k, spikes=()) if p >= 8 \
            else np.sort(rng.uniform(0.3, 3.0, int(p)))
        spectra.append(ev)
        etas.append(float(rng.uniform(0.1, 1.0)) / np.sqrt(max(p, 1)))
    return spectra, etas


@pytest.mark.parametrize("precision", ["f32c", "f64"])
@pytest.mark.parametrize("scaled", [True, False])
@pytest.mark.parametrize("n_streams", [0, 1, 4])
def test_many_equals_single_calls(shk, orc, precision, scaled, n_streams):
    import torch
    from shrinkers_amd.batch import SpectrumBatch
    spectra, etas = small_batch()
    b = SpectrumBatch(spectra, etas)
    t = b.run("blocked", precision, scaled, n_streams, times=True)
    out = b.results()
    assert t.shape == (len(spectra), 2) and np.all(t[:, 1] >= t[:, 0]) and np.all(t >= 0)
    tol = 1e-6 if precision == "f32c" else 1e-13
    for k, (ev, eta) in enumerate(zip(spectra, etas)):
        if ev.size == 0:
            assert out[k][0].size == 0 and tuple(t[k]) == (0.0, 0.0)
            continue
        d = torch.from_numpy(ev).cuda()
        sr, si = shk.all_points_dev(d, eta, "blocked", precision=precision, scaled=scaled)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out[k][0], sr.cpu().numpy())
        np.testing.assert_array_equal(out[k][1], si.cpu().numpy())
        rr, ri = orc.tiled_parallel(ev, eta)
        if scaled:
            rr, ri = rr / ev.size, ri / ev.size
        assert rel_l2(out[k][0], out[k][1], rr, ri) <= tol, (k, ev.size)


def test_many_chebcode_and_empty_spectra(shk):
    import torch
    from shrinkers_amd.batch import all_points_many_dev
    spectra, etas = small_batch(seed=5)
    spectra = spectra[:6] + [np.zeros(0)] + spectra[6:10]
    etas = etas[:6] + [0.1] + etas[6:10]
    offsets = np.concatenate([[0], np.cumsum([s.size for s in spectra])]).astype(np.int64)
    ev = torch.from_numpy(np.concatenate(spectra)).cuda()
    re, im = all_points_many_dev(ev, offsets, etas, "chebcode_fast", "f64", True, 3)
    torch.cuda.synchronize()
    re, im = re.cpu().numpy(), im.cpu().numpy()
    for k, (s, eta) in enumerate(zip(spectra, etas)):
        if s.size == 0:
            continue
        sr, si = shk.all_points_dev(torch.from_numpy(s).cuda(), eta, "chebcode_fast")
        a, b = offsets[k], offsets[k + 1]
        np.testing.assert_array_equal(re[a:b], sr.cpu().numpy())
        np.testing.assert_array_equal(im[a:b], si.cpu().numpy())


def test_many_rejects_bad_arguments(shk):
    import torch
    from shrinkers_amd._lib import ShkError
    from shrinkers_amd.batch import all_points_many_dev
    ev = torch.ones(10, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        all_points_many_dev(ev, [0, 4, 9], [0.1, 0.1])          # does not end at numel
    with pytest.raises(ShkError):
        all_points_many_dev(ev, [0, 4, 10], [0.1, -1.0])        # bad eta
    with pytest.raises(ShkError):
        all_points_many_dev(ev, [0, 4, 10], [0.1, 0.1], n_streams=9)
    with pytest.raises(ShkError):
        all_points_many_dev(ev, [0, 4, 10], [0.1, 0.1], n_streams=-1)
    with pytest.raises(ShkError):
        all_points_many_dev(ev, [0, 6, 4, 10], [0.1, 0.1, 0.1])  # decreasing offsets


@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_segmented_large_spectra(shk, precision):
    """Spectra whose single-spectrum split length exceeds the LDS piece
    (p = 40 000: 768 sources per split; p = 100 003: 4 768) or divides it
    unevenly (p = 20 000: 192), behind a tiny one: the segmented launch
    equals the split-tile single-spectrum kernel bit for bit (raw sums
    through shk_exact_at_points_dev with the spectrum as its own query
    points; the all-points entry takes the symmetric kernel at these
    sizes)."""
    import torch
    from shrinkers_amd import _lib
    from shrinkers_amd import spectra as S
    from shrinkers_amd.batch import SpectrumBatch
    from shrinkers_amd.stieltjes import _PRECISIONS
    sizes = [20000, 3, 100003, 40000]
    spectra = [S.mp_inverse_cdf_spectrum(p, c=0.3, seed=p, spikes=(4.0,)) if p > 8
               else np.array([0.5, 1.0, 1.5]) for p in sizes]
    etas = [1.0 / np.sqrt(p) for p in sizes]
    b = SpectrumBatch(spectra, etas)
    b.run("blocked", precision, False, 0)
    out = b.results()
    for k, (ev, eta) in enumerate(zip(spectra, etas)):
        d = torch.from_numpy(ev).cuda()
        sr, si = torch.empty_like(d), torch.empty_like(d)
        assert _lib.load().shk_exact_at_points_dev(
            d.data_ptr(), d.numel(), d.data_ptr(), d.numel(), eta, _PRECISIONS[precision],
            sr.data_ptr(), si.data_ptr(), None) == _lib.SHK_OK
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out[k][0], sr.cpu().numpy(), err_msg=f"p={ev.size}")
        np.testing.assert_array_equal(out[k][1], si.cpu().numpy(), err_msg=f"p={ev.size}")


def test_config4_workload_sample(shk):
    """The 512-spectrum workload (p ~ U[500, 20000]): 24 spectra checked
    against the GPU FP64 exact kernel at the FP32-compensated bar."""
    import torch
    from shrinkers_amd import spectra as S
    from shrinkers_amd.batch import SpectrumBatch
    spectra, meta = S.batched_workload(n=512, seed=0)
    etas = [1.0 / np.sqrt(s.size) for s in spectra]
    b = SpectrumBatch(spectra, etas)
    b.run("blocked", "f32c", True, 4)
    out4 = b.results()
    b.run("blocked", "f32c", True, 0)
    out = b.results()
    for k in range(len(spectra)):   # segmented launch == stream-pool path, all 512
        np.testing.assert_array_equal(out[k][0], out4[k][0])
        np.testing.assert_array_equal(out[k][1], out4[k][1])
    worst = 0.0
    for k in np.random.default_rng(1).choice(len(spectra), 24, replace=False):
        d = torch.from_numpy(spectra[k]).cuda()
        er, ei = shk.all_points_dev(d, etas[k], "blocked", precision="f64")
        torch.cuda.synchronize()
        worst = max(worst, rel_l2(out[k][0], out[k][1], er.cpu().numpy(), ei.cpu().numpy()))
    print(f"512-spectrum sample: worst rel-L2 vs FP64 {worst:.3e}", flush=True)
    assert worst <= 1e-6
