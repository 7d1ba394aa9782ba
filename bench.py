#!/usr/bin/env python3
"""Benchmark of the Stieltjes-transform hot path on one MI355X.

Workload (the metric's configuration): p = 50 000 eigenvalues, Marchenko-
Pastur c = 0.25, sigma^2 = 1, drawn by inverse-CDF sampling of seeded
uniforms, plus outliers at the forward-BBP images of spikes 12/7/4;
eta = 1/sqrt(p).  One step = one all-points transform s(z_i), z_i = lam_i +
i*eta, for all p eigenvalues (p^2 pair evaluations), with the eigenvalues
already resident in HBM.

Prints ONE JSON line: the FP32-compensated kernel's pair-evals/s as
``value``, the FP64 exact kernel beside it, their relative error against the
FP64 CPU oracle, a roofline object for the dominant kernel (timed with HIP
events on the stream it runs on) and the CPU baseline (the oracle's
restatement of the reference's parallel tiled kernel on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Stieltjes pair-evals/sec at p=50k: FP32-comp (≤1e-6) vs FP64 exact, 1 GPU"
FLOPS_PER_PAIR = 7          # sub, fma(2), rcp, fma(2), add
PEAK_FP32_TFLOPS = 157.3    # MI355X FP32 vector (published)
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector (published)
# HBM bytes per launch of the kernel the headline runs at p = 50 000 (sym_f32c_kernel),
# from one rocprofv3 pass per counter (FETCH_SIZE, WRITE_SIZE) scaled by the readings for a
# 256 MiB device copy (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000): 1681.0 KiB fetched (the
# eigenvalues, once per XCD L2) and 76562.5 KiB written (the FP64 per-tile partials,
# 2 x 98 tiles x 50 000 x 8 B).
TRAFFIC_P50000 = (1681.0 * 2.000 + 76562.5 * 1.000) * 1024
# Fraction of the kernel's time the vector ALUs issue (rocprofv3 derived counter VALUBusy
# = SQ_ACTIVE_INST_VALU / (CU_NUM * GRBM_GUI_ACTIVE)) for sym_f32c_kernel at p = 50 000,
# from `rocprofv3 --pmc VALUBusy VALUUtilization -- python3 bench.py --kernels`; its
# instruction mix is 3.99 VALU instructions per ordered pair (SQ_INSTS_VALU).
VALU_BUSY_P50000 = 0.990
VALU_INSTS_PER_PAIR_P50000 = 3.99


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(fn, steps: int, warmup: int, sync) -> float:
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return time.perf_counter() - t0


def kernel_avg_ms(shk, fn, launches: int, sync) -> float:
    """Average duration of the library's dominant kernel, from HIP events that
    the library records around it on the stream it is launched on."""
    from shrinkers_amd import stieltjes as st
    st.profile_reset()
    st.profile_enable(True)
    for _ in range(launches):
        fn()
    sync()
    st.profile_enable(False)
    ms, n = st.profile_read()
    st.profile_reset()
    return ms / max(n, 1)


def cpu_baseline(ev: np.ndarray, eta: float, budget_s: float):
    """The oracle's C restatement of the reference parallel tiled kernel
    (cacheblock.rs:1454-1513) on all host cores, timed with the reference's
    bench_ms policy (benches/support/mod.rs:90-112), plus one core on a
    bounded sample."""
    import oracle
    oracle.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    p = ev.size

    def once():
        return oracle.tiled_parallel(ev, eta, threads)

    out = once()  # warm-up (also the FP64 CPU result used for the error figures)
    t0 = time.perf_counter()
    once()
    probe = time.perf_counter() - t0
    reps = 15 if probe < 2e-3 else (7 if probe < 20e-3 else 3)
    reps = max(1, min(reps, int(budget_s / max(probe, 1e-6))))
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        once()
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    # one core: sample of rows against all p sources
    rows = max(1, min(p, int(2.0 / max(med * threads / p, 1e-9))))
    sample = ev[:rows].copy()
    t0 = time.perf_counter()
    for x in sample:
        oracle.autovec_sum(x, ev, eta)
    t1 = time.perf_counter() - t0

    def median_time(fn, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    extra = {}
    # sequential symmetric sweep (cacheblock.rs:841-954), one core: each unordered pair once
    t_sym = median_time(lambda: oracle.symmetric_sweep(ev, eta), 1)
    extra["symmetric_sweep_1core"] = {"ms_per_transform": t_sym * 1e3,
                                      "pair_evals_per_s": p * p / t_sym}
    # ChebCode FAST build + traversal (chebcode.rs:509-828), all cores
    oracle.chebcode_preset(ev, eta, "chebcode_fast", threads)
    t_cf = median_time(lambda: oracle.chebcode_preset(ev, eta, "chebcode_fast", threads), 7)
    extra["chebcode_fast_all_cores"] = {"ms_per_transform": t_cf * 1e3}
    # the reference's recorded benchmark spectrum (benches/support/mod.rs:65-82) at p = 50 000
    hs = oracle.harness_spectrum(p)
    he = 1.0 / np.sqrt(p)
    oracle.tiled_parallel(hs, he, threads)
    t_hs = median_time(lambda: oracle.tiled_parallel(hs, he, threads), 3)
    extra["harness_spectrum_tiled_all_cores"] = {"ms_per_transform": t_hs * 1e3,
                                                 "pair_evals_per_s": p * p / t_hs}
    return out, {
        "value": p * p / med,
        "unit": "pair-evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full p={p} transform (oracle C/OpenMP restatement of the reference's "
                  f"parallel tiled kernel, x86-64-v3 + FMA), median of {reps}",
        "ms_per_transform": med * 1e3,
        "single_core": {"value": rows * p / t1, "unit": "pair-evals/s",
                        "sample": f"{rows} rows x {p} sources, 1 thread"},
        "cpu_model": _cpu_model(),
        "extra": extra,
    }


def box_peaks() -> dict:
    """CU count and peak engine clock of the gfx950 agent from rocminfo, and
    the vector peaks they imply (64 lanes per CU per clock, FMA = 2 FLOPs;
    packed FP32 doubles the FP32 figure)."""
    import subprocess
    try:
        txt = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
    except (OSError, subprocess.SubprocessError):
        return {}
    cu = mhz = None
    for block in txt.split("Agent ")[1:]:
        if "gfx950" not in block:
            continue
        for line in block.splitlines():
            s = line.strip()
            if s.startswith("Compute Unit:"):
                cu = int(s.split(":")[1])
            elif s.startswith("Max Clock Freq. (MHz):"):
                mhz = int(s.split(":")[1])
        break
    if not cu or not mhz:
        return {}
    fma_tflops = cu * 64 * 2 * mhz * 1e6 / 1e12
    return {"compute_units": cu, "max_clock_mhz": mhz,
            "fp32_vector_tflops": 2 * fma_tflops, "fp32_vector_nonpacked_tflops": fma_tflops,
            "fp64_vector_tflops": fma_tflops}


def many_leg(shk, sync) -> dict:
    """512 independent spectra of uneven p in one library call
    (shk_all_points_many_dev): the segmented launch (one grid over every
    spectrum's query tiles, FP32-compensated and FP64) and the stream-pool
    path (the single-spectrum kernels, largest spectra first, over 1 or 4
    library streams); then the same spectra as a prepared batch
    (shk_batch_*: device table, launch captured in a hipGraph and replayed,
    configuration tuned: split-tile heights 1, 2, 4 or the symmetric kernel).  Per-spectrum times are device events (stream pool)
    or per-workgroup device wall-clock stamps (segmented launch)."""
    import torch
    from shrinkers_amd import spectra
    from shrinkers_amd.batch import SpectrumBatch
    sp, meta = spectra.batched_workload(n=512, seed=0)
    etas = [1.0 / np.sqrt(s.size) for s in sp]
    b = SpectrumBatch(sp, etas)
    pairs = float(np.sum(np.asarray([s.size for s in sp], dtype=np.float64) ** 2))
    out = {"workload": "512 spectra, p~U[500,20000], c~U[0.05,0.5], sigma2 log-U[0.5,2], 0-5 "
                       "spikes U[2,20]*sigma2 at forward-BBP images; eta=1/sqrt(p); "
                       "exact all-points sums per spectrum",
           "pair_evals_per_batch": pairs}
    for label, ns, prec in (("segmented", 0, "f32c"), ("segmented_f64", 0, "f64"),
                            ("streams_1", 1, "f32c"), ("streams_4", 4, "f32c")):
        for _ in range(2):
            b.run("blocked", prec, True, ns)
        sync()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run("blocked", prec, True, ns)
        sync()
        ms = (time.perf_counter() - t0) / reps * 1e3
        t = b.run("blocked", prec, True, ns, times=True)
        rec = {"precision": prec, "ms_per_batch": ms, "spectra_per_s": 512 / (ms * 1e-3),
               "pair_evals_per_s": pairs / (ms * 1e-3)}
        # segmented: device wall clock per workgroup; stream pool: events
        dur = t[:, 1] - t[:, 0]
        rec.update({"spectrum_ms_p50": float(np.percentile(dur, 50)),
                    "spectrum_ms_p99": float(np.percentile(dur, 99)),
                    "completion_ms_p50": float(np.percentile(t[:, 1], 50)),
                    "completion_ms_p99": float(np.percentile(t[:, 1], 99)),
                    "last_completion_ms": float(t[:, 1].max())})
        out[label] = rec
    # the same spectra as a prepared batch: table on the device, the launch
    # captured once in a hipGraph, tile height tuned on this batch
    from shrinkers_amd.batch import BatchedStieltjes
    gb = BatchedStieltjes(sp, etas)

    def replay_ms(reps=10):
        for _ in range(2):
            gb.run()
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            gb.run()
        sync()
        return (time.perf_counter() - t0) / reps * 1e3

    # every configuration: split-tile heights 1, 2, 4 and the symmetric kernel
    per_config = {}
    for r in (1, 2, 4, 8):
        gb.set_config(r)
        per_config[str(r)] = replay_ms(5)
    r_best, ms_tune = gb.tune()
    ms = replay_ms()
    s0, s1 = gb.spectrum_times_us()
    dur = (s1 - s0) * 1e-3
    gres = gb.results()
    out["hipgraph_replay"] = {
        "precision": "f32c", "tuned_queries_per_lane": r_best, "tune_best_ms": ms_tune,
        "ms_per_batch_by_config": per_config,
        "ms_per_batch": ms, "spectra_per_s": 512 / (ms * 1e-3),
        "pair_evals_per_s": pairs / (ms * 1e-3),
        "spectrum_ms_p50": float(np.percentile(dur, 50)),
        "spectrum_ms_p99": float(np.percentile(dur, 99)),
        "completion_ms_p50": float(np.percentile(s1 * 1e-3, 50)),
        "completion_ms_p99": float(np.percentile(s1 * 1e-3, 99)),
        "last_completion_ms": float(s1.max() * 1e-3)}
    del gb
    b.run("blocked", "f32c", True, 0)
    res = b.results()
    worst = 0.0
    for k in np.random.default_rng(1).choice(len(sp), 8, replace=False):
        d = torch.from_numpy(sp[k]).cuda()
        er, ei = shk.all_points_dev(d, etas[k], "blocked", precision="f64")
        sync()
        er, ei = er.cpu().numpy(), ei.cpu().numpy()
        num = np.sum((res[k][0] - er) ** 2 + (res[k][1] - ei) ** 2)
        worst = max(worst, float(np.sqrt(num / np.sum(er ** 2 + ei ** 2))))
    out["worst_rel_l2_vs_fp64_sample8"] = worst
    worst = 0.0  # the tuned prepared batch, same sample
    for k in np.random.default_rng(1).choice(len(sp), 8, replace=False):
        d = torch.from_numpy(sp[k]).cuda()
        er, ei = shk.all_points_dev(d, etas[k], "blocked", precision="f64")
        sync()
        er, ei = er.cpu().numpy(), ei.cpu().numpy()
        num = np.sum((gres[k][0] - er) ** 2 + (gres[k][1] - ei) ** 2)
        worst = max(worst, float(np.sqrt(num / np.sum(er ** 2 + ei ** 2))))
    out["hipgraph_replay"]["worst_rel_l2_vs_fp64_sample8"] = worst
    return out


def workspace_times(shk, d, eta, method, sync, reps=10) -> dict:
    """The ChebCode one-shot transform built in a caller's workspace (no
    host wait, no allocation): eager calls, and the call captured in a HIP
    graph and replayed."""
    import torch
    nbytes = shk.chebcode_workspace_bytes(d.numel(), method)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    out = (torch.empty_like(d), torch.empty_like(d))
    for _ in range(3):
        shk.chebcode_all_points_ws_dev(d, eta, method, ws, out=out)
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        shk.chebcode_all_points_ws_dev(d, eta, method, ws, out=out)
    sync()
    eager = (time.perf_counter() - t0) / reps * 1e3
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        shk.chebcode_all_points_ws_dev(d, eta, method, ws, out=out, stream=s)
    for _ in range(3):
        g.replay()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    sync()
    graph = (time.perf_counter() - t0) / reps * 1e3
    del g
    return {"ms_workspace_one_shot": eager, "ms_workspace_graph_replay": graph,
            "workspace_bytes": nbytes}


def chebcode_leg(shk, sync) -> dict:
    """ChebCode at p = 2e5 and 1e6 (FAST, DEFAULT and the f64 n = 12 point),
    error against the GPU FP64 exact kernel, build+evaluate and evaluate-only times."""
    import torch
    from shrinkers_amd import spectra
    res = {}
    for p, c in ((200_000, 0.1), (200_000, 0.5), (1_000_000, 0.1), (1_000_000, 0.5)):
        ev = spectra.mp_inverse_cdf_spectrum(p, c=c)
        eta = 1.0 / np.sqrt(p)
        d = torch.from_numpy(ev).cuda()
        er, ei = shk.all_points_dev(d, eta, "blocked", precision="f64", scaled=False)
        sync()
        er, ei = er.cpu().numpy(), ei.cpu().numpy()
        den = np.sum(er ** 2 + ei ** 2)
        row = {}
        for name, (theta, n, leaf, f32) in (("chebcode_fast", (1.0, 8, 32, True)),
                                            ("chebcode", (0.5, 11, 32, False)),
                                            ("chebcode_n12", (0.5, 12, 32, False))):
            ts = []
            for _ in range(7):
                sync()
                t0 = time.perf_counter()
                b = shk.ChebCodeBatch.build_dev(d, theta, n, leaf, f32)
                r, i = b.evaluate_dev(eta)
                sync()
                ts.append(time.perf_counter() - t0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                b.evaluate_dev(eta, out=(r, i))
            e1.record()
            sync()
            err = float(np.sqrt(np.sum((r.cpu().numpy() - er) ** 2 + (i.cpu().numpy() - ei) ** 2) / den))
            ms_eval = e0.elapsed_time(e1) / 5
            near, far = b.count_work(eta)
            flops = 7.0 * near + 8.0 * far   # pair-eval 7 FLOPs, far-field panel term 8
            row[name] = {"ms_build_plus_eval": 1e3 * float(np.median(ts)),
                         "ms_eval": ms_eval, "rel_l2_vs_fp64_exact": err,
                         "nodes": b.node_count, "near_pairs": near, "far_terms": far,
                         "eval_tflops": flops / (ms_eval * 1e-3) / 1e12,
                         "near_field_tflops_lower_bound": 7.0 * near / (ms_eval * 1e-3) / 1e12}
            if not f32:
                # local expansions (M2L into target nodes, L2P per query)
                # instead of the per-query panel sums
                lr, li = b.evaluate_local_dev(eta)
                e0.record()
                for _ in range(5):
                    b.evaluate_local_dev(eta, out=(lr, li))
                e1.record()
                sync()
                row[name]["ms_eval_local"] = e0.elapsed_time(e1) / 5
                row[name]["rel_l2_local_vs_fp64_exact"] = float(np.sqrt(
                    np.sum((lr.cpu().numpy() - er) ** 2 + (li.cpu().numpy() - ei) ** 2) / den))
            if name in ("chebcode_fast", "chebcode"):
                # the one-shot transform (compute_all_stieltjes with the preset):
                # build, evaluate and release in one call
                for _ in range(3):
                    shk.all_points_dev(d, eta, name)
                sync()
                t0 = time.perf_counter()
                for _ in range(10):
                    shk.all_points_dev(d, eta, name)
                sync()
                row[name]["ms_transform_one_shot"] = (time.perf_counter() - t0) / 10 * 1e3
                row[name].update(workspace_times(shk, d, eta, name, sync))
            if name == "chebcode":
                # evaluate_many: four etas, two per traversal (host outputs)
                etas = [eta * f for f in (0.5, 1.0, 2.0, 4.0)]
                b.evaluate_many(etas)
                t0 = time.perf_counter()
                for _ in range(3):
                    b.evaluate_many(etas)
                row[name]["ms_evaluate_many_4eta_host"] = (time.perf_counter() - t0) / 3 * 1e3
                # the same on device buffers (no host copies)
                mo = b.evaluate_many_dev(etas)
                sync()
                t0 = time.perf_counter()
                for _ in range(5):
                    b.evaluate_many_dev(etas, out=mo)
                sync()
                row[name]["ms_evaluate_many_4eta_dev"] = (time.perf_counter() - t0) / 5 * 1e3
            del b
        # CPU comparators on this host's cores: the oracle's restatements of
        # the reference's ChebCode FAST (chebcode.rs:509-828) and of its
        # coarse-grid NUFFT (nufft.rs, fft5.rs; the reference's auto pick
        # above p = 50 000, pareto_autogen.rs:65-71), errors against the GPU
        # FP64 exact sums
        row["cpu"] = cpu_config3(ev, eta, er, ei)
        res[f"p={p},c={c}"] = row
    return res


def cpu_config3(ev, eta, er, ei) -> dict:
    import oracle
    oracle.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    den = np.sum(er ** 2 + ei ** 2)
    out = {"cores": threads}
    for name, fn in (("nufft", lambda: oracle.nufft_all(ev, eta, threads)),
                     ("chebcode_fast", lambda: oracle.chebcode_preset(ev, eta, "chebcode_fast",
                                                                       threads))):
        r, i = fn()  # warm-up
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        out[name] = {"ms_per_transform": 1e3 * float(np.median(ts)),
                     "rel_l2_vs_fp64_exact": float(np.sqrt(np.sum((r - er) ** 2 + (i - ei) ** 2)
                                                           / den))}
    return out


def adjoint_leg(shk, sync, d_ev, eta) -> dict:
    """Reverse mode: (1) the differentiated El Karoui bulk fit at p = 10 000,
    c = 0.25, 256 atoms, 512 grid points (forward + backward per step, FP32-
    compensated and FP64, gradient agreement between them); (2) the all-points
    adjoint on the headline workload (gradient of a linear functional of S
    with respect to every eigenvalue and eta, p^2 ordered pairs, each
    unordered pair evaluated once), beside the two sweeps it replaces."""
    import torch
    from shrinkers_amd import autograd as ag
    from shrinkers_amd import stieltjes as st
    ev, grid, atoms, eta_f = ag.bulk_fit_problem()
    logits = 0.5 * np.random.default_rng(0).standard_normal(atoms.size)
    t_ev = torch.from_numpy(ev).cuda()
    t_grid = torch.from_numpy(grid).cuda()
    t_atoms = torch.from_numpy(atoms).cuda()
    out = {"workload": "bulk fit p=10000 c=0.25 MP (no outliers), K=256 atoms log-spaced "
                       "[0.2,5], softmax weights (seeded logits), M=512 grid points spanning "
                       "the bulk, eta=1/sqrt(p)"}
    grads = {}
    for prec in ("f32c", "f64"):
        def step():
            tev = t_ev.clone().requires_grad_(True)
            te = torch.tensor(eta_f, dtype=torch.float64, device="cuda", requires_grad=True)
            tl = torch.from_numpy(logits).cuda().requires_grad_(True)
            loss = ag.el_karoui_fit_loss(tev, te, tl, t_atoms, 0.25, t_grid, prec)
            loss.backward()
            return loss, tev.grad, te.grad, tl.grad
        for _ in range(3):
            step()
        sync()
        n = 20
        t0 = time.perf_counter()
        for _ in range(n):
            res = step()
        sync()
        ms = (time.perf_counter() - t0) / n * 1e3
        grads[prec] = [r.detach().cpu().numpy() for r in res]
        out[f"fit_{prec}"] = {"ms_forward_backward": ms, "loss": float(grads[prec][0])}

    # the same step as one library call (no autograd, no host round trips)
    w = torch.softmax(torch.from_numpy(logits).cuda(), dim=0)
    for prec in ("f32c", "f64"):
        def fused():
            ag.el_karoui_fit_step(t_ev, eta_f, w, t_atoms, 0.25, t_grid, prec)
        for _ in range(3):
            fused()
        sync()
        n = 50
        t0 = time.perf_counter()
        for _ in range(n):
            fused()
        sync()
        out[f"fit_{prec}"]["ms_fused_step"] = (time.perf_counter() - t0) / n * 1e3

    def rel(a, b):
        return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))
    out["f32c_vs_f64_grad_rel"] = {"eigenvalues": rel(grads["f32c"][1], grads["f64"][1]),
                                   "eta": rel(grads["f32c"][2], grads["f64"][2]),
                                   "logits": rel(grads["f32c"][3], grads["f64"][3])}

    # the all-points adjoint at the headline size, L = sum G_re Re S + G_im Im S:
    # dL/dlam (query and source roles together) and dL/deta in one call, each
    # unordered pair once; beside it the two sweeps (eigenvalue and query)
    p = d_ev.numel()
    rng = np.random.default_rng(5)
    g_re = torch.from_numpy(rng.standard_normal(p)).cuda()
    g_im = torch.from_numpy(rng.standard_normal(p)).cuda()
    gev = torch.empty_like(d_ev)
    gq = torch.empty_like(d_ev)
    ge = torch.empty((), dtype=torch.float64, device="cuda")
    lib = shk._lib.load()
    for prec_name, prec in (("f32c", 1), ("f64", 0)):
        def adj():
            shk._lib.check(lib.shk_exact_adjoint_all_points_dev(
                d_ev.data_ptr(), p, eta, g_re.data_ptr(), g_im.data_ptr(), prec,
                gev.data_ptr(), ge.data_ptr(), st._stream_handle(None)))

        def two_sweeps():
            shk._lib.check(lib.shk_exact_adjoint_dev(
                d_ev.data_ptr(), p, d_ev.data_ptr(), p, eta, g_re.data_ptr(), g_im.data_ptr(),
                prec, gev.data_ptr(), gq.data_ptr(), ge.data_ptr(), st._stream_handle(None)))
        steps = 20 if prec_name == "f32c" else 8
        t = time_steps(adj, steps, 2, sync) / steps
        k_ms = kernel_avg_ms(shk, adj, 5, sync)   # the symmetric adjoint kernel
        t2 = time_steps(two_sweeps, steps, 2, sync) / steps
        out[f"adjoint_all_points_p{p}_{prec_name}"] = {
            "ms_per_backward": t * 1e3,
            "ordered_pairs_per_s": float(p) * p / t,
            "kernel_avg_ms": k_ms,
            "two_sweeps_ms": t2 * 1e3,
        }
    return out


def grid_leg(shk, sync) -> dict:
    """At-points path of the El Karoui consumer: 200 grid points, p = 50 000."""
    import torch
    from shrinkers_amd import spectra
    ev = spectra.mp_inverse_cdf_spectrum(50_000)
    d = torch.from_numpy(ev).cuda()
    q = torch.linspace(float(ev[0]) - 0.4, float(ev[-1]) + 0.4, 200, dtype=torch.float64).cuda()
    out = {}
    for m in ("blocked", "chebcode_fast", "chebcode_balanced"):
        shk.at_points_dev(q, d, 0.1 / np.sqrt(ev.size), m)
        sync()
        t0 = time.perf_counter()
        for _ in range(10):
            shk.at_points_dev(q, d, 0.1 / np.sqrt(ev.size), m)
        sync()
        out[m] = {"ms_end_to_end": (time.perf_counter() - t0) * 100.0}
    return out


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--p", type=int, default=50_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU timing")
    ap.add_argument("--quick", action="store_true",
                    help="only the headline workload (no ChebCode / grid / reverse-mode legs)")
    ap.add_argument("--kernels", action="store_true",
                    help="only run each profiled workload three times (for rocprofv3 counter "
                         "passes): the headline f32c and FP64 transforms at p, then the one-shot "
                         "ChebCode FAST and DEFAULT transforms at 5e4 and 1e6; prints no JSON")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="first copy an MB-sized buffer on the device 5 times (a plain copy "
                         "kernel of known size, to calibrate HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for --gpus {args.gpus}); "
              "the path shards by query rows across GPUs but no multi-GPU run is part of it.",
              file=sys.stderr)
        return 2

    import torch
    import shrinkers_amd as shk
    from shrinkers_amd import spectra

    p = args.p
    ev = spectra.mp_inverse_cdf_spectrum(p, c=0.25, sigma2=1.0, seed=0)
    eta = 1.0 / np.sqrt(p)
    dev = torch.device("cuda:0")
    d_ev = torch.from_numpy(ev).to(dev)
    out32 = (torch.empty_like(d_ev), torch.empty_like(d_ev))
    out64 = (torch.empty_like(d_ev), torch.empty_like(d_ev))
    sync = torch.cuda.synchronize
    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        sync()
        log(f"copy probe: 5 x {args.copy_probe} MiB device copies")
        del src, dst

    def step32():
        shk.all_points_dev(d_ev, eta, "blocked", precision="f32c", scaled=True, out=out32)

    def step64():
        shk.all_points_dev(d_ev, eta, "blocked", precision="f64", scaled=True, out=out64)

    if args.kernels:
        for _ in range(3):
            step32()
            step64()
        sync()
        for pc in (50_000, 1_000_000):
            evc = torch.from_numpy(spectra.mp_inverse_cdf_spectrum(pc, c=0.25)).to(dev)
            for m in ("chebcode_fast", "chebcode"):
                for _ in range(3):
                    shk.all_points_dev(evc, 1.0 / np.sqrt(pc), m)
            sync()
            # the same transforms built in a caller's workspace (no host wait)
            for m in ("chebcode_fast", "chebcode"):
                ws = torch.empty(shk.chebcode_workspace_bytes(pc, m), dtype=torch.uint8,
                                 device=dev)
                for _ in range(3):
                    shk.chebcode_all_points_ws_dev(evc, 1.0 / np.sqrt(pc), m, ws)
                sync()
                del ws
        log("kernel workloads done")
        return 0

    log(f"p={p} eta={eta:.6g}: timing FP32-compensated ({args.steps} steps)")
    t32 = time_steps(step32, args.steps, args.warmup, sync)
    log("timing FP64 exact")
    steps64 = max(1, args.steps // 4)
    t64 = time_steps(step64, steps64, max(1, args.warmup // 4), sync)
    log("kernel event timing")
    k32 = kernel_avg_ms(shk, step32, 20, sync)
    k64 = kernel_avg_ms(shk, step64, 10, sync)

    # ChebCode FAST (build + evaluate) and DEFAULT at the same p, for reference
    def cheb(m):
        shk.all_points_dev(d_ev, eta, m, scaled=True, out=out64)
    tcf = time_steps(lambda: cheb("chebcode_fast"), 10, 2, sync) / 10
    kcf = kernel_avg_ms(shk, lambda: cheb("chebcode_fast"), 5, sync)

    r32 = (out32[0].cpu().numpy(), out32[1].cpu().numpy())
    r64 = (out64[0].cpu().numpy(), out64[1].cpu().numpy())
    step64()
    sync()
    r64 = (out64[0].cpu().numpy(), out64[1].cpu().numpy())

    def rel(a, b):
        num = np.sum((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2)
        return float(np.sqrt(num / np.sum(b[0] ** 2 + b[1] ** 2)))

    pairs = float(p) * float(p)
    # the range in which the all-points FP32-compensated entry takes the symmetric
    # kernel (each unordered pair once); the split-tile kernel runs outside it
    sym_lo, sym_hi = shk.symmetric_range("f32c")
    ms32 = t32 / args.steps * 1e3
    ms64 = t64 / steps64 * 1e3
    value = pairs * args.steps / t32
    achieved = FLOPS_PER_PAIR * pairs / (k32 * 1e-3) / 1e12

    cpu = None
    err_oracle_32 = err_oracle_64 = None
    if not args.no_cpu:
        log("CPU baseline (oracle restatement, all host cores)")
        ref, cpu = cpu_baseline(ev, eta, args.cpu_budget)
        ref = (ref[0] / p, ref[1] / p)
        err_oracle_32 = rel(r32, ref)
        err_oracle_64 = rel(r64, ref)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "exact all-points Stieltjes transform, MP inverse-CDF spectrum "
                        "c=0.25 sigma2=1 seed 0 + forward-BBP outliers of spikes 12/7/4, "
                        "eta=1/sqrt(p), FP32-compensated kernel",
            "p": p,
            "eta": eta,
            "pair_evals_per_step": pairs,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": TRAFFIC_P50000 if p == 50_000 else None,
            "traffic_unit": "bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "valu_busy": VALU_BUSY_P50000 if p == 50_000 else None,
            "valu_insts_per_ordered_pair": VALU_INSTS_PER_PAIR_P50000 if p == 50_000 else None,
            "kernel": "sym_f32c_kernel" if sym_lo <= p <= sym_hi else "exact_f32c_kernel",
            "kernel_avg_ms": k32,
            "flops_per_launch": FLOPS_PER_PAIR * pairs,
            "peak_box": box_peaks(),
        },
        "cpu_baseline": cpu,
        "fp64_exact": {
            "value": pairs / (ms64 * 1e-3),
            "unit": "pair-evals/s",
            "ms_per_step": ms64,
            "kernel_avg_ms": k64,
            "achieved_tflops": FLOPS_PER_PAIR * pairs / (k64 * 1e-3) / 1e12,
            "peak_tflops": PEAK_FP64_TFLOPS,
            "speedup_f32c_over_f64": ms64 / ms32,
        },
        "rel_l2_f32c_vs_fp64_gpu": rel(r32, r64),
        "rel_l2_f32c_vs_fp64_oracle": err_oracle_32,
        "rel_l2_fp64_gpu_vs_fp64_oracle": err_oracle_64,
        "chebcode_fast": {"ms_end_to_end": tcf * 1e3, "eval_kernel_ms": kcf,
                          **workspace_times(shk, d_ev, eta, "chebcode_fast", sync)},
    }
    if cpu:
        line["cpu_baseline"]["gpu_over_cpu"] = value / cpu["value"]
    if not args.quick:
        # 512 spectra: one call (segmented launch / stream pool) and the
        # prepared batch replayed from a hipGraph
        log("512 spectra in one call")
        line["batched"] = many_leg(shk, sync)
        log("ChebCode at p = 2e5 / 1e6")
        line["chebcode_large"] = chebcode_leg(shk, sync)
        log("grid (at-points) leg")
        line["grid_at_points_p50000_nq200"] = grid_leg(shk, sync)
        log("reverse mode (bulk fit, adjoint kernels)")
        line["reverse_mode"] = adjoint_leg(shk, sync, d_ev, eta)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
