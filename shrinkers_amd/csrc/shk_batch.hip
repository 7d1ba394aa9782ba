// Many spectra of uneven size as one prepared launch replayed from a hipGraph.
//
// There is no batched path in the reference (its batching is across eta on
// one spectrum, chebcode.rs:1312); each spectrum's result here is its own
// all-points transform (compute_all_stieltjes with the exact family,
// src/stieltjes/mod.rs:709-723) in FP32-compensated arithmetic.
//
// The launch is the segmented all-points kernel of shk_exact.hip (one
// one-wavefront workgroup per (spectrum, query tile), the tiles of all
// spectra in one grid, largest spectra first), or with r = 8 the segmented
// symmetric kernel of shk_sym.hip and its folds.  A batch keeps its spectrum
// table on the device and captures the launch (a single kernel node; a
// launch and a fold per group for r = 8) in a hipGraph on first use; later
// runs replay the graph.  Every workgroup stamps
// the device wall clock at its start and end, from which the per-spectrum
// start and end times are derived.
//
// Tuning: the configuration is timed on the actual batch and the fastest
// kept.  r = 1, 2 or 4 queries per lane are the split-tile kernel's tiles of
// 64, 128 or 256 queries: with r = 2 the tiles are the single-spectrum
// kernel's, so every result equals that kernel's bit for bit; r = 1 and 4 use
// other tiles (their own centres and pair forms) and agree within the
// FP32-compensated bound.  r = 8 is the symmetric kernel (tiles of 512, 8
// queries per lane), which evaluates each unordered pair once: its tile pairs
// of all spectra run in groups whose partial sums fit the symmetric scratch
// limit, each group's launch followed by an in-order fold, and every result
// equals the single-spectrum symmetric transform's bit for bit (the
// all-points entry's own result inside its symmetric range).
#include "shk_internal.hpp"
#include "shrinkers_stieltjes.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

using namespace shk;

namespace {

constexpr int kRs[4] = {1, 2, 4, 8};
constexpr int kSymR = 8;  // the symmetric kernel's queries per lane

int r_index(int r) { return r == 1 ? 0 : (r == 2 ? 1 : (r == 4 ? 2 : 3)); }

}  // namespace

struct shk_batch {
    const double *d_ev = nullptr;
    size_t n = 0;
    std::vector<int64_t> off;
    std::vector<double> eta;
    int scaled = 1;
    std::vector<size_t> order;   // spectra by descending size (the grid order)
    int r = 2;
    // per tile height: device spectrum table and its workgroup count (query
    // tiles; tile pairs for r = 8)
    SegSpectrum *d_table[3] = {nullptr, nullptr, nullptr};
    int n_table = 0;
    int64_t n_tiles[4] = {0, 0, 0, 0};
    std::vector<int64_t> tile0[4];  // host copy of each table's first workgroups
    long long *d_stamps = nullptr;  // 2 per tile of the largest tile count (r = 1)
    // r = 8: symmetric table, groups, partial-sum scratch and stamps
    SymSegSpectrum *d_sym = nullptr;
    std::vector<SymSegGroup> sym_groups;
    double *d_sym_part = nullptr;
    long long *d_sym_stamps = nullptr;
    // captured graph
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    double *g_re = nullptr, *g_im = nullptr;
    int g_r = 0;
    bool ran = false;
};

namespace {

shk_status bfail(shk_status s, const std::string &m) {
    set_error(m);
    return s;
}

#define SHK_BHIP(expr)                                                          \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
            return SHK_ERR_HIP;                                                 \
        }                                                                       \
    } while (0)

// Device table for tile height r (built once per r).
shk_status prepare_sym(shk_batch *b) {
    if (b->d_sym) return SHK_OK;
    std::vector<SymSegSpectrum> table(b->n);
    std::vector<SymSegGroup> groups;
    int64_t pairs = 0;
    const int st = sym_many_plan(b->off.data(), b->order.data(), b->n, b->eta.data(), b->scaled,
                                 sym_scratch_budget() / sizeof(double), table.data(), groups,
                                 &pairs);
    if (st) return (shk_status)st;
    size_t part = 0;
    for (const SymSegGroup &g : groups) part = std::max(part, g.part_doubles);
    hipError_t e = hipMalloc(&b->d_sym, sizeof(SymSegSpectrum) * b->n);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_sym, table.data(), sizeof(SymSegSpectrum) * b->n,
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&b->d_sym_part, sizeof(double) * std::max<size_t>(1, part));
    if (e == hipSuccess)
        e = hipMalloc(&b->d_sym_stamps, sizeof(long long) * 2 * (size_t)std::max<int64_t>(1, pairs));
    if (e != hipSuccess) {
        set_error(std::string("batch (r = 8): ") + hipGetErrorString(e));
        (void)hipFree(b->d_sym);
        (void)hipFree(b->d_sym_part);
        (void)hipFree(b->d_sym_stamps);
        b->d_sym = nullptr;
        b->d_sym_part = nullptr;
        b->d_sym_stamps = nullptr;
        return SHK_ERR_HIP;
    }
    b->sym_groups = groups;
    b->n_tiles[3] = pairs;
    b->tile0[3].resize(b->n);
    for (size_t k = 0; k < b->n; ++k) {
        const SymSegGroup *g = nullptr;
        for (const SymSegGroup &h : groups)
            if ((int)k >= h.first && (int)k < h.first + h.count) g = &h;
        b->tile0[3][k] = g->pair_base + table[k].pair0;
    }
    return SHK_OK;
}

shk_status prepare(shk_batch *b, int r) {
    if (r == kSymR) return prepare_sym(b);
    const int ri = r_index(r);
    if (b->d_table[ri]) return SHK_OK;
    std::vector<SegSpectrum> table(b->n);
    int n_table = 0, n_tiles = 0;
    const int st = exact_many_plan(b->off.data(), b->order.data(), b->n, b->eta.data(), b->scaled,
                                   table.data(), &n_table, &n_tiles, r);
    if (st) return (shk_status)st;
    SHK_BHIP(hipMalloc(&b->d_table[ri], sizeof(SegSpectrum) * std::max(1, n_table)));
    SHK_BHIP(hipMemcpy(b->d_table[ri], table.data(), sizeof(SegSpectrum) * n_table,
                       hipMemcpyHostToDevice));
    b->n_table = n_table;
    b->n_tiles[ri] = n_tiles;
    b->tile0[ri].resize(n_table);
    for (int e = 0; e < n_table; ++e) b->tile0[ri][e] = table[e].tile0;
    return SHK_OK;
}

shk_status capture(shk_batch *b, double *d_re, double *d_im) {
    if (b->exec) {
        SHK_BHIP(hipDeviceSynchronize());  // the old graph may still be executing
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    const int ri = r_index(b->r);
    hipGraph_t g = nullptr;
    SHK_BHIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    const int st = (b->r == kSymR)
                       ? sym_many(b->d_ev, b->d_sym, b->sym_groups, b->d_sym_part, d_re, d_im,
                                  b->d_sym_stamps, b->cap)
                       : exact_many(b->d_ev, b->d_table[ri], b->n_table, (int)b->n_tiles[ri], 1,
                                    d_re, d_im, b->d_stamps, b->cap, b->r);
    const hipError_t ce = hipStreamEndCapture(b->cap, &g);
    if (st || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (st) return (shk_status)st;
        return bfail(SHK_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(ce));
    }
    const hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return bfail(SHK_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    b->g_re = d_re;
    b->g_im = d_im;
    b->g_r = b->r;
    return SHK_OK;
}

shk_status launch(shk_batch *b, double *d_re, double *d_im, hipStream_t s) {
    if (!b->exec || b->g_re != d_re || b->g_im != d_im || b->g_r != b->r) {
        shk_status st = capture(b, d_re, d_im);
        if (st != SHK_OK) return st;
    }
    ProfileMark m = profile_begin(s);
    SHK_BHIP(hipGraphLaunch(b->exec, s));
    profile_end(m, s);
    b->ran = true;
    return SHK_OK;
}

}  // namespace

extern "C" {

shk_status shk_batch_create(const double *d_eigenvalues, const int64_t *offsets, size_t n_spectra,
                            const double *etas, int scaled, shk_batch **out) {
    if (!out) return bfail(SHK_ERR_BAD_ARG, "NULL handle pointer");
    *out = nullptr;
    if (n_spectra == 0) return bfail(SHK_ERR_EMPTY, "the batch must hold at least one spectrum");
    if (!d_eigenvalues || !offsets || !etas) return bfail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (offsets[0] != 0) return bfail(SHK_ERR_BAD_ARG, "offsets[0] must be 0");
    for (size_t s = 0; s < n_spectra; ++s) {
        const int64_t p = offsets[s + 1] - offsets[s];
        if (p <= 0) return bfail(SHK_ERR_EMPTY, "every spectrum must be non-empty");
        if (p > INT32_MAX) return bfail(SHK_ERR_BAD_ARG, "spectrum larger than 2^31-1");
        if (!(std::isfinite(etas[s]) && etas[s] > 0.0))
            return bfail(SHK_ERR_BAD_ETA, "eta must be a positive finite number, got " +
                                              std::to_string(etas[s]));
        // the prepared batch runs the FP32-compensated kernel only
        if (!f32c_eta_ok(etas[s]))
            return bfail(SHK_ERR_BAD_ETA,
                         "eta outside the FP32-compensated range [1e-18, 1e18] (use "
                         "shk_all_points_many_dev, which runs FP64 there), got " +
                             std::to_string(etas[s]));
    }
    if (const int dv = device_ready()) return (shk_status)dv;
    shk_batch *b = new shk_batch();
    b->d_ev = d_eigenvalues;
    b->n = n_spectra;
    b->off.assign(offsets, offsets + n_spectra + 1);
    b->eta.assign(etas, etas + n_spectra);
    b->scaled = scaled ? 1 : 0;
    b->order.resize(n_spectra);
    std::iota(b->order.begin(), b->order.end(), (size_t)0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](size_t a, size_t c) {
        return (b->off[a + 1] - b->off[a]) > (b->off[c + 1] - b->off[c]);
    });
    // stamps for the largest tile count (r = 1)
    int64_t max_tiles = 0;
    for (size_t s = 0; s < n_spectra; ++s) max_tiles += (b->off[s + 1] - b->off[s] + 63) / 64;
    hipError_t e = hipMalloc(&b->d_stamps, sizeof(long long) * 2 * (size_t)std::max<int64_t>(1, max_tiles));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error(std::string("shk_batch_create: ") + hipGetErrorString(e));
        shk_batch_destroy(b);
        return SHK_ERR_HIP;
    }
    const shk_status st = prepare(b, 2);
    if (st != SHK_OK) {
        shk_batch_destroy(b);
        return st;
    }
    *out = b;
    return SHK_OK;
}

shk_status shk_batch_run(shk_batch *b, double *d_out_re, double *d_out_im, void *stream) {
    if (!b || !d_out_re || !d_out_im) return bfail(SHK_ERR_BAD_ARG, "NULL pointer");
    return launch(b, d_out_re, d_out_im, (hipStream_t)stream);
}

shk_status shk_batch_tune(shk_batch *b, double *d_out_re, double *d_out_im, void *stream,
                          int *best_r, double *best_ms) {
    if (!b || !d_out_re || !d_out_im) return bfail(SHK_ERR_BAD_ARG, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    double best = 1e300;
    int bR = 2;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SHK_BHIP(hipEventCreate(&e0));
    SHK_BHIP(hipEventCreate(&e1));
    shk_status st = SHK_OK;
    for (int R : kRs) {
        if ((st = prepare(b, R)) != SHK_OK) {
            // the symmetric kernel is refused for spectra whose partial sums
            // alone exceed the symmetric scratch limit: the split-tile
            // heights remain
            if (R != kSymR) break;
            st = SHK_OK;
            set_error("");
            continue;
        }
        b->r = R;
        if ((st = launch(b, d_out_re, d_out_im, s)) != SHK_OK) break;  // capture + warm-up
        float ms = 0.f, tot = 1e30f;
        for (int rep = 0; rep < 3 && st == SHK_OK; ++rep) {
            if (hipEventRecord(e0, s) != hipSuccess) st = SHK_ERR_HIP;
            if (st == SHK_OK) st = launch(b, d_out_re, d_out_im, s);
            if (st == SHK_OK && hipEventRecord(e1, s) != hipSuccess) st = SHK_ERR_HIP;
            if (st == SHK_OK && hipEventSynchronize(e1) != hipSuccess) st = SHK_ERR_HIP;
            if (st == SHK_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = SHK_ERR_HIP;
            tot = std::min(tot, ms);
        }
        if (st != SHK_OK) break;
        if (tot < best) {
            best = tot;
            bR = R;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != SHK_OK) {
        if (std::string(last_error()).empty()) set_error("batch tuning failed");
        return st;
    }
    b->r = bR;
    if (best_r) *best_r = bR;
    if (best_ms) *best_ms = best;
    return SHK_OK;
}

shk_status shk_batch_set_config(shk_batch *b, int r) {
    if (!b) return bfail(SHK_ERR_BAD_ARG, "NULL handle");
    if (r != 1 && r != 2 && r != 4 && r != kSymR)
        return bfail(SHK_ERR_BAD_ARG, "r must be 1, 2, 4 or 8");
    const shk_status st = prepare(b, r);
    if (st != SHK_OK) return st;
    b->r = r;
    return SHK_OK;
}

shk_status shk_batch_spectrum_times(const shk_batch *b, double *start_us, double *end_us) {
    if (!b || !start_us || !end_us) return bfail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (!b->ran || !b->exec) return bfail(SHK_ERR_BAD_ARG, "the batch has not run yet");
    const int ri = r_index(b->g_r);
    const int64_t nt = b->n_tiles[ri];
    std::vector<long long> t(2 * (size_t)std::max<int64_t>(1, nt));
    SHK_BHIP(hipMemcpy(t.data(), b->g_r == kSymR ? b->d_sym_stamps : b->d_stamps,
                       sizeof(long long) * 2 * (size_t)nt, hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    SHK_BHIP(hipGetDevice(&dev));
    SHK_BHIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) return bfail(SHK_ERR_HIP, "device wall-clock rate unavailable");
    const double us_per_tick = 1e3 / (double)khz;
    long long origin = t[0];
    for (int64_t k = 0; k < nt; ++k) origin = std::min(origin, t[2 * (size_t)k]);
    const std::vector<int64_t> &t0 = b->tile0[ri];
    for (int e = 0; e < b->n_table; ++e) {
        const int64_t a = t0[e], z = (e + 1 < b->n_table) ? t0[e + 1] : nt;
        long long lo = t[2 * (size_t)a], hi = t[2 * (size_t)a + 1];
        for (int64_t k = a; k < z; ++k) {
            lo = std::min(lo, t[2 * (size_t)k]);
            hi = std::max(hi, t[2 * (size_t)k + 1]);
        }
        const size_t s = b->order[e];  // every spectrum is non-empty: entry e is order[e]
        start_us[s] = (double)(lo - origin) * us_per_tick;
        end_us[s] = (double)(hi - origin) * us_per_tick;
    }
    return SHK_OK;
}

shk_status shk_batch_destroy(shk_batch *b) {
    if (!b) return SHK_OK;
    if (b->exec || b->ran) (void)hipDeviceSynchronize();  // a replay may still read the table
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (SegSpectrum *t : b->d_table)
        if (t) (void)hipFree(t);
    if (b->d_stamps) (void)hipFree(b->d_stamps);
    if (b->d_sym) (void)hipFree(b->d_sym);
    if (b->d_sym_part) (void)hipFree(b->d_sym_part);
    if (b->d_sym_stamps) (void)hipFree(b->d_sym_stamps);
    delete b;
    return SHK_OK;
}

}  // extern "C"
