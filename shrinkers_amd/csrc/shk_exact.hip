// Exact O(p^2) Stieltjes kernels for gfx950 (CDNA4).
//
// Reference arithmetic: src/stieltjes/term.rs:26-45 (term),
// src/stieltjes/cacheblock.rs:1286-1424 (full-square tiled body) and
// :1454-1513 (disjoint output spans).  The GPU decomposition keeps the
// reference's "query tile owns its outputs, sweeps the sources" shape:
//
//   grid.x = query tiles of kQT = kWG*kR queries (kR per thread, strided so
//            loads and stores coalesce), grid.y = source splits (about
//            kTargetWG workgroups in flight).
//   Each workgroup stages its source split through LDS in pieces of kPiece
//   values; every wavefront then reads the same LDS word (a broadcast, no
//   bank conflicts) and updates kR independent query accumulators, so one
//   LDS read feeds kR pair evaluations and every reciprocal chain is
//   independent of its neighbours.
//   Splits write FP64 partial sums to scratch ([acc][split][query]); a second
//   kernel adds them in split order (deterministic) and applies eta and the
//   1/p scale once per output, as the reference's eta-hoisted form does
//   (term.rs:33-45, cacheblock.rs:218-221).
//
// FP32-compensated variant (exact_f32c): the pair arithmetic is FP32 on
// eigenvalues re-expressed relative to a per-tile centre c (computed in FP64
// from the tile's query range).  Per workgroup one of three forms is chosen
// from the tile half-width M:
//   M <= 2*eta  : d = qh - sh                  (rounding of x_i - c costs at
//                                               most eps/2*M <= eps*eta)
//   M <= 32*eta : d = (qh - sh) + ql           (query residual restored)
//   otherwise   : d = (qh - sh) + (ql - sl)    (both residuals restored)
// where qh/ql and sh/sl are the f32 hi/lo parts of x_i - c and lam_j - c.
// For the near pairs that dominate Im S, qh - sh is exact (Sterbenz), so the
// pair difference carries at most one f32 rounding.  The reciprocal is
// v_rcp_f32 (1 ulp).  Partial sums run over kF32Run sources in FP32 and are
// folded into FP64 accumulators; the cross-split reduction is FP64.
#include "shk_internal.hpp"
#include "shk_pair.hpp"

#include <algorithm>
#include <vector>

namespace shk {
namespace {

// Tile shape; the defaults are the measured choice on MI355X (one wavefront
// per workgroup, 2 queries per lane: 128-query tiles, narrow enough that
// the one-subtraction pair difference applies to most tiles; about 16 384
// workgroups: as fast as 32 768 at half the partial-sum traffic, while 8 192
// and fewer leave the last round of workgroups unbalanced) and can be
// overridden at compile time for tuning sweeps.
#ifndef SHK_EXACT_WG
#define SHK_EXACT_WG 64
#endif
#ifndef SHK_EXACT_R
#define SHK_EXACT_R 2
#endif
#ifndef SHK_EXACT_PIECE
#define SHK_EXACT_PIECE 512
#endif
#ifndef SHK_EXACT_TARGET_WG
#define SHK_EXACT_TARGET_WG 16384
#endif
#ifndef SHK_EXACT_MIN_SRC
#define SHK_EXACT_MIN_SRC 64
#endif
constexpr int kWG = SHK_EXACT_WG;          // threads per workgroup
constexpr int kR = SHK_EXACT_R;            // queries per thread
constexpr int kQT = kWG * kR;              // queries per tile
constexpr int kPiece = SHK_EXACT_PIECE;    // sources staged in LDS per pass
constexpr int kTargetWG = SHK_EXACT_TARGET_WG;  // workgroups to aim for
// Sources per split at least.  Measured (all-points, ms, f64 / f32c): 64 ->
// p = 1 000: 0.014 / 0.015, 5 000: 0.029 / 0.023, 8 000: 0.045 / 0.035;
// 128 -> 0.017 / 0.013, 0.025 / 0.018, 0.040 / 0.029; 256 -> 0.027 / 0.019,
// 0.031 / 0.021, 0.040 / 0.026; grids of 8 000 and 20 000 points against
// 5e4 sources and the 512-spectrum batch unchanged.  64 keeps the FP64
// transform at p ~ 1 000 (the spiked-Wishart workload) fastest.
constexpr int kMinSrc = SHK_EXACT_MIN_SRC;

template <int KIND>
struct NAcc {
    static constexpr int value = (KIND == EXACT_DERIV) ? 4 : 2;
};

template <int KIND>
__global__ __launch_bounds__(kWG) void exact_f64_kernel(
    const double *__restrict__ q, int nq, const double *__restrict__ src, int p, double eta,
    double cut, int src_per_split, double *__restrict__ part) {
    constexpr int NA = NAcc<KIND>::value;
    __shared__ double s_src[kPiece];
    const int tid = threadIdx.x;
    const int split = blockIdx.y;
    const int n_split = gridDim.y;
    const int j_begin = split * src_per_split;
    const int j_end = min(j_begin + src_per_split, p);

    double x[kR], e2[kR], acc[NA][kR];
#pragma unroll
    for (int k = 0; k < kR; ++k) {
        const int i = blockIdx.x * kQT + k * kWG + tid;
        x[k] = (i < nq) ? q[i] : 0.0;
        const double ek = (KIND == EXACT_RAY) ? eta * x[k] : eta;
        e2[k] = ek * ek;
#pragma unroll
        for (int a = 0; a < NA; ++a) acc[a][k] = 0.0;
    }

    for (int base = j_begin; base < j_end; base += kPiece) {
        const int cnt = min(kPiece, j_end - base);
        __syncthreads();
        for (int j = tid; j < cnt; j += kWG) s_src[j] = src[base + j];
        __syncthreads();
        // Two sources per step (one combined skip-branch for their paired
        // reciprocals); each pair's reciprocal depends on its own
        // denominators only, so the grouping does not change any result.
        auto accumulate = [&](const double (&dd)[kR], const double (&rr)[kR]) {
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                const double d = dd[k];
                const double r = rr[k];
                if constexpr (KIND == EXACT_CUTOFF) {
                    const double rm = (__builtin_fabs(d) <= cut) ? r : 0.0;
                    acc[0][k] = __builtin_fma(d, rm, acc[0][k]);
                    acc[1][k] += rm;
                } else if constexpr (KIND == EXACT_HYBRID) {
                    acc[0][k] = __builtin_fma(d, r, acc[0][k]);
                    acc[1][k] += (__builtin_fabs(d) <= cut) ? r : 0.0;
                } else if constexpr (KIND == EXACT_DERIV) {
                    // S' = -sum (d^2 - eta^2) inv^2  +  i * (-sum 2 d eta inv^2)
                    acc[0][k] = __builtin_fma(d, r, acc[0][k]);
                    acc[1][k] += r;
                    const double r2 = r * r;
                    acc[2][k] = __builtin_fma(-__builtin_fma(d, d, -e2[k]), r2, acc[2][k]);
                    acc[3][k] = __builtin_fma(-2.0 * d * eta, r2, acc[3][k]);
                } else {
                    acc[0][k] = __builtin_fma(d, r, acc[0][k]);
                    acc[1][k] += r;
                }
            }
        };
        int j = 0;
        for (; j + 2 <= cnt; j += 2) {
            const double s0 = s_src[j], s1 = s_src[j + 1];
            double dd[2 * kR], tt[2 * kR], rr[2 * kR];
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                dd[k] = x[k] - s0;
                dd[kR + k] = x[k] - s1;
                tt[k] = __builtin_fma(dd[k], dd[k], e2[k]);
                tt[kR + k] = __builtin_fma(dd[kR + k], dd[kR + k], e2[k]);
            }
            rcp_many_f64(tt, rr);
            double d0[kR], r0[kR], d1[kR], r1[kR];
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                d0[k] = dd[k];
                r0[k] = rr[k];
                d1[k] = dd[kR + k];
                r1[k] = rr[kR + k];
            }
            accumulate(d0, r0);
            accumulate(d1, r1);
        }
        if (j < cnt) {
            const double s0 = s_src[j];
            double dd[kR], tt[kR], rr[kR];
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                dd[k] = x[k] - s0;
                tt[k] = __builtin_fma(dd[k], dd[k], e2[k]);
            }
            rcp_many_f64(tt, rr);
            accumulate(dd, rr);
        }
    }

#pragma unroll
    for (int k = 0; k < kR; ++k) {
        const int i = blockIdx.x * kQT + k * kWG + tid;
        if (i < nq) {
#pragma unroll
            for (int a = 0; a < NA; ++a)
                part[((size_t)a * n_split + split) * nq + i] = acc[a][k];
        }
    }
}

// ---- FP32-compensated pair kernel ------------------------------------------

// Loads the R queries each lane owns in query tile `tile` (kWG*R queries)
// of q[0, nq), finds the tile centre c (midpoint of its query range, FP64)
// and the pair form, and splits x - c into FP32 hi/lo parts.  Shared by the
// single-spectrum and the segmented kernels so that both see the same tile
// arithmetic.
template <int R>
__device__ __forceinline__ int f32c_tile_setup(const double *__restrict__ q, int nq, int tile,
                                               double eta, double (*s_red)[kWG / 64],
                                               double *c_out, float (&qh)[R],
                                               float (&ql)[R]) {
    const int tid = threadIdx.x;
    double x[R];
    double mn = __builtin_inf(), mx = -__builtin_inf();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tile * (kWG * R) + k * kWG + tid;
        x[k] = (i < nq) ? q[i] : 0.0;
        if (i < nq) {
            mn = fmin(mn, x[k]);
            mx = fmax(mx, x[k]);
        }
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    if ((tid & 63) == 0) {
        s_red[0][tid >> 6] = mn;
        s_red[1][tid >> 6] = mx;
    }
    __syncthreads();
    mn = s_red[0][0];
    mx = s_red[1][0];
#pragma unroll
    for (int w = 1; w < kWG / 64; ++w) {
        mn = fmin(mn, s_red[0][w]);
        mx = fmax(mx, s_red[1][w]);
    }
    const double c = 0.5 * (mn + mx);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double t = x[k] - c;
        qh[k] = (float)t;
        ql[k] = (float)(t - (double)qh[k]);
    }
    *c_out = c;
    return f32c_form(0.5 * (mx - mn), eta);
}

// Stages cnt sources as FP32 hi/lo parts of (src - c) in LDS.
__device__ __forceinline__ void f32c_stage(const double *__restrict__ src, int cnt, double c,
                                           float *__restrict__ s_hi, float *__restrict__ s_lo) {
    for (int j = threadIdx.x; j < cnt; j += kWG) {
        const double t = src[j] - c;
        const float h = (float)t;
        s_hi[j] = h;
        s_lo[j] = (float)(t - (double)h);
    }
}

template <int R>
__device__ __forceinline__ void f32c_sweep_any(int form, const float *s_hi, const float *s_lo,
                                               int cnt, const float (&qh)[R],
                                               const float (&ql)[R], float eta2f,
                                               double (&re)[R], double (&iv)[R]) {
    if (form == 0)
        f32c_sweep<R, 0>(s_hi, s_lo, cnt, qh, ql, eta2f, re, iv);
    else if (form == 1)
        f32c_sweep<R, 1>(s_hi, s_lo, cnt, qh, ql, eta2f, re, iv);
    else
        f32c_sweep<R, 2>(s_hi, s_lo, cnt, qh, ql, eta2f, re, iv);
}

__global__ __launch_bounds__(kWG) void exact_f32c_kernel(const double *__restrict__ q, int nq,
                                                         const double *__restrict__ src, int p,
                                                         double eta, int src_per_split,
                                                         double *__restrict__ part) {
    __shared__ float s_hi[kPiece];
    __shared__ float s_lo[kPiece];
    __shared__ double s_red[2][kWG / 64];
    const int tid = threadIdx.x;
    const int split = blockIdx.y;
    const int n_split = gridDim.y;
    const int j_begin = split * src_per_split;
    const int j_end = min(j_begin + src_per_split, p);

    double c;
    float qh[kR], ql[kR];
    const int form = f32c_tile_setup<kR>(q, nq, blockIdx.x, eta, s_red, &c, qh, ql);
    double re[kR], iv[kR];
#pragma unroll
    for (int k = 0; k < kR; ++k) re[k] = iv[k] = 0.0;
    const float eta2f = (float)(eta * eta);

    for (int base = j_begin; base < j_end; base += kPiece) {
        const int cnt = min(kPiece, j_end - base);
        __syncthreads();
        f32c_stage(src + base, cnt, c, s_hi, s_lo);
        __syncthreads();
        f32c_sweep_any<kR>(form, s_hi, s_lo, cnt, qh, ql, eta2f, re, iv);
    }

#pragma unroll
    for (int k = 0; k < kR; ++k) {
        const int i = blockIdx.x * kQT + k * kWG + tid;
        if (i < nq) {
            part[((size_t)0 * n_split + split) * nq + i] = re[k];
            part[((size_t)1 * n_split + split) * nq + i] = iv[k];
        }
    }
}

// ---- deterministic cross-split reduction ------------------------------------

// One lane per query, one wavefront per workgroup (nq/64 workgroups spread
// over every CU).  Loads run kRedAhead splits ahead of the additions, which
// stay in split order.
constexpr int kRedWG = 64;
constexpr int kRedAhead = 8;

template <int KIND, typename OutT>
__global__ __launch_bounds__(kRedWG) void reduce_kernel(const double *__restrict__ part,
                                                        int n_split, int nq, double eta,
                                                        double scale,
                                                        const double *__restrict__ q,
                                                        OutT *__restrict__ o0,
                                                        OutT *__restrict__ o1,
                                                        OutT *__restrict__ o2,
                                                        OutT *__restrict__ o3) {
    constexpr int NA = NAcc<KIND>::value;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    double a[NA];
#pragma unroll
    for (int t = 0; t < NA; ++t) a[t] = 0.0;
    int s = 0;
    for (; s + kRedAhead <= n_split; s += kRedAhead) {
        double v[kRedAhead][NA];
#pragma unroll
        for (int u = 0; u < kRedAhead; ++u)
#pragma unroll
            for (int t = 0; t < NA; ++t) v[u][t] = part[((size_t)t * n_split + s + u) * nq + i];
#pragma unroll
        for (int u = 0; u < kRedAhead; ++u)
#pragma unroll
            for (int t = 0; t < NA; ++t) a[t] += v[u][t];
    }
    for (; s < n_split; ++s) {
#pragma unroll
        for (int t = 0; t < NA; ++t) a[t] += part[((size_t)t * n_split + s) * nq + i];
    }
    const double eta_i = (KIND == EXACT_RAY) ? eta * q[i] : eta;
    o0[i] = (OutT)(a[0] * scale);
    o1[i] = (OutT)((eta_i * a[1]) * scale);
    if constexpr (NA == 4) {
        o2[i] = (OutT)(a[2] * scale);
        o3[i] = (OutT)(a[3] * scale);
    }
}

__global__ void widen_kernel(const float *__restrict__ in, double *__restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

// ---- segmented launch: the all-points sums of many spectra at once ----------
//
// One workgroup per (spectrum, query tile); the tiles of all spectra form one
// grid, enumerated spectrum by spectrum in the order of the table (largest
// spectra first, so the long tiles start first and the short ones fill in).
// A workgroup sweeps its spectrum's whole source range with the per-tile
// arithmetic of the single-spectrum kernels, and keeps a chunk accumulator
// that it adds to the running total every `per` sources -- `per` being the
// split length that the single-spectrum plan picks for that spectrum.  Runs
// of kF32Run sources start at multiples of 32 from each split start in both
// (per and kPiece are multiples of 32), so the FP64 additions happen in the
// same order as split partials followed by the in-order cross-split
// reduction: every output equals the single-spectrum call bit for bit, with
// no partial-sum scratch and no second kernel.

__device__ __forceinline__ int seg_find(const SegSpectrum *__restrict__ specs, int n, int tile) {
    int lo = 0, hi = n - 1;  // the last entry whose first tile is <= tile
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (specs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Device wall clock (constant rate) at a workgroup's start (which = 0) or
// end (which = 1), written by lane 0 to stamps[2 * blockIdx.x + which].
__device__ __forceinline__ void seg_stamp(long long *stamps, int which) {
    if (which) __syncthreads();  // the end stamp follows every lane's stores
    if (threadIdx.x == 0) stamps[2 * (size_t)blockIdx.x + which] = (long long)wall_clock64();
}

// Sources [base, base + cnt) are in LDS; sweeps them in sub-ranges cut at the
// chunk ends, folding the chunk accumulator into the total at each end.
template <int R, typename Sweep>
__device__ __forceinline__ void seg_sweep_piece(int base, int cnt, int per, int p,
                                                int &chunk_end, double (&re)[R],
                                                double (&iv)[R], double (&tr)[R],
                                                double (&ti)[R], Sweep sweep) {
    int j = 0;
    while (j < cnt) {
        const int stop = min(cnt, chunk_end - base);
        sweep(j, stop - j);
        j = stop;
        if (base + j == chunk_end) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                tr[k] += re[k];
                ti[k] += iv[k];
                re[k] = iv[k] = 0.0;
            }
            chunk_end = (p - chunk_end > per) ? chunk_end + per : p;
        }
    }
}

// R queries per lane (tiles of kWG*R); R = kR gives the single-spectrum
// tiles and hence the same bits as exact_f32c.
template <int R, bool STAMP>
__global__ __launch_bounds__(kWG) void many_f32c_kernel(const double *__restrict__ ev,
                                                        const SegSpectrum *__restrict__ specs,
                                                        int n_specs, double *__restrict__ out_re,
                                                        double *__restrict__ out_im,
                                                        long long *__restrict__ stamps) {
    __shared__ float s_hi[kPiece];
    __shared__ float s_lo[kPiece];
    __shared__ double s_red[2][kWG / 64];
    if (STAMP) seg_stamp(stamps, 0);
    const SegSpectrum sp = specs[seg_find(specs, n_specs, (int)blockIdx.x)];
    const double *__restrict__ src = ev + sp.off;
    const int p = sp.p;
    const int tile = (int)blockIdx.x - sp.tile0;

    double c;
    float qh[R], ql[R];
    const int form = f32c_tile_setup<R>(src, p, tile, sp.eta, s_red, &c, qh, ql);
    double re[R], iv[R], tr[R], ti[R];
#pragma unroll
    for (int k = 0; k < R; ++k) re[k] = iv[k] = tr[k] = ti[k] = 0.0;
    const float eta2f = (float)(sp.eta * sp.eta);

    int chunk_end = min(sp.per, p);
    for (int base = 0; base < p; base += kPiece) {
        const int cnt = min(kPiece, p - base);
        __syncthreads();
        f32c_stage(src + base, cnt, c, s_hi, s_lo);
        __syncthreads();
        seg_sweep_piece<R>(base, cnt, sp.per, p, chunk_end, re, iv, tr, ti, [&](int j0, int n) {
            f32c_sweep_any<R>(form, s_hi + j0, s_lo + j0, n, qh, ql, eta2f, re, iv);
        });
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tile * (kWG * R) + k * kWG + (int)threadIdx.x;
        if (i < p) {
            out_re[sp.off + i] = tr[k] * sp.scale;
            out_im[sp.off + i] = (sp.eta * ti[k]) * sp.scale;
        }
    }
    if (STAMP) seg_stamp(stamps, 1);
}

template <bool STAMP>
__global__ __launch_bounds__(kWG) void many_f64_kernel(const double *__restrict__ ev,
                                                       const SegSpectrum *__restrict__ specs,
                                                       int n_specs, double *__restrict__ out_re,
                                                       double *__restrict__ out_im,
                                                       long long *__restrict__ stamps) {
    __shared__ double s_src[kPiece];
    if (STAMP) seg_stamp(stamps, 0);
    const SegSpectrum sp = specs[seg_find(specs, n_specs, (int)blockIdx.x)];
    const double *__restrict__ src = ev + sp.off;
    const int p = sp.p;
    const int tile = (int)blockIdx.x - sp.tile0;
    const int tid = threadIdx.x;

    double x[kR], e2[kR], re[kR], iv[kR], tr[kR], ti[kR];
#pragma unroll
    for (int k = 0; k < kR; ++k) {
        const int i = tile * kQT + k * kWG + tid;
        x[k] = (i < p) ? src[i] : 0.0;
        e2[k] = sp.eta * sp.eta;
        re[k] = iv[k] = tr[k] = ti[k] = 0.0;
    }
    int chunk_end = min(sp.per, p);
    for (int base = 0; base < p; base += kPiece) {
        const int cnt = min(kPiece, p - base);
        __syncthreads();
        for (int j = tid; j < cnt; j += kWG) s_src[j] = src[base + j];
        __syncthreads();
        seg_sweep_piece<kR>(base, cnt, sp.per, p, chunk_end, re, iv, tr, ti, [&](int j0, int n) {
            // two sources per step, as in exact_f64_kernel (same bits)
            int j = j0;
            for (; j + 2 <= j0 + n; j += 2) {
                const double s0 = s_src[j], s1 = s_src[j + 1];
                double dd[2 * kR], tt[2 * kR], rr[2 * kR];
#pragma unroll
                for (int k = 0; k < kR; ++k) {
                    dd[k] = x[k] - s0;
                    dd[kR + k] = x[k] - s1;
                    tt[k] = __builtin_fma(dd[k], dd[k], e2[k]);
                    tt[kR + k] = __builtin_fma(dd[kR + k], dd[kR + k], e2[k]);
                }
                rcp_many_f64(tt, rr);
#pragma unroll
                for (int k = 0; k < kR; ++k) {
                    re[k] = __builtin_fma(dd[k], rr[k], re[k]);
                    iv[k] += rr[k];
                }
#pragma unroll
                for (int k = 0; k < kR; ++k) {
                    re[k] = __builtin_fma(dd[kR + k], rr[kR + k], re[k]);
                    iv[k] += rr[kR + k];
                }
            }
            if (j < j0 + n) {
                const double s0 = s_src[j];
                double dd[kR], tt[kR], rr[kR];
#pragma unroll
                for (int k = 0; k < kR; ++k) {
                    dd[k] = x[k] - s0;
                    tt[k] = __builtin_fma(dd[k], dd[k], e2[k]);
                }
                rcp_many_f64(tt, rr);
#pragma unroll
                for (int k = 0; k < kR; ++k) {
                    re[k] = __builtin_fma(dd[k], rr[k], re[k]);
                    iv[k] += rr[k];
                }
            }
        });
    }
#pragma unroll
    for (int k = 0; k < kR; ++k) {
        const int i = tile * kQT + k * kWG + tid;
        if (i < p) {
            out_re[sp.off + i] = tr[k] * sp.scale;
            out_im[sp.off + i] = (sp.eta * ti[k]) * sp.scale;
        }
    }
    if (STAMP) seg_stamp(stamps, 1);
}

// Source splits so that about kTargetWG workgroups are in flight.
void plan_splits(size_t nq, size_t p, int *n_qtiles, int *n_split, int *src_per_split) {
    const int qt = (int)((nq + kQT - 1) / kQT);
    int ns = std::max(1, (kTargetWG + qt - 1) / qt);
    const int max_split = (int)std::max<size_t>(1, (p + kMinSrc - 1) / kMinSrc);
    ns = std::min(ns, max_split);
    int per = (int)((p + ns - 1) / ns);
    per = (per + 31) / 32 * 32;
    ns = (int)((p + per - 1) / per);
    *n_qtiles = qt;
    *n_split = std::max(ns, 1);
    *src_per_split = per;
}

template <int KIND>
int launch_f64(const double *q, size_t nq, const double *src, size_t p, double eta, double cut,
               int n_qtiles, int n_split, int per, double *part, hipStream_t stream) {
    dim3 grid(n_qtiles, n_split);
    ProfileMark m = profile_begin(stream);
    hipLaunchKernelGGL(exact_f64_kernel<KIND>, grid, dim3(kWG), 0, stream, q, (int)nq, src,
                       (int)p, eta, cut, per, part);
    profile_end(m, stream);
    SHK_HIP(hipGetLastError());
    return 0;
}

template <int KIND, typename OutT>
int launch_reduce(const double *part, int n_split, size_t nq, double eta, double scale,
                  const double *q, OutT *o0, OutT *o1, OutT *o2, OutT *o3, hipStream_t stream) {
    const int blocks = (int)((nq + kRedWG - 1) / kRedWG);
    hipLaunchKernelGGL((reduce_kernel<KIND, OutT>), dim3(blocks), dim3(kRedWG), 0, stream, part,
                       n_split, (int)nq, eta, scale, q, o0, o1, o2, o3);
    SHK_HIP(hipGetLastError());
    return 0;
}

// ---- few query points against many sources -----------------------------------
//
// A grid of query points (the El Karoui consumer's 200, the bulk fit's 512)
// against a whole spectrum: with few query tiles the split plan above gives
// each query hundreds of splits, and the in-order reduction of those
// partials is a chain of dependent loads (about 60 us for 200 points against
// 5e4 sources).  Here one workgroup takes one query point: thread t sums the
// sources t, t + kGridWG, ... in order (coalesced loads), then the thread
// sums are added in a fixed tree through LDS, and eta and the scale are
// applied once (as reduce_kernel does).  Deterministic; only the summation
// order differs from the split plan (both within the family's tolerances).
// FP32-compensated: the pair difference is the FP64 difference rounded to
// f32 (one rounding, exact for the near pairs by Sterbenz), the reciprocal
// v_rcp_f32, and each thread's FP32 sums over kGridUnroll sources are folded
// into FP64.  FP64: reciprocals four per v_rcp_f64 (rcp_many_f64) and FP64
// sums.
constexpr int kGridWG = kGridThreads;
constexpr int kGridUnroll = 8;  // sources per thread and step: their loads go out together

template <bool F32C>
__global__ __launch_bounds__(kGridWG) void grid_query_kernel(const double *__restrict__ q,
                                                             const double *__restrict__ src,
                                                             int p, double eta, double scale,
                                                             double *__restrict__ out_re,
                                                             double *__restrict__ out_im) {
    constexpr int U = kGridUnroll;
    __shared__ double s_re[kGridWG], s_im[kGridWG];
    const int t = threadIdx.x;
    const double x = q[blockIdx.x];
    const double e2 = eta * eta;
    const float e2f = (float)e2;
    double re = 0.0, im = 0.0;
    int j = t;
    // steps of U sources (j, j + kGridWG, ...): FP32-compensated sums each
    // step's terms in FP32 and folds them into FP64
    for (; j + (U - 1) * kGridWG < p; j += U * kGridWG) {
        double d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] = x - src[j + u * kGridWG];
        if constexpr (F32C) {
            float ar = 0.f, ai = 0.f;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float df = (float)d[u];
                const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(df, df, e2f));
                ar = __builtin_fmaf(df, r, ar);
                ai += r;
            }
            re += (double)ar;
            im += (double)ai;
        } else {
            double tt[U], rr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) tt[u] = __builtin_fma(d[u], d[u], e2);
            rcp_many_f64<U>(tt, rr);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                re = __builtin_fma(d[u], rr[u], re);
                im += rr[u];
            }
        }
    }
    for (; j < p; j += kGridWG) {  // the last few sources, one at a time
        const double d = x - src[j];
        if constexpr (F32C) {
            const float df = (float)d;
            const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(df, df, e2f));
            re += (double)(df * r);
            im += (double)r;
        } else {
            const double r = rcp_f64(__builtin_fma(d, d, e2));
            re = __builtin_fma(d, r, re);
            im += r;
        }
    }
    s_re[t] = re;
    s_im[t] = im;
    __syncthreads();
#pragma unroll
    for (int w = kGridWG / 2; w > 0; w >>= 1) {
        if (t < w) {
            s_re[t] += s_re[t + w];
            s_im[t] += s_im[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out_re[blockIdx.x] = s_re[0] * scale;
        out_im[blockIdx.x] = (eta * s_im[0]) * scale;
    }
}

}  // namespace

int exact_reduce_plain(const double *part, int n_split, size_t nq, double eta, double scale,
                       const double *q, double *out_re, double *out_im, hipStream_t stream) {
    return launch_reduce<EXACT_PLAIN, double>(part, n_split, nq, eta, scale, q, out_re, out_im,
                                              nullptr, nullptr, stream);
}

size_t exact_workspace_doubles(size_t nq, size_t p, ExactKind kind) {
    if (nq == 0 || p == 0) return 0;
    int qt, ns, per;
    plan_splits(nq, p, &qt, &ns, &per);
    return (size_t)((kind == EXACT_DERIV) ? 4 : 2) * (size_t)ns * nq;
}

int exact_f64(const double *q, size_t nq, const double *src, size_t p, double eta,
              ExactKind kind, double cut, double scale, double *const *out,
              hipStream_t stream, double *workspace) {
    if (nq == 0 || p == 0) return 0;
    if (nq > (size_t)INT32_MAX || p > (size_t)INT32_MAX) {
        set_error("exact kernel: sizes above 2^31-1 are not supported");
        return 6;
    }
    if (kind == EXACT_PLAIN && grid_queries(nq, p)) {
        hipLaunchKernelGGL(grid_query_kernel<false>, dim3((unsigned)nq), dim3(kGridWG), 0, stream,
                           q, src, (int)p, eta, scale, out[0], out[1]);
        SHK_HIP(hipGetLastError());
        return 0;
    }
    int qt, ns, per;
    plan_splits(nq, p, &qt, &ns, &per);
    const int na = (kind == EXACT_DERIV) ? 4 : 2;
    double *part = workspace;
    CachedBuf cache(CACHE_EXACT_PART, stream);
    if (!part)
        if (int e = cache.get(na * (size_t)ns * nq, &part)) return e;
    int st = 0;
    switch (kind) {
        case EXACT_PLAIN:
            st = launch_f64<EXACT_PLAIN>(q, nq, src, p, eta, cut, qt, ns, per, part, stream);
            if (!st) st = launch_reduce<EXACT_PLAIN, double>(part, ns, nq, eta, scale, q, out[0], out[1], nullptr, nullptr, stream);
            break;
        case EXACT_CUTOFF:
            st = launch_f64<EXACT_CUTOFF>(q, nq, src, p, eta, cut, qt, ns, per, part, stream);
            if (!st) st = launch_reduce<EXACT_CUTOFF, double>(part, ns, nq, eta, scale, q, out[0], out[1], nullptr, nullptr, stream);
            break;
        case EXACT_HYBRID:
            st = launch_f64<EXACT_HYBRID>(q, nq, src, p, eta, cut, qt, ns, per, part, stream);
            if (!st) st = launch_reduce<EXACT_HYBRID, double>(part, ns, nq, eta, scale, q, out[0], out[1], nullptr, nullptr, stream);
            break;
        case EXACT_DERIV:
            st = launch_f64<EXACT_DERIV>(q, nq, src, p, eta, cut, qt, ns, per, part, stream);
            if (!st) st = launch_reduce<EXACT_DERIV, double>(part, ns, nq, eta, scale, q, out[0], out[1], out[2], out[3], stream);
            break;
        case EXACT_RAY:
            st = launch_f64<EXACT_RAY>(q, nq, src, p, eta, cut, qt, ns, per, part, stream);
            if (!st) st = launch_reduce<EXACT_RAY, double>(part, ns, nq, eta, scale, q, out[0], out[1], nullptr, nullptr, stream);
            break;
        default:
            st = 6;
            set_error("exact kernel: unknown kind");
    }
    return st;
}

int exact_f32c(const double *q, size_t nq, const double *src, size_t p, double eta,
               double scale, void *out_re, void *out_im, int out_f32, hipStream_t stream,
               double *workspace) {
    if (nq == 0 || p == 0) return 0;
    if (nq > (size_t)INT32_MAX || p > (size_t)INT32_MAX) {
        set_error("exact kernel: sizes above 2^31-1 are not supported");
        return 6;
    }
    if (!out_f32 && grid_queries(nq, p)) {
        hipLaunchKernelGGL(grid_query_kernel<true>, dim3((unsigned)nq), dim3(kGridWG), 0, stream,
                           q, src, (int)p, eta, scale, (double *)out_re, (double *)out_im);
        SHK_HIP(hipGetLastError());
        return 0;
    }
    int qt, ns, per;
    plan_splits(nq, p, &qt, &ns, &per);
    double *part = workspace;
    CachedBuf cache(CACHE_EXACT_PART, stream);
    if (!part)
        if (int e = cache.get(2 * (size_t)ns * nq, &part)) return e;
    dim3 grid(qt, ns);
    ProfileMark m = profile_begin(stream);
    hipLaunchKernelGGL(exact_f32c_kernel, grid, dim3(kWG), 0, stream, q, (int)nq, src, (int)p,
                       eta, per, part);
    profile_end(m, stream);
    int st = 0;
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        set_error(std::string("exact_f32c_kernel launch: ") + hipGetErrorString(le));
        st = 7;
    }
    if (!st) {
        if (out_f32)
            st = launch_reduce<EXACT_PLAIN, float>(part, ns, nq, eta, scale, q, (float *)out_re,
                                                   (float *)out_im, nullptr, nullptr, stream);
        else
            st = launch_reduce<EXACT_PLAIN, double>(part, ns, nq, eta, scale, q,
                                                    (double *)out_re, (double *)out_im, nullptr,
                                                    nullptr, stream);
    }
    return st;
}

int exact_many_plan(const int64_t *offsets, const size_t *order, size_t n, const double *etas,
                    int scaled, SegSpectrum *table, int *n_table, int *n_tiles, int r) {
    if (r != 1 && r != 2 && r != 4) {
        set_error("segmented launch: queries per lane must be 1, 2 or 4");
        return 6;
    }
    const int64_t tile_q = (int64_t)kWG * r;
    int64_t tiles = 0;
    int m = 0;
    for (size_t idx = 0; idx < n; ++idx) {
        const size_t k = order[idx];
        const int64_t p = offsets[k + 1] - offsets[k];
        if (p <= 0) continue;
        if (p > (int64_t)INT32_MAX) {
            set_error("exact kernel: sizes above 2^31-1 are not supported");
            return 6;
        }
        int qt, ns, per;
        plan_splits((size_t)p, (size_t)p, &qt, &ns, &per);
        SegSpectrum &e = table[m++];
        e.off = offsets[k];
        e.p = (int32_t)p;
        e.per = per;
        e.tile0 = (int32_t)tiles;
        e.pad = 0;
        e.eta = etas[k];
        e.scale = scaled ? 1.0 / (double)p : 1.0;
        tiles += (p + tile_q - 1) / tile_q;
        if (tiles > (int64_t)INT32_MAX) {
            set_error("segmented launch: more than 2^31-1 query tiles");
            return 6;
        }
    }
    *n_table = m;
    *n_tiles = (int)tiles;
    return 0;
}

namespace {

template <int R>
void launch_many_f32c(const dim3 &grid, const double *ev, const SegSpectrum *d_table,
                      int n_table, double *out_re, double *out_im, long long *d_stamps,
                      hipStream_t stream) {
    if (d_stamps)
        hipLaunchKernelGGL((many_f32c_kernel<R, true>), grid, dim3(kWG), 0, stream, ev, d_table,
                           n_table, out_re, out_im, d_stamps);
    else
        hipLaunchKernelGGL((many_f32c_kernel<R, false>), grid, dim3(kWG), 0, stream, ev, d_table,
                           n_table, out_re, out_im, nullptr);
}

}  // namespace

int exact_many(const double *ev, const SegSpectrum *d_table, int n_table, int n_tiles,
               int f32c, double *out_re, double *out_im, long long *d_stamps,
               hipStream_t stream, int r) {
    if (n_table == 0 || n_tiles == 0) return 0;
    if (r != 1 && r != 2 && r != 4) {
        set_error("segmented launch: queries per lane must be 1, 2 or 4");
        return 6;
    }
    if (!f32c && r != kR) {
        set_error("segmented launch: the FP64 kernel takes 2 queries per lane");
        return 6;
    }
    ProfileMark m = profile_begin(stream);
    const dim3 grid(n_tiles), block(kWG);
    if (f32c && r == 1)
        launch_many_f32c<1>(grid, ev, d_table, n_table, out_re, out_im, d_stamps, stream);
    else if (f32c && r == 2)
        launch_many_f32c<2>(grid, ev, d_table, n_table, out_re, out_im, d_stamps, stream);
    else if (f32c)
        launch_many_f32c<4>(grid, ev, d_table, n_table, out_re, out_im, d_stamps, stream);
    else if (d_stamps)
        hipLaunchKernelGGL(many_f64_kernel<true>, grid, block, 0, stream, ev, d_table, n_table,
                           out_re, out_im, d_stamps);
    else
        hipLaunchKernelGGL(many_f64_kernel<false>, grid, block, 0, stream, ev, d_table, n_table,
                           out_re, out_im, nullptr);
    profile_end(m, stream);
    SHK_HIP(hipGetLastError());
    return 0;
}

int widen_f32(const float *in, double *out, size_t n, hipStream_t stream) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, in,
                       out, n);
    SHK_HIP(hipGetLastError());
    return 0;
}

}  // namespace shk
