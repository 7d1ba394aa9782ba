// Symmetric all-points kernel, FP32-compensated (queries are the sources).
//
// Reference: the sequential exact path visits each unordered pair once and
// credits both ends -- Re antisymmetric, Im symmetric, one shared reciprocal
// (symmetric_sweep, src/stieltjes/cacheblock.rs:841-954); its parallel path
// is the full square (:1454-1513).  This kernel takes the half-work form on
// the GPU:
//
//   tiles of kST = 64*kSR (512) eigenvalues; workgroup (a, b), a <= b, pairs
//   query tile a with source tile b.  Each lane holds kSR contiguous queries
//   of tile a with its own FP64 centre c_l (their midpoint), so the pair
//   difference of the FP32-compensated forms (shk_exact.hip) is taken
//   relative to a centre at most a few eigenvalue spacings away; the form is
//   chosen per workgroup from the widest lane.  The four waves split tile
//   b's sources.
//   a < b: every pair (i in a, j in b) adds d r and r to query i (i-side, as
//   the full-square kernel does) and -d r and r to query j (j-side).  A lane
//   sums its kSR pairs of a source in FP32; the 64 lane sums of each source
//   go through LDS and are added, per (source, component), in four FP32
//   chains of 16 (two lanes, two chains each) joined in FP64, once per batch
//   of kSB sources.
//   a = b: the full square of tile a, i-side only (the diagonal term
//   included, d = 0 exactly).
//
// Each query i of tile t receives exactly one FP64 partial per tile m (its
// slot m): the i-side of workgroup (t, m) for m >= t and the j-side of
// workgroup (m, t) for m < t.  The slots are added in m order (deterministic)
// and eta and the 1/p scale applied once (sym_band_reduce_kernel).  When the
// triangle's partials (16 nt p bytes) exceed the scratch budget, the rows of
// the triangle run in bands: after each band the band tiles are finished and
// the tiles beyond it carry running sums of their slots so far, so the
// additions per point keep the order 0..nt-1 (SymBand, sym_run_bands).
#include "shk_internal.hpp"
#include "shk_pair.hpp"

#include <algorithm>

namespace shk {
namespace {

// Tuning knobs (compile-time, for sweeps): queries per lane, and whether the
// j-side lane sums are handed over with a wavefront-local fence (each wave
// owns its LDS rows) instead of a workgroup barrier.
// Defaults measured on MI355X at p = 2e4 / 5e4 / 2e5 (ms per transform):
// 8 per lane with the wave-local handover 0.088 / 0.352 / 4.68; 16 per lane
// 0.159 / 0.370 / 4.68 (too few tile pairs at 5e4 to fill the chip evenly);
// a workgroup barrier instead costs 1-4 %.  SHK_SYM_RUN is the length of the
// FP32 i-side runs: 32 instead of 16 gains under 1 % at 5e4 and raises the
// error against FP64 from 1.8e-8 to 2.3e-8.
#ifndef SHK_SYM_R
#define SHK_SYM_R 8
#endif
#ifndef SHK_SYM_WAVE_SYNC
#define SHK_SYM_WAVE_SYNC 1
#endif
#ifndef SHK_SYM_RUN
#define SHK_SYM_RUN 16
#endif
constexpr int kSR = SHK_SYM_R;      // queries per lane (contiguous)
constexpr int kSL = 64;             // lanes per wavefront
constexpr int kSW = 4;              // wavefronts per workgroup
constexpr int kST = kSL * kSR;      // tile: 512 eigenvalues at kSR = 8
constexpr int kSQ = kST / kSW;      // sources per wavefront
constexpr int kSB = 16;             // sources per batch (j-side reduction)
constexpr int kSRun = SHK_SYM_RUN;  // sources per FP32 i-side run (a multiple of kSB)
constexpr int kJRow = kSL + 1;      // padded LDS row of 64 lane sums (no bank conflicts)

// Makes this wave's LDS writes visible to its own lanes' reads (each wave
// uses its own rows): a wavefront-scope fence pair around a wave barrier, or
// a workgroup barrier.
__device__ __forceinline__ void sym_handover() {
#if SHK_SYM_WAVE_SYNC
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#else
    __syncthreads();
#endif
}

// One source against this lane's kSR queries: the FP32 i-side runs (ar, ai)
// and, when JSIDE, the source's lane sums over the kSR pairs (bre, bim).
// FIRST starts the runs with the batch's first source: a run that starts
// from its first term equals one that starts from 0 and adds it.
template <int FORM, bool JSIDE, bool FIRST>
__device__ __forceinline__ void sym_source(double sv, double c, const float (&qh)[kSR],
                                           const float (&ql)[kSR], float eta2f,
                                           float (&ar)[kSR], float (&ai)[kSR], float &bre,
                                           float &bim) {
    const double t = sv - c;
    const float sh = (float)t;
    const float sl = (FORM == 2) ? (float)(t - (double)sh) : 0.0f;
#pragma unroll
    for (int k = 0; k < kSR; ++k) {
        const float d = pair_diff<FORM>(qh[k], ql[k], sh, sl);
        const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2f));
        if (FIRST) {
            ar[k] = d * r;
            ai[k] = r;
        } else {
            ar[k] = __builtin_fmaf(d, r, ar[k]);
            ai[k] += r;
        }
        if (JSIDE) {
            if (k == 0) {
                bre = d * r;
                bim = r;
            } else {
                bre = __builtin_fmaf(d, r, bre);
                bim += r;
            }
        }
    }
}

// One batch of kSB sources from j0 (cnt of them, cnt <= 0 past the end):
// the i-side runs continue (START: begin with this batch), and with JSIDE the
// lane sums of each source go through LDS and are reduced per (source,
// component).  n_src is wave-uniform (a scalar), so a full batch runs with
// no per-source test; only the last tile's sources can end inside a batch.
template <int FORM, bool JSIDE, bool START>
__device__ __forceinline__ void sym_batch(const double *__restrict__ s_src, int j0, int cnt,
                                          double c, const float (&qh)[kSR],
                                          const float (&ql)[kSR], float eta2f, float (&ar)[kSR],
                                          float (&ai)[kSR], float *__restrict__ s_jb,
                                          double *__restrict__ j_re, double *__restrict__ j_im,
                                          int lane) {
    if (cnt > 0) {
        float bre, bim;
        sym_source<FORM, JSIDE, START>(s_src[j0], c, qh, ql, eta2f, ar, ai, bre, bim);
        if (JSIDE) {
            s_jb[0 * kJRow + lane] = bre;
            s_jb[1 * kJRow + lane] = bim;
        }
        if (cnt == kSB) {
#pragma unroll
            for (int u = 1; u < kSB; ++u) {
                sym_source<FORM, JSIDE, false>(s_src[j0 + u], c, qh, ql, eta2f, ar, ai, bre, bim);
                if (JSIDE) {
                    s_jb[(2 * u + 0) * kJRow + lane] = bre;
                    s_jb[(2 * u + 1) * kJRow + lane] = bim;
                }
            }
        } else {
            for (int u = 1; u < cnt; ++u) {
                sym_source<FORM, JSIDE, false>(s_src[j0 + u], c, qh, ql, eta2f, ar, ai, bre, bim);
                if (JSIDE) {
                    s_jb[(2 * u + 0) * kJRow + lane] = bre;
                    s_jb[(2 * u + 1) * kJRow + lane] = bim;
                }
            }
        }
    }
    if (JSIDE) {
        sym_handover();
        {
            // two lanes per (source, component) row: lane h = 0 adds
            // row[0, 32) and h = 1 row[32, 64), each as two FP32 chains
            // of 16 joined in FP64; the row total is (h0 + h1) in that
            // order on both lanes, ((c0 + c1) + (c2 + c3)) overall
            const int row_id = lane >> 1, h = lane & 1;
            const int u = row_id >> 1, comp = row_id & 1;
            const float *row = s_jb + (2 * u + comp) * kJRow + 32 * h;
            float ca = row[0], cb = row[16];
#pragma unroll
            for (int l = 1; l < 16; ++l) {
                ca += row[l];
                cb += row[16 + l];
            }
            const double mine = (double)ca + (double)cb;
            const double other = __shfl_xor(mine, 1, 64);
            const double tot = h ? other + mine : mine + other;
            if (h == 0 && u < cnt) {
                if (comp) j_im[j0 + u] = tot;
                else j_re[j0 + u] = -tot;
            }
        }
        sym_handover();  // the next batch overwrites s_jb
    }
}

// The wave's sources in i-side runs of kSRun (FP32, folded into FP64 after
// each run), each run in batches of kSB.
template <int FORM, bool JSIDE>
__device__ __forceinline__ void sym_sweep(const double *__restrict__ s_src, int n_src, double c,
                                          const float (&qh)[kSR], const float (&ql)[kSR],
                                          float eta2f, double (&re)[kSR], double (&iv)[kSR],
                                          float *__restrict__ s_jb, double *__restrict__ j_re,
                                          double *__restrict__ j_im, int lane) {
    for (int j0 = 0; j0 < kSQ; j0 += kSRun) {  // same trip count in every wave
        float ar[kSR], ai[kSR];
        const int cnt0 = min(kSB, n_src - j0);
        sym_batch<FORM, JSIDE, true>(s_src, j0, cnt0, c, qh, ql, eta2f, ar, ai, s_jb, j_re, j_im,
                                     lane);
#pragma unroll
        for (int q = 1; q < kSRun / kSB; ++q) {
            const int jq = j0 + q * kSB;
            sym_batch<FORM, JSIDE, false>(s_src, jq, min(kSB, n_src - jq), c, qh, ql, eta2f, ar,
                                          ai, s_jb, j_re, j_im, lane);
        }
        if (cnt0 > 0) {
#pragma unroll
            for (int k = 0; k < kSR; ++k) {
                re[k] += (double)ar[k];
                iv[k] += (double)ai[k];
            }
        }
    }
}

// i-side finish: waves 1..3 hand their partial sums to wave 0 in turn (fixed
// order, through `buf`, 2 * kST doubles), and wave 0 writes tile a's slot b
// (i_re / i_im point at the tile's first point).
__device__ __forceinline__ void sym_finish(double (&re)[kSR], double (&iv)[kSR],
                                           double *__restrict__ buf, int w, int lane,
                                           double *__restrict__ i_re, double *__restrict__ i_im,
                                           int p, int q0) {
    for (int r = 1; r < kSW; ++r) {
        __syncthreads();
        if (w == r) {
#pragma unroll
            for (int k = 0; k < kSR; ++k) {
                buf[lane * kSR + k] = re[k];
                buf[kST + lane * kSR + k] = iv[k];
            }
        }
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int k = 0; k < kSR; ++k) {
                re[k] += buf[lane * kSR + k];
                iv[k] += buf[kST + lane * kSR + k];
            }
        }
    }
    if (w == 0) {
        const int l0 = lane * kSR;  // this lane's first query within the tile
#pragma unroll
        for (int k = 0; k < kSR; ++k) {
            if (q0 + k < p) {
                i_re[l0 + k] = re[k];
                i_im[l0 + k] = iv[k];
            }
        }
    }
}

// FP64 variant: the j-side lane sums travel through LDS in FP64, kSB64
// sources per batch (8: the same LDS bytes as the FP32 batches).  Measured
// on MI355X (FP64 all-points, ms at p = 2e4 / 5e4 / 1e5): 8 per batch
// 0.118 / 0.461 / 1.78; 4 (half the LDS, 7 instead of 4 workgroups per CU)
// 0.121 / 0.477 / 1.82; 16: 0.123 / 0.504 / 1.94.
#ifndef SHK_SYM_SB64
#define SHK_SYM_SB64 8
#endif
constexpr int kSB64 = SHK_SYM_SB64;
static_assert(kSB64 == 4 || kSB64 == 8 || kSB64 == 16, "a batch's rows split the wavefront evenly");

template <bool JSIDE>
__device__ __forceinline__ void sym_sweep_f64(const double *__restrict__ s_src, int n_src,
                                              const double (&x)[kSR], double e2,
                                              double (&re)[kSR], double (&iv)[kSR],
                                              double *__restrict__ s_jb, double *__restrict__ j_re,
                                              double *__restrict__ j_im, int lane) {
    for (int j0 = 0; j0 < kSQ; j0 += kSB64) {  // same trip count in every wave
        for (int u = 0; u < kSB64; ++u) {
            const int j = j0 + u;
            if (j < n_src) {  // wave-uniform (a scalar test)
                const double sv = s_src[j];
                double d[kSR], t[kSR], r[kSR];
#pragma unroll
                for (int k = 0; k < kSR; ++k) {
                    d[k] = x[k] - sv;
                    t[k] = __builtin_fma(d[k], d[k], e2);
                }
                rcp_many_f64(t, r);
                // the lane sums start from the first pair's term
                double bre = d[0] * r[0], bim = r[0];
                re[0] = __builtin_fma(d[0], r[0], re[0]);
                iv[0] += r[0];
#pragma unroll
                for (int k = 1; k < kSR; ++k) {
                    re[k] = __builtin_fma(d[k], r[k], re[k]);
                    iv[k] += r[k];
                    if (JSIDE) {
                        bre = __builtin_fma(d[k], r[k], bre);
                        bim += r[k];
                    }
                }
                if (JSIDE) {
                    s_jb[(2 * u + 0) * kJRow + lane] = bre;
                    s_jb[(2 * u + 1) * kJRow + lane] = bim;
                }
            }
        }
        if (JSIDE) {
            sym_handover();
            {
                // kL lanes per (source, component) row, one chain of kC
                // each, joined in a fixed tree on every lane: at kSB64 = 8,
                // four chains of 16 as (c0 + c1) + (c2 + c3)
                constexpr int kL = kSL / (2 * kSB64), kC = kSL / kL;
                const int row_id = lane / kL, q = lane % kL;
                const int u = row_id >> 1, comp = row_id & 1;
                const double *row = s_jb + (2 * u + comp) * kJRow + kC * q;
                double tot = row[0];
#pragma unroll
                for (int l = 1; l < kC; ++l) tot += row[l];
#pragma unroll
                for (int o = 1; o < kL; o <<= 1) {
                    const double other = __shfl_xor(tot, o, 64);
                    tot = (q & o) ? other + tot : tot + other;
                }
                if (q == 0 && j0 + u < n_src) {
                    if (comp) j_im[j0 + u] = tot;
                    else j_re[j0 + u] = -tot;
                }
            }
            sym_handover();  // the next batch overwrites s_jb
        }
    }
}

__global__ __launch_bounds__(kSL * kSW) void sym_f64_kernel(const double *__restrict__ ev, int p,
                                                            double eta, SymBand bd, int id0) {
    int a, b;  // query tile a (in the band), source tile b >= a
    tile_pair_of((int)blockIdx.x + id0, bd.nt, a, b);
    __shared__ double s_src[kSW][kSQ];
    __shared__ double s_jbuf[kSW * 2 * kSB64 * kJRow];  // j-side lane sums / i-side combine
    // w is wave-uniform; in a scalar register the source counts and batch
    // tests below stay scalar
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int q0 = a * kST + lane * kSR;
    double x[kSR], re[kSR], iv[kSR];
#pragma unroll
    for (int k = 0; k < kSR; ++k) {
        const int i = q0 + k;
        x[k] = (i < p) ? ev[i] : 0.0;  // only the last tile (a = b) is partial
        re[k] = iv[k] = 0.0;
    }
    const int js = b * kST + w * kSQ;
    const int n_src = max(0, min(kSQ, p - js));
    for (int u = lane; u < n_src; u += kSL) s_src[w][u] = ev[js + u];
    __syncthreads();
    double *s_jb = s_jbuf + w * (2 * kSB64 * kJRow);
    double *j_re = bd.slot(0, b, a) + w * kSQ;  // tile b, slot a, this wave's sources
    double *j_im = bd.slot(1, b, a) + w * kSQ;
    const double e2 = eta * eta;
    if (a < b)
        sym_sweep_f64<true>(s_src[w], n_src, x, e2, re, iv, s_jb, j_re, j_im, lane);
    else
        sym_sweep_f64<false>(s_src[w], n_src, x, e2, re, iv, s_jb, j_re, j_im, lane);
    sym_finish(re, iv, s_jbuf, w, lane, bd.slot(0, a, b), bd.slot(1, a, b), p, q0);
}

static_assert(kSW * 2 * kSB64 * kJRow >= 2 * kST, "FP64 combine buffer fits in s_jbuf");

// Tile pair (a, b), a <= b, of one spectrum (ev, p) into the band bd: the
// body of the single-spectrum and the segmented FP32-compensated kernels.
__device__ __forceinline__ void sym_f32c_pair(const double *__restrict__ ev, int p, double eta,
                                              const SymBand &bd, int a, int b) {
    __shared__ double s_src[kSW][kSQ];                      // 8 KB
    __shared__ double s_jbuf[kSW * 2 * kSB * kJRow / 2];    // j-side lane sums / i-side combine
    // w is wave-uniform; in a scalar register the source counts and batch
    // tests below stay scalar
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const bool jside = a < b;

    // queries: kSR contiguous per lane; only the last tile can be partial, and
    // it only meets itself (a = b), where no j-side is formed
    const int q0 = a * kST + lane * kSR;
    double x[kSR];
    double mn = __builtin_inf(), mx = -__builtin_inf();
#pragma unroll
    for (int k = 0; k < kSR; ++k) {
        const int i = q0 + k;
        x[k] = (i < p) ? ev[i] : 0.0;
        if (i < p) {
            mn = fmin(mn, x[k]);
            mx = fmax(mx, x[k]);
        }
    }
    const bool any = mn <= mx;
    const double c = any ? 0.5 * (mn + mx) : 0.0;
    const double half_w = any ? 0.5 * (mx - mn) : 0.0;
    const double wmax = wave_max(half_w);  // every wave holds the same queries
    const int form = f32c_form(wmax, eta);
    float qh[kSR], ql[kSR];
    double re[kSR], iv[kSR];
#pragma unroll
    for (int k = 0; k < kSR; ++k) {
        const double t = x[k] - c;
        qh[k] = (float)t;
        ql[k] = (float)(t - (double)qh[k]);
        re[k] = iv[k] = 0.0;
    }

    // this wave's sources: tile b, quarter w
    const int js = b * kST + w * kSQ;
    const int n_src = max(0, min(kSQ, p - js));
    for (int u = lane; u < n_src; u += kSL) s_src[w][u] = ev[js + u];
    __syncthreads();

    float *s_jb = reinterpret_cast<float *>(s_jbuf) + w * (2 * kSB * kJRow);
    double *j_re = bd.slot(0, b, a) + w * kSQ;  // tile b, slot a, this wave's sources
    double *j_im = bd.slot(1, b, a) + w * kSQ;
    const float eta2f = (float)(eta * eta);
    if (jside) {
        if (form == 0)
            sym_sweep<0, true>(s_src[w], n_src, c, qh, ql, eta2f, re, iv, s_jb, j_re, j_im, lane);
        else if (form == 1)
            sym_sweep<1, true>(s_src[w], n_src, c, qh, ql, eta2f, re, iv, s_jb, j_re, j_im, lane);
        else
            sym_sweep<2, true>(s_src[w], n_src, c, qh, ql, eta2f, re, iv, s_jb, j_re, j_im, lane);
    } else {
        if (form == 0)
            sym_sweep<0, false>(s_src[w], n_src, c, qh, ql, eta2f, re, iv, s_jb, j_re, j_im, lane);
        else if (form == 1)
            sym_sweep<1, false>(s_src[w], n_src, c, qh, ql, eta2f, re, iv, s_jb, j_re, j_im, lane);
        else
            sym_sweep<2, false>(s_src[w], n_src, c, qh, ql, eta2f, re, iv, s_jb, j_re, j_im, lane);
    }

    sym_finish(re, iv, s_jbuf, w, lane, bd.slot(0, a, b), bd.slot(1, a, b), p, q0);
}

__global__ __launch_bounds__(kSL * kSW) void sym_f32c_kernel(const double *__restrict__ ev, int p,
                                                             double eta, SymBand bd, int id0) {
    int a, b;  // query tile a (in the band), source tile b >= a
    tile_pair_of((int)blockIdx.x + id0, bd.nt, a, b);
    sym_f32c_pair(ev, p, eta, bd, a, b);
}

static_assert(kSW * 2 * kSB * kJRow / 2 >= 2 * kST, "i-side combine buffer fits in s_jbuf");
static_assert(kSQ % kSRun == 0 && kSRun % kSB == 0, "runs and batches tile the wave's sources");
static_assert(kST == kSymTile, "the band layout's tile");

// The in-order fold of one band's partials, one thread per point of
// [r0 T, p): band points add their running sums (none in the first band) and
// slots r0..nt-1 and store out0 = s0 * scale, out1 = (mul * s1) * scale;
// points beyond the band add slots r0..r1-1 to their running sums
// (acc[0, p) and acc[p, 2p)).  Loads run kFoldAhead slots ahead of the
// additions.  A group of 64 points never straddles a tile.
constexpr int kFoldWG = 64;
constexpr int kFoldAhead = 8;

// Adds slots 0..n-1 (stride apart) of both components to s0 and s1 in slot
// order, loading kFoldAhead slots ahead of the additions.
__device__ __forceinline__ void fold_slots(const double *__restrict__ c0,
                                           const double *__restrict__ c1, int n, size_t stride,
                                           double &s0, double &s1) {
    int m = 0;
    for (; m + kFoldAhead <= n; m += kFoldAhead) {
        double v0[kFoldAhead], v1[kFoldAhead];
#pragma unroll
        for (int u = 0; u < kFoldAhead; ++u) {
            v0[u] = c0[(size_t)(m + u) * stride];
            v1[u] = c1[(size_t)(m + u) * stride];
        }
#pragma unroll
        for (int u = 0; u < kFoldAhead; ++u) {
            s0 += v0[u];
            s1 += v1[u];
        }
    }
    for (; m < n; ++m) {
        s0 += c0[(size_t)m * stride];
        s1 += c1[(size_t)m * stride];
    }
}

__global__ __launch_bounds__(kFoldWG) void sym_band_reduce_kernel(SymBand bd, int p, double mul,
                                                                 double scale,
                                                                 double *__restrict__ acc,
                                                                 double *__restrict__ o0,
                                                                 double *__restrict__ o1) {
    const int i = bd.r0 * kSymTile + (int)(blockIdx.x * kFoldWG + threadIdx.x);
    if (i >= p) return;
    const int t = i / kSymTile, loc = i - t * kSymTile;
    const bool band = t < bd.r1;
    const int n = (band ? bd.nt : bd.r1) - bd.r0;
    const size_t stride = band ? (size_t)bd.na : (size_t)bd.nb;
    const double *c0 = bd.slot(0, t, bd.r0) + loc;
    const double *c1 = bd.slot(1, t, bd.r0) + loc;
    double s0 = 0.0, s1 = 0.0;
    if (bd.r0 > 0) {
        s0 = acc[i];
        s1 = acc[(size_t)p + i];
    }
    fold_slots(c0, c1, n, stride, s0, s1);
    if (band) {
        o0[i] = s0 * scale;
        o1[i] = (mul * s1) * scale;
    } else {
        acc[i] = s0;
        acc[(size_t)p + i] = s1;
    }
}

// ---- segmented launch: many spectra, one group of them per launch ---------
//
// The workgroups of a group are its spectra's tile pairs, spectrum by
// spectrum (largest first); each spectrum's partials take one band of the
// band layout (r0 = 0, r1 = nt) at its own offset in the scratch.  The fold
// adds each point's slots in order and scales as sym_band_reduce_kernel does
// for a single band, so every result equals exact_sym's bit for bit.

// The last entry of specs[0, n) whose member `first` is <= id.
template <int32_t SymSegSpectrum::*FIRST>
__device__ __forceinline__ int sym_seg_find(const SymSegSpectrum *__restrict__ specs, int n,
                                            int id) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (specs[mid].*FIRST <= id) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ SymBand sym_seg_band(const SymSegSpectrum &sp, double *scratch) {
    return SymBand{scratch + sp.part, nullptr, 0, sp.nt, sp.nt, sp.p, 0};
}

template <bool STAMP>
__global__ __launch_bounds__(kSL * kSW) void sym_many_f32c_kernel(
    const double *__restrict__ ev, const SymSegSpectrum *__restrict__ specs, int n_specs,
    double *__restrict__ scratch, long long *__restrict__ stamps) {
    if (STAMP && threadIdx.x == 0)
        stamps[2 * (size_t)blockIdx.x] = (long long)wall_clock64();
    const SymSegSpectrum sp =
        specs[sym_seg_find<&SymSegSpectrum::pair0>(specs, n_specs, (int)blockIdx.x)];
    int a, b;
    tile_pair_of((int)blockIdx.x - sp.pair0, sp.nt, a, b);
    sym_f32c_pair(ev + sp.off, sp.p, sp.eta, sym_seg_band(sp, scratch), a, b);
    if (STAMP) {
        __syncthreads();  // the end stamp follows every lane's stores
        if (threadIdx.x == 0) stamps[2 * (size_t)blockIdx.x + 1] = (long long)wall_clock64();
    }
}

__global__ __launch_bounds__(kFoldWG) void sym_many_fold_kernel(
    const SymSegSpectrum *__restrict__ specs, int n_specs, const double *__restrict__ scratch,
    double *__restrict__ out_re, double *__restrict__ out_im) {
    const SymSegSpectrum sp =
        specs[sym_seg_find<&SymSegSpectrum::fold0>(specs, n_specs, (int)blockIdx.x)];
    const int i = ((int)blockIdx.x - sp.fold0) * kFoldWG + (int)threadIdx.x;
    if (i >= sp.p) return;
    const int t = i / kSymTile, loc = i - t * kSymTile;
    const SymBand bd = sym_seg_band(sp, const_cast<double *>(scratch));
    double s0 = 0.0, s1 = 0.0;
    fold_slots(bd.slot(0, t, 0) + loc, bd.slot(1, t, 0) + loc, sp.nt, (size_t)sp.p, s0, s1);
    out_re[sp.off + i] = s0 * sp.scale;
    out_im[sp.off + i] = (sp.eta * s1) * sp.scale;
}

// index of row x's first tile pair in the row-major triangle of nt tiles
inline int64_t row_start(int64_t x, int64_t nt) { return x * (2 * nt - x + 1) / 2; }

struct SymCtx {
    const double *ev;
    int p;
    double eta;
    int f64;
    hipStream_t stream;
};

int sym_launch(const SymBand &bd, int grid, int id0, void *vctx) {
    const SymCtx &c = *static_cast<const SymCtx *>(vctx);
    ProfileMark pm = profile_begin(c.stream);
    if (c.f64)
        hipLaunchKernelGGL(sym_f64_kernel, dim3(grid), dim3(kSL * kSW), 0, c.stream, c.ev, c.p,
                           c.eta, bd, id0);
    else
        hipLaunchKernelGGL(sym_f32c_kernel, dim3(grid), dim3(kSL * kSW), 0, c.stream, c.ev, c.p,
                           c.eta, bd, id0);
    profile_end(pm, c.stream);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        set_error(std::string("symmetric kernel launch: ") + hipGetErrorString(le));
        return 7;
    }
    return 0;
}

}  // namespace

int sym_band_rows(size_t p, size_t avail) {
    const size_t nt = (p + kSymTile - 1) / kSymTile;
    if (2 * nt * p <= avail) return (int)nt;
    const size_t per_row = 2 * nt * kSymTile + 2 * p;  // A and B per band row
    if (avail < 2 * p + per_row) return 0;
    return (int)std::min(nt, (avail - 2 * p) / per_row);
}

size_t sym_scratch_doubles(size_t p, int R) {
    const size_t nt = (p + kSymTile - 1) / kSymTile;
    if ((size_t)R >= nt) return 2 * nt * p;
    return 2 * nt * std::min(p, (size_t)R * kSymTile) + 2 * (size_t)R * p + 2 * p;
}

int sym_run_bands(size_t p, double *scratch, size_t avail, double mul, double scale,
                  double *out0, double *out1, SymLaunch launch, void *ctx, hipStream_t stream) {
    if (p == 0) return 0;
    const int R = sym_band_rows(p, avail);
    if (R <= 0) {
        set_error("symmetric kernel: scratch too small for one band");
        return 6;
    }
    const int nt = (int)((p + kSymTile - 1) / kSymTile);
    // A first (the largest band's), then B, then the running sums
    const size_t a_max = 2 * (size_t)nt * std::min(p, (size_t)R * kSymTile);
    double *A = scratch, *B = scratch + a_max;
    double *acc = B + 2 * (size_t)R * p;
    for (int r0 = 0; r0 < nt; r0 += R) {
        const int r1 = std::min(nt, r0 + R);
        const size_t lo = (size_t)r0 * kSymTile, mid = std::min(p, (size_t)r1 * kSymTile);
        const SymBand bd{A, B, r0, r1, nt, (int)(mid - lo), (int)(p - mid)};
        const int64_t id0 = row_start(r0, nt), id1 = row_start(r1, nt);
        if (int st = launch(bd, (int)(id1 - id0), (int)id0, ctx)) return st;
        hipLaunchKernelGGL(sym_band_reduce_kernel,
                           dim3((unsigned)((p - lo + kFoldWG - 1) / kFoldWG)), dim3(kFoldWG), 0,
                           stream, bd, (int)p, mul, scale, acc, out0, out1);
        SHK_HIP(hipGetLastError());
    }
    return 0;
}

int exact_sym(const double *ev, size_t p, double eta, double scale, int f64, double *out_re,
              double *out_im, hipStream_t stream) {
    if (p == 0) return 0;
    if (p > (size_t)kSymMaxP) {
        set_error("symmetric kernel: p above its maximum");
        return 6;
    }
    const int R = std::max(1, sym_band_rows(p, sym_scratch_budget() / sizeof(double)));
    const size_t need = sym_scratch_doubles(p, R);
    double *part = nullptr;
    CachedBuf cache(CACHE_EXACT_PART, stream);
    if (int e = cache.get(need, &part)) return e;
    return exact_sym_on(ev, p, eta, scale, f64, out_re, out_im, part, need, stream);
}

int exact_sym_on(const double *ev, size_t p, double eta, double scale, int f64, double *out_re,
                 double *out_im, double *scratch, size_t scratch_doubles, hipStream_t stream) {
    if (p == 0) return 0;
    if (p > (size_t)kSymMaxP) {
        set_error("symmetric kernel: p above its maximum");
        return 6;
    }
    SymCtx ctx{ev, (int)p, eta, f64, stream};
    return sym_run_bands(p, scratch, scratch_doubles, eta, scale, out_re, out_im, sym_launch,
                         &ctx, stream);
}

int sym_many_plan(const int64_t *offsets, const size_t *order, size_t n, const double *etas,
                  int scaled, size_t budget_doubles, SymSegSpectrum *table,
                  std::vector<SymSegGroup> &groups, int64_t *n_pairs) {
    groups.clear();
    int64_t pair_base = 0;
    SymSegGroup g{0, 0, 0, 0, 0, 0};
    auto close = [&]() {
        if (g.count > 0) {
            groups.push_back(g);
            pair_base += g.pairs;
        }
        g = SymSegGroup{g.first + g.count, 0, 0, 0, pair_base, 0};
    };
    for (size_t idx = 0; idx < n; ++idx) {
        const size_t k = order[idx];
        const int64_t p = offsets[k + 1] - offsets[k];
        if (p <= 0 || p > (int64_t)kSymMaxP) {
            set_error("segmented symmetric launch: every spectrum needs 1 <= p <= 4194304");
            return 6;
        }
        const int64_t nt = (p + kSymTile - 1) / kSymTile;
        const size_t part = 2 * (size_t)nt * (size_t)p;
        if (part > budget_doubles) {
            set_error("segmented symmetric launch: the partial sums of a spectrum of " +
                      std::to_string(p) + " eigenvalues exceed the symmetric scratch limit");
            return 6;
        }
        const int64_t pairs = nt * (nt + 1) / 2, folds = (p + kFoldWG - 1) / kFoldWG;
        if (g.part_doubles + part > budget_doubles || g.pairs + pairs > (int64_t)INT32_MAX ||
            g.folds + folds > (int64_t)INT32_MAX)
            close();
        SymSegSpectrum &e = table[idx];
        e.off = offsets[k];
        e.part = (int64_t)g.part_doubles;
        e.p = (int32_t)p;
        e.nt = (int32_t)nt;
        e.pair0 = g.pairs;
        e.fold0 = g.folds;
        e.eta = etas[k];
        e.scale = scaled ? 1.0 / (double)p : 1.0;
        g.count += 1;
        g.pairs += (int)pairs;
        g.folds += (int)folds;
        g.part_doubles += part;
    }
    close();
    *n_pairs = pair_base;
    return 0;
}

int sym_many(const double *ev, const SymSegSpectrum *d_table,
             const std::vector<SymSegGroup> &groups, double *scratch, double *out_re,
             double *out_im, long long *d_stamps, hipStream_t stream) {
    for (const SymSegGroup &g : groups) {
        ProfileMark pm = profile_begin(stream);
        const SymSegSpectrum *tab = d_table + g.first;
        if (d_stamps)
            hipLaunchKernelGGL(sym_many_f32c_kernel<true>, dim3(g.pairs), dim3(kSL * kSW), 0,
                               stream, ev, tab, g.count, scratch, d_stamps + 2 * g.pair_base);
        else
            hipLaunchKernelGGL(sym_many_f32c_kernel<false>, dim3(g.pairs), dim3(kSL * kSW), 0,
                               stream, ev, tab, g.count, scratch, nullptr);
        profile_end(pm, stream);
        SHK_HIP(hipGetLastError());
        hipLaunchKernelGGL(sym_many_fold_kernel, dim3(g.folds), dim3(kFoldWG), 0, stream, tab,
                           g.count, scratch, out_re, out_im);
        SHK_HIP(hipGetLastError());
    }
    return 0;
}

}  // namespace shk
