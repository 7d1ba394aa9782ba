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
// Each query i of tile t receives exactly one FP64 partial per tile m:
// part[.][m][i] holds the i-side of workgroup (t, m) for m >= t and the
// j-side of workgroup (m, t) for m < t.  reduce_kernel adds the nt partials
// in m order (deterministic) and applies eta and the 1/p scale.
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
// a workgroup barrier instead costs 1-4 %.
#ifndef SHK_SYM_R
#define SHK_SYM_R 8
#endif
#ifndef SHK_SYM_WAVE_SYNC
#define SHK_SYM_WAVE_SYNC 1
#endif
constexpr int kSR = SHK_SYM_R;      // queries per lane (contiguous)
constexpr int kSL = 64;             // lanes per wavefront
constexpr int kSW = 4;              // wavefronts per workgroup
constexpr int kST = kSL * kSR;      // tile: 512 eigenvalues at kSR = 8
constexpr int kSQ = kST / kSW;      // sources per wavefront
constexpr int kSB = 16;             // sources per batch (FP32 run and j-side reduction)
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

template <int FORM, bool JSIDE>
__device__ __forceinline__ void sym_sweep(const double *__restrict__ s_src, int n_src, double c,
                                          const float (&qh)[kSR], const float (&ql)[kSR],
                                          float eta2f, double (&re)[kSR], double (&iv)[kSR],
                                          float *__restrict__ s_jb, double *__restrict__ j_re,
                                          double *__restrict__ j_im, int lane) {
    for (int j0 = 0; j0 < kSQ; j0 += kSB) {  // same trip count in every wave
        float ar[kSR], ai[kSR];
#pragma unroll
        for (int k = 0; k < kSR; ++k) ar[k] = ai[k] = 0.0f;
        for (int u = 0; u < kSB; ++u) {
            const int j = j0 + u;
            if (j < n_src) {  // wave-uniform
                const double t = s_src[j] - c;
                const float sh = (float)t;
                const float sl = (FORM == 2) ? (float)(t - (double)sh) : 0.0f;
                float bre = 0.0f, bim = 0.0f;
#pragma unroll
                for (int k = 0; k < kSR; ++k) {
                    const float d = pair_diff<FORM>(qh[k], ql[k], sh, sl);
                    const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2f));
                    ar[k] = __builtin_fmaf(d, r, ar[k]);
                    ai[k] += r;
                    if (JSIDE) {
                        bre = __builtin_fmaf(d, r, bre);
                        bim += r;
                    }
                }
                if (JSIDE) {
                    s_jb[(2 * u + 0) * kJRow + lane] = bre;
                    s_jb[(2 * u + 1) * kJRow + lane] = bim;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kSR; ++k) {
            re[k] += (double)ar[k];
            iv[k] += (double)ai[k];
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
                float ca = 0.0f, cb = 0.0f;
#pragma unroll
                for (int l = 0; l < 16; ++l) {
                    ca += row[l];
                    cb += row[16 + l];
                }
                const double mine = (double)ca + (double)cb;
                const double other = __shfl_xor(mine, 1, 64);
                const double tot = h ? other + mine : mine + other;
                if (h == 0 && j0 + u < n_src) {
                    if (comp) j_im[j0 + u] = tot;
                    else j_re[j0 + u] = -tot;
                }
            }
            sym_handover();  // the next batch overwrites s_jb
        }
    }
}

// i-side finish: waves 1..3 hand their partial sums to wave 0 in turn (fixed
// order, through `buf`, 2 * kST doubles), and wave 0 writes part[.][b][i].
__device__ __forceinline__ void sym_finish(double (&re)[kSR], double (&iv)[kSR],
                                           double *__restrict__ buf, int w, int lane,
                                           double *__restrict__ part, int nt, int p, int b,
                                           int q0) {
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
        double *i_re = part + (size_t)b * p;          // part[0][b][i]
        double *i_im = part + ((size_t)nt + b) * p;   // part[1][b][i]
#pragma unroll
        for (int k = 0; k < kSR; ++k) {
            const int i = q0 + k;
            if (i < p) {
                i_re[i] = re[k];
                i_im[i] = iv[k];
            }
        }
    }
}

// FP64 variant: the j-side lane sums travel through LDS in FP64, kSB64
// sources per batch (the same LDS bytes as the FP32 batches).
constexpr int kSB64 = kSB / 2;

template <bool JSIDE>
__device__ __forceinline__ void sym_sweep_f64(const double *__restrict__ s_src, int n_src,
                                              const double (&x)[kSR], double e2,
                                              double (&re)[kSR], double (&iv)[kSR],
                                              double *__restrict__ s_jb, double *__restrict__ j_re,
                                              double *__restrict__ j_im, int lane) {
    for (int j0 = 0; j0 < kSQ; j0 += kSB64) {  // same trip count in every wave
        for (int u = 0; u < kSB64; ++u) {
            const int j = j0 + u;
            if (j < n_src) {  // wave-uniform
                const double sv = s_src[j];
                double d[kSR], t[kSR], r[kSR];
#pragma unroll
                for (int k = 0; k < kSR; ++k) {
                    d[k] = x[k] - sv;
                    t[k] = __builtin_fma(d[k], d[k], e2);
                }
                rcp_pairs_f64(t, r);
                double bre = 0.0, bim = 0.0;
#pragma unroll
                for (int k = 0; k < kSR; ++k) {
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
                // four lanes per (source, component) row, one chain of 16
                // each; joined as (c0 + c1) + (c2 + c3) on every lane
                const int row_id = lane >> 2, q = lane & 3;
                const int u = row_id >> 1, comp = row_id & 1;
                const double *row = s_jb + (2 * u + comp) * kJRow + 16 * q;
                double cq = 0.0;
#pragma unroll
                for (int l = 0; l < 16; ++l) cq += row[l];
                const double o1 = __shfl_xor(cq, 1, 64);
                const double pair = (q & 1) ? o1 + cq : cq + o1;   // c0+c1 or c2+c3
                const double o2 = __shfl_xor(pair, 2, 64);
                const double tot = (q & 2) ? o2 + pair : pair + o2;
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
                                                            double eta, int nt,
                                                            double *__restrict__ part) {
    int a, b;  // query tile a, source tile b >= a
    tile_pair_of((int)blockIdx.x, nt, a, b);
    __shared__ double s_src[kSW][kSQ];
    __shared__ double s_jbuf[kSW * 2 * kSB64 * kJRow];  // j-side lane sums / i-side combine
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
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
    double *j_re = part + (size_t)a * p + js;         // part[0][a][j]
    double *j_im = part + ((size_t)nt + a) * p + js;  // part[1][a][j]
    const double e2 = eta * eta;
    if (a < b)
        sym_sweep_f64<true>(s_src[w], n_src, x, e2, re, iv, s_jb, j_re, j_im, lane);
    else
        sym_sweep_f64<false>(s_src[w], n_src, x, e2, re, iv, s_jb, j_re, j_im, lane);
    sym_finish(re, iv, s_jbuf, w, lane, part, nt, p, b, q0);
}

static_assert(kSW * 2 * kSB64 * kJRow >= 2 * kST, "FP64 combine buffer fits in s_jbuf");

__global__ __launch_bounds__(kSL * kSW) void sym_f32c_kernel(const double *__restrict__ ev, int p,
                                                             double eta, int nt,
                                                             double *__restrict__ part) {
    int a, b;  // query tile a, source tile b >= a
    tile_pair_of((int)blockIdx.x, nt, a, b);
    __shared__ double s_src[kSW][kSQ];                      // 8 KB
    __shared__ double s_jbuf[kSW * 2 * kSB * kJRow / 2];    // j-side lane sums / i-side combine
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
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
    double *j_re = part + (size_t)a * p + js;                  // part[0][a][j]
    double *j_im = part + ((size_t)nt + a) * p + js;           // part[1][a][j]
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

    sym_finish(re, iv, s_jbuf, w, lane, part, nt, p, b, q0);
}

static_assert(kSW * 2 * kSB * kJRow / 2 >= 2 * kST, "i-side combine buffer fits in s_jbuf");
static_assert(kSQ % kSB == 0, "batches tile the wave's sources");

}  // namespace

int exact_sym(const double *ev, size_t p, double eta, double scale, int f64, double *out_re,
              double *out_im, hipStream_t stream) {
    if (p == 0) return 0;
    if (p > (size_t)kSymMaxP) {
        set_error("symmetric kernel: p above its maximum");
        return 6;
    }
    const int nt = (int)((p + kST - 1) / kST);
    double *part = nullptr;
    CachedBuf cache(CACHE_EXACT_PART, stream);
    if (int e = cache.get(2 * (size_t)nt * p, &part)) return e;
    ProfileMark m = profile_begin(stream);
    if (f64)
        hipLaunchKernelGGL(sym_f64_kernel, dim3(nt * (nt + 1) / 2), dim3(kSL * kSW), 0, stream,
                           ev, (int)p, eta, nt, part);
    else
        hipLaunchKernelGGL(sym_f32c_kernel, dim3(nt * (nt + 1) / 2), dim3(kSL * kSW), 0, stream,
                           ev, (int)p, eta, nt, part);
    profile_end(m, stream);
    int st = 0;
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        set_error(std::string("symmetric kernel launch: ") + hipGetErrorString(le));
        st = 7;
    }
    if (!st) st = exact_reduce_plain(part, nt, p, eta, scale, ev, out_re, out_im, stream);
    return st;
}

}  // namespace shk
