// Symmetric all-points adjoint: the gradient of a real loss of the raw
// all-points sums S_i = sum_j 1/((lam_i - lam_j) - i eta) (diagonal
// included) with respect to every eigenvalue -- through the query and the
// source role together -- and with respect to eta, evaluating each unordered
// pair once.
//
// Reference: only dS/dx exists there (stieltjes_with_deriv_sum,
// src/stieltjes/autovec.rs:38-59); the half-work pair visit is the one of the
// sequential exact path (symmetric_sweep, src/stieltjes/cacheblock.rs:
// 841-954), applied here to the adjoint.
//
// With upstream gradients (Gr_i, Gi_i) = (dL/dRe S_i, dL/dIm S_i) and, for a
// pair (i, j), d = lam_i - lam_j, D = d^2 + eta^2,
//   alpha = (d^2 - eta^2) / D^2    (symmetric in i, j)
//   beta  = 2 eta d / D^2          (antisymmetric),
// the pair adds T = alpha Gr_j - beta Gi_j - alpha Gr_i - beta Gi_i to
// dL/dlam_i and -T to dL/dlam_j (i as query of source j and as source of
// query j; the diagonal gives T = 0 and is left out of the gradient sums,
// where its alpha = -1/eta^2 would otherwise cancel only to rounding), and
// the ordered pair (i, j) adds
// Gi_i alpha - Gr_i beta to dL/deta.  So each end needs only three sums over
// its partners: for end i (sums over j)
//   P = sum alpha Gr_j - beta' Gi'_j,  A = sum alpha,  B = sum beta'
// with beta' = d / D^2, Gi' = 2 eta Gi, giving
//   dL/dlam_i += P - Gr_i A - Gi'_i B,   dL/deta += Gi_i A - 2 eta Gr_i B;
// for end j (sums over i, d still lam_i - lam_j)
//   Q = sum alpha Gr_i + beta' Gi'_i,  A' = sum alpha,  B' = sum beta'
//   dL/dlam_j += Q - Gr_j A' + Gi'_j B',  dL/deta += Gi_j A' + 2 eta Gr_j B'.
//
// Layout (as the symmetric forward kernels): tiles of 512 eigenvalues;
// workgroup (a, b), a <= b, holds query tile a, eight contiguous queries per
// lane, and the four waves split source tile b.  a < b: i-side sums in the
// lanes, j-side lane sums of each source through LDS, reduced per batch in a
// fixed order.  a = b: the full square of tile a, i-side only.  Each end
// gets one (gradient, eta term) partial per tile m (its slot m), the i-side
// of workgroup (t, m) for m >= t and the j-side of (m, t) for m < t; the
// in-order reduction and an ordered sum of the eta terms finish.  Every sum
// runs in a fixed order: results are bitwise reproducible.
//
// Per pair: r = 1/D, alpha = r - 2 eta^2 r^2 (one fused multiply-add),
// beta' = d r^2, then two fused multiply-adds and two additions at each end.
// FP32-compensated form: the pair difference of the forward kernels
// (per-lane FP64 centre, f32 hi/lo parts); runs of kAB sources summed in f32
// and folded into FP64.  FP64 form: reciprocals four per v_rcp_f64 (rcp_many_f64).
#include "shk_internal.hpp"
#include "shk_pair.hpp"

#include <algorithm>
#include <string>

namespace shk {
namespace {

constexpr int kJR = 8;              // queries per lane (contiguous)
constexpr int kJL = 64;             // lanes per wavefront
constexpr int kJW = 4;              // wavefronts per workgroup
constexpr int kJT = kJL * kJR;      // tile: 512 eigenvalues
constexpr int kJQ = kJT / kJW;      // sources per wavefront
constexpr int kAB = 8;              // sources per batch, FP32 form
constexpr int kAB64 = 4;            // sources per batch, FP64 form (same LDS bytes)
constexpr int kARow = kJL + 1;      // padded LDS row of 64 lane sums
constexpr int kComp = 3;            // j-side sums per source: Q, A', B'

__device__ __forceinline__ void wave_handover() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// j-side partials of source j from its three sums (FP64, fixed order).
__device__ __forceinline__ void jside_store(double Q, double Ap, double Bp, double g_r,
                                            double g_i, double two_eta, double *grad,
                                            double *eta_t) {
    *grad = (Q - g_r * Ap) + (two_eta * g_i) * Bp;
    *eta_t = g_i * Ap + (two_eta * g_r) * Bp;
}

// One source against this lane's kJR queries: the FP32 i-side runs (rp, ra,
// rb) and, when JSIDE, the source's lane sums (bq, ba, bb).  FIRST starts the
// runs with the batch's first source: a run that starts from its first term
// equals one that starts from 0 and adds it.
template <int FORM, bool JSIDE, bool FIRST>
__device__ __forceinline__ void adj_source_f32c(double sv, float grj, float gij, int j, double c,
                                                const float (&qh)[kJR], const float (&ql)[kJR],
                                                const float (&gr)[kJR], const float (&gi)[kJR],
                                                float eta2f, float m2eta2f, float (&rp)[kJR],
                                                float (&ra)[kJR], float (&rb)[kJR], float &bq,
                                                float &ba, float &bb, int qoff) {
    const double t = sv - c;
    const float sh = (float)t;
    const float sl = (FORM == 2) ? (float)(t - (double)sh) : 0.0f;
#pragma unroll
    for (int k = 0; k < kJR; ++k) {
        const float d = pair_diff<FORM>(qh[k], ql[k], sh, sl);
        const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2f));
        const float r2 = r * r;
        float al = __builtin_fmaf(m2eta2f, r2, r);  // (d^2 - eta^2) / D^2
        float be = d * r2;
        if (!JSIDE) {
            // the pair (i, i): T = 0, so it adds nothing.  Left in, its
            // alpha = -1/eta^2 would enter P and A and cancel only in
            // P - Gr_i A, to the rounding of 1/eta^2.
            const bool self = qoff + k == j;
            al = self ? 0.0f : al;
            be = self ? 0.0f : be;
        }
        if (FIRST) {
            rp[k] = __builtin_fmaf(-be, gij, al * grj);
            ra[k] = al;
            rb[k] = be;
        } else {
            rp[k] = __builtin_fmaf(al, grj, rp[k]);
            rp[k] = __builtin_fmaf(-be, gij, rp[k]);
            ra[k] += al;
            rb[k] += be;
        }
        if (JSIDE) {
            if (k == 0) {
                bq = __builtin_fmaf(be, gi[k], al * gr[k]);
                ba = al;
                bb = be;
            } else {
                bq = __builtin_fmaf(al, gr[k], bq);
                bq = __builtin_fmaf(be, gi[k], bq);
                ba += al;
                bb += be;
            }
        }
    }
}

// The wave's sources in batches of kAB.  n_src is wave-uniform (a scalar), so
// a full batch runs with no per-source test; only the last tile's sources can
// end inside a batch.
template <int FORM, bool JSIDE>
__device__ __forceinline__ void adj_sweep_f32c(
    const double *__restrict__ s_src, const float *__restrict__ s_gr,
    const float *__restrict__ s_gi, int n_src, double c, const float (&qh)[kJR],
    const float (&ql)[kJR], const float (&gr)[kJR], const float (&gi)[kJR], float eta2f,
    float m2eta2f, double (&P)[kJR], double (&A)[kJR], double (&B)[kJR], float *__restrict__ s_jb,
    const double *__restrict__ g_re, const double *__restrict__ g_im, double two_eta,
    double *__restrict__ j_grad, double *__restrict__ j_eta, int lane, int qoff) {
    for (int j0 = 0; j0 < kJQ; j0 += kAB) {  // same trip count in every wave
        const int cnt = min(kAB, n_src - j0);
        if (cnt > 0) {
            float rp[kJR], ra[kJR], rb[kJR], bq, ba, bb;
            auto store = [&](int u) {
                if (JSIDE) {
                    s_jb[(kComp * u + 0) * kARow + lane] = bq;
                    s_jb[(kComp * u + 1) * kARow + lane] = ba;
                    s_jb[(kComp * u + 2) * kARow + lane] = bb;
                }
            };
            adj_source_f32c<FORM, JSIDE, true>(s_src[j0], s_gr[j0], s_gi[j0], j0, c, qh, ql, gr,
                                               gi, eta2f, m2eta2f, rp, ra, rb, bq, ba, bb, qoff);
            store(0);
            if (cnt == kAB) {
#pragma unroll
                for (int u = 1; u < kAB; ++u) {
                    const int j = j0 + u;
                    adj_source_f32c<FORM, JSIDE, false>(s_src[j], s_gr[j], s_gi[j], j, c, qh, ql,
                                                        gr, gi, eta2f, m2eta2f, rp, ra, rb, bq,
                                                        ba, bb, qoff);
                    store(u);
                }
            } else {
                for (int u = 1; u < cnt; ++u) {
                    const int j = j0 + u;
                    adj_source_f32c<FORM, JSIDE, false>(s_src[j], s_gr[j], s_gi[j], j, c, qh, ql,
                                                        gr, gi, eta2f, m2eta2f, rp, ra, rb, bq,
                                                        ba, bb, qoff);
                    store(u);
                }
            }
#pragma unroll
            for (int k = 0; k < kJR; ++k) {
                P[k] += (double)rp[k];
                A[k] += (double)ra[k];
                B[k] += (double)rb[k];
            }
        }
        if (JSIDE) {
            wave_handover();
            // rows (u, comp) = lane >> 1 for lanes < 48, two lanes per row
            // (h = 0: row[0, 32), h = 1: row[32, 64)), each as two f32 chains
            // of 16 joined in FP64; the row total is (h0 + h1) on both lanes
            const int row_id = lane >> 1, h = lane & 1;
            double mine = 0.0;
            if (row_id < kComp * kAB) {
                const float *row = s_jb + row_id * kARow + 32 * h;
                float ca = row[0], cb = row[16];
#pragma unroll
                for (int l = 1; l < 16; ++l) {
                    ca += row[l];
                    cb += row[16 + l];
                }
                mine = (double)ca + (double)cb;
            }
            const double other = __shfl_xor(mine, 1, 64);
            const double tot = h ? other + mine : mine + other;
            // source u's sums sit on lanes 6u, 6u + 2, 6u + 4
            const double t1 = __shfl_down(tot, 2, 64), t2 = __shfl_down(tot, 4, 64);
            if (lane < 2 * kComp * kAB && lane % (2 * kComp) == 0) {
                const int u = lane / (2 * kComp);
                if (u < cnt)
                    jside_store(tot, t1, t2, g_re[j0 + u], g_im[j0 + u], two_eta,
                                j_grad + j0 + u, j_eta + j0 + u);
            }
            wave_handover();  // the next batch overwrites s_jb
        }
    }
}

template <bool JSIDE>
__device__ __forceinline__ void adj_sweep_f64(
    const double *__restrict__ s_src, const double *__restrict__ s_gr,
    const double *__restrict__ s_gi, int n_src, const double (&x)[kJR],
    const double (&gr)[kJR], const double (&gi)[kJR], double e2, double m2e2, double (&P)[kJR],
    double (&A)[kJR], double (&B)[kJR], double *__restrict__ s_jb,
    const double *__restrict__ g_re, const double *__restrict__ g_im, double two_eta,
    double *__restrict__ j_grad, double *__restrict__ j_eta, int lane, int qoff) {
    for (int j0 = 0; j0 < kJQ; j0 += kAB64) {  // same trip count in every wave
        for (int u = 0; u < kAB64; ++u) {
            const int j = j0 + u;
            if (j < n_src) {  // wave-uniform
                const double sv = s_src[j], grj = s_gr[j], gij = s_gi[j];
                double d[kJR], t[kJR], r[kJR];
#pragma unroll
                for (int k = 0; k < kJR; ++k) {
                    d[k] = x[k] - sv;
                    t[k] = __builtin_fma(d[k], d[k], e2);
                }
                rcp_many_f64(t, r);
                double bq = 0.0, ba = 0.0, bb = 0.0;
#pragma unroll
                for (int k = 0; k < kJR; ++k) {
                    const double r2 = r[k] * r[k];
                    double al = __builtin_fma(m2e2, r2, r[k]);  // (d^2 - eta^2) / D^2
                    double be = d[k] * r2;
                    if (!JSIDE) {  // the pair (i, i) adds nothing (as in the FP32 form)
                        const bool self = qoff + k == j;
                        al = self ? 0.0 : al;
                        be = self ? 0.0 : be;
                    }
                    P[k] = __builtin_fma(al, grj, P[k]);
                    P[k] = __builtin_fma(-be, gij, P[k]);
                    A[k] += al;
                    B[k] += be;
                    if (JSIDE) {
                        if (k == 0) {  // the lane sums start from the first pair's term
                            bq = __builtin_fma(be, gi[k], al * gr[k]);
                            ba = al;
                            bb = be;
                        } else {
                            bq = __builtin_fma(al, gr[k], bq);
                            bq = __builtin_fma(be, gi[k], bq);
                            ba += al;
                            bb += be;
                        }
                    }
                }
                if (JSIDE) {
                    s_jb[(kComp * u + 0) * kARow + lane] = bq;
                    s_jb[(kComp * u + 1) * kARow + lane] = ba;
                    s_jb[(kComp * u + 2) * kARow + lane] = bb;
                }
            }
        }
        if (JSIDE) {
            wave_handover();
            // rows (u, comp) = lane >> 2 for lanes < 48, four lanes per row,
            // one chain of 16 each, joined as (c0 + c1) + (c2 + c3)
            const int row_id = lane >> 2, q = lane & 3;
            double cq = 0.0;
            if (row_id < kComp * kAB64) {
                const double *row = s_jb + row_id * kARow + 16 * q;
                cq = row[0];
#pragma unroll
                for (int l = 1; l < 16; ++l) cq += row[l];
            }
            const double o1 = __shfl_xor(cq, 1, 64);
            const double pair = (q & 1) ? o1 + cq : cq + o1;
            const double o2 = __shfl_xor(pair, 2, 64);
            const double tot = (q & 2) ? o2 + pair : pair + o2;
            // source u's sums sit on lanes 12u, 12u + 4, 12u + 8
            const double t1 = __shfl_down(tot, 4, 64), t2 = __shfl_down(tot, 8, 64);
            if (lane < 4 * kComp * kAB64 && lane % (4 * kComp) == 0) {
                const int j = j0 + lane / (4 * kComp);
                if (j < n_src)
                    jside_store(tot, t1, t2, g_re[j], g_im[j], two_eta, j_grad + j, j_eta + j);
            }
            wave_handover();  // the next batch overwrites s_jb
        }
    }
}

// i-side finish: waves 1..3 hand their sums to wave 0 in turn (fixed order,
// through `buf`, 3 * kJT doubles); wave 0 writes tile a's slot b (i_grad /
// i_eta point at the tile's first point).  self_alpha
// is the pair (i, i)'s alpha, -1/eta^2, in a diagonal tile (0 otherwise): it
// was left out of the sums, and its eta term Gi_i alpha belongs to dL/deta.
__device__ __forceinline__ void adj_finish(double (&P)[kJR], double (&A)[kJR], double (&B)[kJR],
                                           double *__restrict__ buf, int w, int lane,
                                           const double *__restrict__ g_re,
                                           const double *__restrict__ g_im, double two_eta,
                                           double *__restrict__ i_grad,
                                           double *__restrict__ i_eta, int p, int q0,
                                           double self_alpha) {
    for (int r = 1; r < kJW; ++r) {
        __syncthreads();
        if (w == r) {
#pragma unroll
            for (int k = 0; k < kJR; ++k) {
                buf[lane * kJR + k] = P[k];
                buf[kJT + lane * kJR + k] = A[k];
                buf[2 * kJT + lane * kJR + k] = B[k];
            }
        }
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int k = 0; k < kJR; ++k) {
                P[k] += buf[lane * kJR + k];
                A[k] += buf[kJT + lane * kJR + k];
                B[k] += buf[2 * kJT + lane * kJR + k];
            }
        }
    }
    if (w == 0) {
        const int l0 = lane * kJR;  // this lane's first query within the tile
#pragma unroll
        for (int k = 0; k < kJR; ++k) {
            const int i = q0 + k;
            if (i < p) {
                const double g_r = g_re[i], g_i = g_im[i];
                i_grad[l0 + k] = (P[k] - g_r * A[k]) - (two_eta * g_i) * B[k];
                i_eta[l0 + k] = g_i * (A[k] + self_alpha) - (two_eta * g_r) * B[k];
            }
        }
    }
}

// Shared memory of both kernels: staged sources and their coefficients, and
// the j-side rows (reused for the i-side combine).
constexpr int kJbufBytes = kJW * kComp * kAB * kARow * 4;  // FP32 rows (= FP64 rows)
static_assert(kComp * kAB * 4 == kComp * kAB64 * 8, "both forms use the same row bytes");
static_assert(kJbufBytes >= 3 * kJT * 8, "i-side combine fits in the row buffer");
static_assert(kJQ % kAB == 0 && kJQ % kAB64 == 0, "batches tile the wave's sources");

__global__ __launch_bounds__(kJL * kJW) __attribute__((amdgpu_waves_per_eu(3))) void symadj_f32c_kernel(
    const double *__restrict__ ev, const double *__restrict__ g_re,
    const double *__restrict__ g_im, int p, double eta, SymBand bd, int id0) {
    int a, b;  // query tile a (in the band), source tile b >= a
    tile_pair_of((int)blockIdx.x + id0, bd.nt, a, b);
    __shared__ double s_src[kJW][kJQ];
    __shared__ float s_gr[kJW][kJQ], s_gi[kJW][kJQ];
    __shared__ double s_jbuf[kJbufBytes / 8];
    // w is wave-uniform; in a scalar register the source counts and batch
    // tests below stay scalar
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const double two_eta = 2.0 * eta;

    // queries: kJR contiguous per lane (only the last tile, which meets only
    // itself, can be partial), centred per lane as in the forward kernels
    const int q0 = a * kJT + lane * kJR;
    double x[kJR];
    double mn = __builtin_inf(), mx = -__builtin_inf();
#pragma unroll
    for (int k = 0; k < kJR; ++k) {
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
    const int form = f32c_form(wave_max(half_w), eta);
    float qh[kJR], ql[kJR], gr[kJR], gi[kJR];
    double P[kJR], A[kJR], B[kJR];
#pragma unroll
    for (int k = 0; k < kJR; ++k) {
        const int i = q0 + k;
        const double t = x[k] - c;
        qh[k] = (float)t;
        ql[k] = (float)(t - (double)qh[k]);
        gr[k] = (i < p) ? (float)g_re[i] : 0.0f;
        gi[k] = (i < p) ? (float)(two_eta * g_im[i]) : 0.0f;
        P[k] = A[k] = B[k] = 0.0;
    }

    // this wave's sources: tile b, quarter w
    const int js = b * kJT + w * kJQ;
    const int n_src = max(0, min(kJQ, p - js));
    for (int u = lane; u < n_src; u += kJL) {
        s_src[w][u] = ev[js + u];
        s_gr[w][u] = (float)g_re[js + u];
        s_gi[w][u] = (float)(two_eta * g_im[js + u]);
    }
    __syncthreads();

    float *s_jb = reinterpret_cast<float *>(s_jbuf) + w * (kComp * kAB * kARow);
    double *j_grad = bd.slot(0, b, a) + w * kJQ;  // tile b, slot a, this wave's sources
    double *j_eta = bd.slot(1, b, a) + w * kJQ;
    const float eta2f = (float)(eta * eta);
    const float m2eta2f = -2.0f * eta2f;
#define SHK_ADJ_SWEEP(F, J)                                                                     \
    adj_sweep_f32c<F, J>(s_src[w], s_gr[w], s_gi[w], n_src, c, qh, ql, gr, gi, eta2f, m2eta2f,  \
                         P, A, B,                                                               \
                         s_jb, g_re + js, g_im + js, two_eta, j_grad, j_eta, lane,     \
                         q0 - js)
    if (a < b) {
        if (form == 0) SHK_ADJ_SWEEP(0, true);
        else if (form == 1) SHK_ADJ_SWEEP(1, true);
        else SHK_ADJ_SWEEP(2, true);
    } else {
        if (form == 0) SHK_ADJ_SWEEP(0, false);
        else if (form == 1) SHK_ADJ_SWEEP(1, false);
        else SHK_ADJ_SWEEP(2, false);
    }
#undef SHK_ADJ_SWEEP
    adj_finish(P, A, B, s_jbuf, w, lane, g_re, g_im, two_eta, bd.slot(0, a, b),
               bd.slot(1, a, b), p, q0, a == b ? -1.0 / (eta * eta) : 0.0);
}

__global__ __launch_bounds__(kJL * kJW) void symadj_f64_kernel(
    const double *__restrict__ ev, const double *__restrict__ g_re,
    const double *__restrict__ g_im, int p, double eta, SymBand bd, int id0) {
    int a, b;  // query tile a (in the band), source tile b >= a
    tile_pair_of((int)blockIdx.x + id0, bd.nt, a, b);
    __shared__ double s_src[kJW][kJQ];
    __shared__ double s_gr[kJW][kJQ], s_gi[kJW][kJQ];
    __shared__ double s_jbuf[kJbufBytes / 8];
    // w is wave-uniform; in a scalar register the source counts and batch
    // tests below stay scalar
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const double two_eta = 2.0 * eta;
    const int q0 = a * kJT + lane * kJR;
    double x[kJR], gr[kJR], gi[kJR], P[kJR], A[kJR], B[kJR];
#pragma unroll
    for (int k = 0; k < kJR; ++k) {
        const int i = q0 + k;
        x[k] = (i < p) ? ev[i] : 0.0;
        gr[k] = (i < p) ? g_re[i] : 0.0;
        gi[k] = (i < p) ? two_eta * g_im[i] : 0.0;
        P[k] = A[k] = B[k] = 0.0;
    }
    const int js = b * kJT + w * kJQ;
    const int n_src = max(0, min(kJQ, p - js));
    for (int u = lane; u < n_src; u += kJL) {
        s_src[w][u] = ev[js + u];
        s_gr[w][u] = g_re[js + u];
        s_gi[w][u] = two_eta * g_im[js + u];
    }
    __syncthreads();
    double *s_jb = s_jbuf + w * (kComp * kAB64 * kARow);
    double *j_grad = bd.slot(0, b, a) + w * kJQ;
    double *j_eta = bd.slot(1, b, a) + w * kJQ;
    const double e2 = eta * eta, m2e2 = -2.0 * e2;
    if (a < b)
        adj_sweep_f64<true>(s_src[w], s_gr[w], s_gi[w], n_src, x, gr, gi, e2, m2e2, P, A, B, s_jb,
                            g_re + js, g_im + js, two_eta, j_grad, j_eta, lane,
                            q0 - js);
    else
        adj_sweep_f64<false>(s_src[w], s_gr[w], s_gi[w], n_src, x, gr, gi, e2, m2e2, P, A, B, s_jb,
                             g_re + js, g_im + js, two_eta, j_grad, j_eta, lane,
                             q0 - js);
    adj_finish(P, A, B, s_jbuf, w, lane, g_re, g_im, two_eta, bd.slot(0, a, b),
               bd.slot(1, a, b), p, q0, a == b ? -1.0 / (eta * eta) : 0.0);
}

static_assert(kJT == kSymTile, "the band layout's tile");

struct AdjCtx {
    const double *ev, *gre, *gim;
    int p;
    double eta;
    int f64;
    hipStream_t stream;
};

int symadj_launch(const SymBand &bd, int grid, int id0, void *vctx) {
    const AdjCtx &c = *static_cast<const AdjCtx *>(vctx);
    ProfileMark pm = profile_begin(c.stream);
    if (c.f64)
        hipLaunchKernelGGL(symadj_f64_kernel, dim3(grid), dim3(kJL * kJW), 0, c.stream, c.ev,
                           c.gre, c.gim, c.p, c.eta, bd, id0);
    else
        hipLaunchKernelGGL(symadj_f32c_kernel, dim3(grid), dim3(kJL * kJW), 0, c.stream, c.ev,
                           c.gre, c.gim, c.p, c.eta, bd, id0);
    profile_end(pm, c.stream);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        set_error(std::string("symmetric adjoint launch: ") + hipGetErrorString(le));
        return 7;
    }
    return 0;
}

}  // namespace

int exact_adjoint_sym(const double *ev, size_t p, double eta, const double *gre,
                      const double *gim, int f64, double *grad, double *grad_eta,
                      hipStream_t stream, int accumulate_eta) {
    if (p == 0) return 0;
    if (p > (size_t)kSymMaxP) {
        set_error("symmetric adjoint: p above its maximum");
        return 6;
    }
    // partials (banded within the budget) and the per-end eta terms
    const int R = std::max(1, sym_band_rows(p, sym_scratch_budget() / sizeof(double)));
    const size_t need = sym_scratch_doubles(p, R);
    double *part = nullptr;
    CachedBuf cache(CACHE_ADJ_PART, stream);
    if (int e = cache.get(need + p, &part)) return e;
    double *eta_term = part + need;
    AdjCtx ctx{ev, gre, gim, (int)p, eta, f64, stream};
    // in-order sums over the nt slots: gradient and per-end eta terms
    if (int st = sym_run_bands(p, part, need, 1.0, 1.0, grad, eta_term, symadj_launch, &ctx,
                               stream))
        return st;
    if (grad_eta) return adjoint_ordered_sum(eta_term, p, grad_eta, accumulate_eta, stream);
    return 0;
}

}  // namespace shk
