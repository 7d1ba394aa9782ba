// Reverse-mode kernels for the exact at-points sums and the El Karoui fit.
//
// The reference differentiates the transform only with respect to the query
// point (stieltjes_with_deriv_sum, src/stieltjes/autovec.rs:38-59); these
// kernels supply the rest of the chain for a real loss of the raw sums
// S_m = sum_j 1/((x_m - lam_j) - i*eta) with upstream gradients
// (G_re[m], G_im[m]) = (dL/dRe S_m, dL/dIm S_m).  With d = x_m - lam_j and
// D = d^2 + eta^2:
//
//   grad_lam[j] = sum_m (G_re (d^2 - eta^2) + 2 G_im eta d) / D^2        (A)
//   dS_m/dx     = sum_j ((eta^2 - d^2) - i 2 eta d) / D^2                 (B)
//   grad_x[m]   = G_re Re dS/dx + G_im Im dS/dx
//   grad_eta    = sum_m G_re Im dS/dx - G_im Re dS/dx     (dS/deta = -i dS/dx)
//
// (A) is the transpose sweep: a workgroup owns a tile of eigenvalues and
// sweeps the query points (with their upstream gradients) staged in LDS.  (B)
// is the forward sweep with two derivative accumulators.  Both use the tile
// layout of the exact kernels (one wavefront per workgroup, two outputs per
// lane, query/source splits to fill the chip, a deterministic FP64 split
// reduction) and come in FP64 and FP32-compensated form; the compensated
// form re-expresses both operands relative to the owned tile's centre in f32
// hi/lo parts exactly as the forward kernel does, sums runs of 32 terms in
// f32 and folds them into FP64.  Each term is evaluated as r * (r * poly)
// with r = 1/D, so no intermediate exceeds the range of the term itself.
//
// The El Karoui fit (no reference counterpart; the reference inverts the MP
// equation pointwise, src/deconvolution/mod.rs:116-259) is O(M K) FP64 work:
// one wavefront per grid point forms a = 1 - c z g, the residual
// r = g - sum_k w_k / (t_k a - z) and its gradients; one wavefront per atom
// forms dL/dw_k; a one-workgroup tree sum gives the loss and the direct eta
// term.  All sums run in a fixed order, so results are bitwise reproducible.
#include "shk_internal.hpp"
#include "shk_pair.hpp"

#include <algorithm>

namespace shk {
namespace {

constexpr int kAWG = 64;           // one wavefront per workgroup
constexpr int kAR = 2;             // owned outputs per lane
constexpr int kAT = kAWG * kAR;    // owned outputs per tile
constexpr int kAPiece = 512;       // swept values staged in LDS per pass
constexpr int kATargetWG = 16384;  // workgroups to aim for
constexpr int kSumThreads = 256;

void plan(size_t n_own, size_t n_sweep, int *tiles, int *n_split, int *per) {
    const int qt = (int)((n_own + kAT - 1) / kAT);
    int ns = std::max(1, (kATargetWG + qt - 1) / qt);
    const int max_split = (int)std::max<size_t>(1, (n_sweep + 63) / 64);
    ns = std::min(ns, max_split);
    int pp = (int)((n_sweep + ns - 1) / ns);
    pp = (pp + 31) / 32 * 32;
    ns = (int)((n_sweep + pp - 1) / pp);
    *tiles = qt;
    *n_split = std::max(ns, 1);
    *per = pp;
}

// Tile centre and pair-difference form from the owned values (one wavefront).
__device__ __forceinline__ void tile_centre(const double (&v)[kAR], const bool (&ok)[kAR],
                                            double eta, double *c, int *form) {
    double mn = __builtin_inf(), mx = -__builtin_inf();
#pragma unroll
    for (int k = 0; k < kAR; ++k)
        if (ok[k]) {
            mn = fmin(mn, v[k]);
            mx = fmax(mx, v[k]);
        }
    mn = wave_min(mn);
    mx = wave_max(mx);
    *c = 0.5 * (mn + mx);
    *form = f32c_form(0.5 * (mx - mn), eta);
}

// ---- (A) gradient with respect to the eigenvalues ---------------------------

// Staged per query m: x_m relative to the tile centre (hi, lo) and the term
// coefficients ga = G_re, gb = -2 eta G_im, gc = -eta^2 G_re, so that with
// dd = lam - x the term is (ga dd^2 + gb dd + gc) / D^2.
template <int FORM>
__device__ __forceinline__ void src_sweep_f32c(const float *__restrict__ xh,
                                               const float *__restrict__ xl,
                                               const float *__restrict__ ga,
                                               const float *__restrict__ gb,
                                               const float *__restrict__ gc, int cnt,
                                               const float (&lh)[kAR], const float (&ll)[kAR],
                                               float eta2f, double (&acc)[kAR]) {
    int j = 0;
    while (j < cnt) {
        const int run = min(kF32Run, cnt - j);
        float a[kAR];
#pragma unroll
        for (int k = 0; k < kAR; ++k) a[k] = 0.0f;
        if (run == kF32Run) {
#pragma unroll 8
            for (int u = 0; u < kF32Run; ++u) {
                const float sh = xh[j + u], sl = (FORM == 2) ? xl[j + u] : 0.0f;
                const float A = ga[j + u], B = gb[j + u], C = gc[j + u];
#pragma unroll
                for (int k = 0; k < kAR; ++k) {
                    const float dd = pair_diff<FORM>(lh[k], ll[k], sh, sl);
                    const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(dd, dd, eta2f));
                    const float poly = __builtin_fmaf(__builtin_fmaf(A, dd, B), dd, C);
                    a[k] = __builtin_fmaf(r, r * poly, a[k]);
                }
            }
        } else {
            for (int u = 0; u < run; ++u) {
                const float sh = xh[j + u], sl = (FORM == 2) ? xl[j + u] : 0.0f;
                const float A = ga[j + u], B = gb[j + u], C = gc[j + u];
#pragma unroll
                for (int k = 0; k < kAR; ++k) {
                    const float dd = pair_diff<FORM>(lh[k], ll[k], sh, sl);
                    const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(dd, dd, eta2f));
                    const float poly = __builtin_fmaf(__builtin_fmaf(A, dd, B), dd, C);
                    a[k] = __builtin_fmaf(r, r * poly, a[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kAR; ++k) acc[k] += (double)a[k];
        j += run;
    }
}

__global__ __launch_bounds__(kAWG) void adj_src_f32c_kernel(
    const double *__restrict__ q, int nq, const double *__restrict__ gre,
    const double *__restrict__ gim, const double *__restrict__ ev, int p, double eta, int per,
    double *__restrict__ part) {
    __shared__ float s_xh[kAPiece], s_xl[kAPiece], s_a[kAPiece], s_b[kAPiece], s_c[kAPiece];
    const int lane = threadIdx.x;
    const int split = blockIdx.y;
    const int m_begin = split * per;
    const int m_end = min(m_begin + per, nq);
    double lam[kAR];
    bool ok[kAR];
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int j = blockIdx.x * kAT + k * kAWG + lane;
        ok[k] = j < p;
        lam[k] = ok[k] ? ev[j] : 0.0;
    }
    double c;
    int form;
    tile_centre(lam, ok, eta, &c, &form);
    float lh[kAR], ll[kAR];
    double acc[kAR];
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const double t = lam[k] - c;
        lh[k] = (float)t;
        ll[k] = (float)(t - (double)lh[k]);
        acc[k] = 0.0;
    }
    const float eta2f = (float)(eta * eta);
    for (int base = m_begin; base < m_end; base += kAPiece) {
        const int cnt = min(kAPiece, m_end - base);
        __syncthreads();
        for (int u = lane; u < cnt; u += kAWG) {
            const double t = q[base + u] - c;
            const float h = (float)t;
            s_xh[u] = h;
            s_xl[u] = (float)(t - (double)h);
            const double a = gre[base + u], b = gim[base + u];
            s_a[u] = (float)a;
            s_b[u] = (float)(-2.0 * eta * b);
            s_c[u] = (float)(-(eta * eta) * a);
        }
        __syncthreads();
        if (form == 0)
            src_sweep_f32c<0>(s_xh, s_xl, s_a, s_b, s_c, cnt, lh, ll, eta2f, acc);
        else if (form == 1)
            src_sweep_f32c<1>(s_xh, s_xl, s_a, s_b, s_c, cnt, lh, ll, eta2f, acc);
        else
            src_sweep_f32c<2>(s_xh, s_xl, s_a, s_b, s_c, cnt, lh, ll, eta2f, acc);
    }
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int j = blockIdx.x * kAT + k * kAWG + lane;
        if (j < p) part[(size_t)split * p + j] = acc[k];
    }
}

__global__ __launch_bounds__(kAWG) void adj_src_f64_kernel(
    const double *__restrict__ q, int nq, const double *__restrict__ gre,
    const double *__restrict__ gim, const double *__restrict__ ev, int p, double eta, int per,
    double *__restrict__ part) {
    __shared__ double s_x[kAPiece], s_a[kAPiece], s_b[kAPiece], s_c[kAPiece];
    const int lane = threadIdx.x;
    const int split = blockIdx.y;
    const int m_begin = split * per;
    const int m_end = min(m_begin + per, nq);
    const double e2 = eta * eta;
    double lam[kAR], acc[kAR];
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int j = blockIdx.x * kAT + k * kAWG + lane;
        lam[k] = (j < p) ? ev[j] : 0.0;
        acc[k] = 0.0;
    }
    for (int base = m_begin; base < m_end; base += kAPiece) {
        const int cnt = min(kAPiece, m_end - base);
        __syncthreads();
        for (int u = lane; u < cnt; u += kAWG) {
            const double a = gre[base + u], b = gim[base + u];
            s_x[u] = q[base + u];
            s_a[u] = a;
            s_b[u] = -2.0 * eta * b;
            s_c[u] = -e2 * a;
        }
        __syncthreads();
        // two query points per step; their 2 kAR reciprocals come four per
        // v_rcp_f64 (rcp_many_f64)
        auto step = [&](int u, const double (&dd)[kAR], const double (&rr)[kAR]) {
            const double A = s_a[u], B = s_b[u], C = s_c[u];
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                const double poly = __builtin_fma(__builtin_fma(A, dd[k], B), dd[k], C);
                acc[k] = __builtin_fma(rr[k], rr[k] * poly, acc[k]);
            }
        };
        int u = 0;
        for (; u + 2 <= cnt; u += 2) {
            const double x0 = s_x[u], x1 = s_x[u + 1];
            double dd[2 * kAR], tt[2 * kAR], rr[2 * kAR];
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                dd[k] = lam[k] - x0;
                dd[kAR + k] = lam[k] - x1;
                tt[k] = __builtin_fma(dd[k], dd[k], e2);
                tt[kAR + k] = __builtin_fma(dd[kAR + k], dd[kAR + k], e2);
            }
            rcp_many_f64(tt, rr);
            double d0[kAR], r0[kAR], d1[kAR], r1[kAR];
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                d0[k] = dd[k];
                r0[k] = rr[k];
                d1[k] = dd[kAR + k];
                r1[k] = rr[kAR + k];
            }
            step(u, d0, r0);
            step(u + 1, d1, r1);
        }
        if (u < cnt) {
            double dd[kAR], tt[kAR], rr[kAR];
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                dd[k] = lam[k] - s_x[u];
                tt[k] = __builtin_fma(dd[k], dd[k], e2);
            }
            rcp_many_f64(tt, rr);
            step(u, dd, rr);
        }
    }
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int j = blockIdx.x * kAT + k * kAWG + lane;
        if (j < p) part[(size_t)split * p + j] = acc[k];
    }
}

// Split reductions: one lane per output, one wavefront per workgroup (so
// they spread over every CU), loads kRedAhead splits ahead of the additions,
// which stay in split order.
constexpr int kRedWG = 64;
constexpr int kRedAhead = 8;

__global__ __launch_bounds__(kRedWG) void adj_src_reduce_kernel(const double *__restrict__ part,
                                                                int n_split, int p,
                                                                double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p) return;
    double a = 0.0;
    int s = 0;
    for (; s + kRedAhead <= n_split; s += kRedAhead) {
        double v[kRedAhead];
#pragma unroll
        for (int u = 0; u < kRedAhead; ++u) v[u] = part[(size_t)(s + u) * p + j];
#pragma unroll
        for (int u = 0; u < kRedAhead; ++u) a += v[u];
    }
    for (; s < n_split; ++s) a += part[(size_t)s * p + j];
    out[j] = a;
}

// ---- (B) derivative sums at the query points ---------------------------------

// Per query: P = sum (eta^2 - d^2)/D^2 and Q = sum d/D^2; dS/dx = P - 2 i eta Q.
template <int FORM>
__device__ __forceinline__ void qry_sweep_f32c(const float *__restrict__ sh_lds,
                                               const float *__restrict__ sl_lds, int cnt,
                                               const float (&qh)[kAR], const float (&ql)[kAR],
                                               float eta2f, double (&P)[kAR], double (&Q)[kAR]) {
    int j = 0;
    while (j < cnt) {
        const int run = min(kF32Run, cnt - j);
        float a[kAR], b[kAR];
#pragma unroll
        for (int k = 0; k < kAR; ++k) a[k] = b[k] = 0.0f;
        if (run == kF32Run) {
#pragma unroll 8
            for (int u = 0; u < kF32Run; ++u) {
                const float sh = sh_lds[j + u], sl = (FORM == 2) ? sl_lds[j + u] : 0.0f;
#pragma unroll
                for (int k = 0; k < kAR; ++k) {
                    const float d = pair_diff<FORM>(qh[k], ql[k], sh, sl);
                    const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2f));
                    const float e = __builtin_fmaf(-d, d, eta2f);
                    a[k] = __builtin_fmaf(r, r * e, a[k]);
                    b[k] = __builtin_fmaf(r, r * d, b[k]);
                }
            }
        } else {
            for (int u = 0; u < run; ++u) {
                const float sh = sh_lds[j + u], sl = (FORM == 2) ? sl_lds[j + u] : 0.0f;
#pragma unroll
                for (int k = 0; k < kAR; ++k) {
                    const float d = pair_diff<FORM>(qh[k], ql[k], sh, sl);
                    const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2f));
                    const float e = __builtin_fmaf(-d, d, eta2f);
                    a[k] = __builtin_fmaf(r, r * e, a[k]);
                    b[k] = __builtin_fmaf(r, r * d, b[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kAR; ++k) {
            P[k] += (double)a[k];
            Q[k] += (double)b[k];
        }
        j += run;
    }
}

__global__ __launch_bounds__(kAWG) void adj_qry_f32c_kernel(const double *__restrict__ q, int nq,
                                                            const double *__restrict__ ev, int p,
                                                            double eta, int per,
                                                            double *__restrict__ part) {
    __shared__ float s_hi[kAPiece], s_lo[kAPiece];
    const int lane = threadIdx.x;
    const int split = blockIdx.y;
    const int n_split = gridDim.y;
    const int j_begin = split * per;
    const int j_end = min(j_begin + per, p);
    double x[kAR];
    bool ok[kAR];
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int i = blockIdx.x * kAT + k * kAWG + lane;
        ok[k] = i < nq;
        x[k] = ok[k] ? q[i] : 0.0;
    }
    double c;
    int form;
    tile_centre(x, ok, eta, &c, &form);
    float qh[kAR], ql[kAR];
    double P[kAR], Q[kAR];
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const double t = x[k] - c;
        qh[k] = (float)t;
        ql[k] = (float)(t - (double)qh[k]);
        P[k] = Q[k] = 0.0;
    }
    const float eta2f = (float)(eta * eta);
    for (int base = j_begin; base < j_end; base += kAPiece) {
        const int cnt = min(kAPiece, j_end - base);
        __syncthreads();
        for (int u = lane; u < cnt; u += kAWG) {
            const double t = ev[base + u] - c;
            const float h = (float)t;
            s_hi[u] = h;
            s_lo[u] = (float)(t - (double)h);
        }
        __syncthreads();
        if (form == 0)
            qry_sweep_f32c<0>(s_hi, s_lo, cnt, qh, ql, eta2f, P, Q);
        else if (form == 1)
            qry_sweep_f32c<1>(s_hi, s_lo, cnt, qh, ql, eta2f, P, Q);
        else
            qry_sweep_f32c<2>(s_hi, s_lo, cnt, qh, ql, eta2f, P, Q);
    }
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int i = blockIdx.x * kAT + k * kAWG + lane;
        if (i < nq) {
            part[((size_t)0 * n_split + split) * nq + i] = P[k];
            part[((size_t)1 * n_split + split) * nq + i] = Q[k];
        }
    }
}

__global__ __launch_bounds__(kAWG) void adj_qry_f64_kernel(const double *__restrict__ q, int nq,
                                                           const double *__restrict__ ev, int p,
                                                           double eta, int per,
                                                           double *__restrict__ part) {
    __shared__ double s_src[kAPiece];
    const int lane = threadIdx.x;
    const int split = blockIdx.y;
    const int n_split = gridDim.y;
    const int j_begin = split * per;
    const int j_end = min(j_begin + per, p);
    const double e2 = eta * eta;
    double x[kAR], P[kAR], Q[kAR];
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int i = blockIdx.x * kAT + k * kAWG + lane;
        x[k] = (i < nq) ? q[i] : 0.0;
        P[k] = Q[k] = 0.0;
    }
    for (int base = j_begin; base < j_end; base += kAPiece) {
        const int cnt = min(kAPiece, j_end - base);
        __syncthreads();
        for (int u = lane; u < cnt; u += kAWG) s_src[u] = ev[base + u];
        __syncthreads();
        // two sources per step; their 2 kAR reciprocals come four per
        // v_rcp_f64 (rcp_many_f64)
        auto step = [&](const double (&dd)[kAR], const double (&rr)[kAR]) {
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                const double d = dd[k], r = rr[k];
                const double e = __builtin_fma(-d, d, e2);
                P[k] = __builtin_fma(r, r * e, P[k]);
                Q[k] = __builtin_fma(r, r * d, Q[k]);
            }
        };
        int u = 0;
        for (; u + 2 <= cnt; u += 2) {
            const double s0 = s_src[u], s1 = s_src[u + 1];
            double dd[2 * kAR], tt[2 * kAR], rr[2 * kAR];
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                dd[k] = x[k] - s0;
                dd[kAR + k] = x[k] - s1;
                tt[k] = __builtin_fma(dd[k], dd[k], e2);
                tt[kAR + k] = __builtin_fma(dd[kAR + k], dd[kAR + k], e2);
            }
            rcp_many_f64(tt, rr);
            double d0[kAR], r0[kAR], d1[kAR], r1[kAR];
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                d0[k] = dd[k];
                r0[k] = rr[k];
                d1[k] = dd[kAR + k];
                r1[k] = rr[kAR + k];
            }
            step(d0, r0);
            step(d1, r1);
        }
        if (u < cnt) {
            double dd[kAR], tt[kAR], rr[kAR];
#pragma unroll
            for (int k = 0; k < kAR; ++k) {
                dd[k] = x[k] - s_src[u];
                tt[k] = __builtin_fma(dd[k], dd[k], e2);
            }
            rcp_many_f64(tt, rr);
            step(dd, rr);
        }
    }
#pragma unroll
    for (int k = 0; k < kAR; ++k) {
        const int i = blockIdx.x * kAT + k * kAWG + lane;
        if (i < nq) {
            part[((size_t)0 * n_split + split) * nq + i] = P[k];
            part[((size_t)1 * n_split + split) * nq + i] = Q[k];
        }
    }
}

// The query sweep at a few points against many sources (grid_queries): one
// workgroup per point, thread t sums the sources t, t + kGridThreads, ...,
// eight per step with their loads issued together (FP32-compensated: the
// FP64 difference rounded to f32, each step's FP32 sums folded into FP64),
// the thread sums are added in a fixed tree, and grad_x and the eta term
// follow as in adj_qry_reduce_kernel.  No partial sums.
constexpr int kGridStep = 8;

template <bool F32C>
__global__ __launch_bounds__(kGridThreads) void adj_grid_qry_kernel(
    const double *__restrict__ q, const double *__restrict__ ev, int p, double eta,
    const double *__restrict__ gre, const double *__restrict__ gim, double *__restrict__ grad_q,
    double *__restrict__ eta_term) {
    constexpr int U = kGridStep;
    __shared__ double s_p[kGridThreads], s_q[kGridThreads];
    const int t = threadIdx.x, i = blockIdx.x;
    const double x = q[i];
    const double e2 = eta * eta;
    const float e2f = (float)e2;
    double P = 0.0, Q = 0.0;
    int j = t;
    for (; j + (U - 1) * kGridThreads < p; j += U * kGridThreads) {
        double d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] = x - ev[j + u * kGridThreads];
        if constexpr (F32C) {
            float a = 0.f, b = 0.f;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float df = (float)d[u];
                const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(df, df, e2f));
                const float e = __builtin_fmaf(-df, df, e2f);
                a = __builtin_fmaf(r, r * e, a);
                b = __builtin_fmaf(r, r * df, b);
            }
            P += (double)a;
            Q += (double)b;
        } else {
            double tt[U], rr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) tt[u] = __builtin_fma(d[u], d[u], e2);
            rcp_many_f64<U>(tt, rr);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double e = __builtin_fma(-d[u], d[u], e2);
                P = __builtin_fma(rr[u], rr[u] * e, P);
                Q = __builtin_fma(rr[u], rr[u] * d[u], Q);
            }
        }
    }
    for (; j < p; j += kGridThreads) {  // the last few sources, one at a time
        const double d = x - ev[j];
        if constexpr (F32C) {
            const float df = (float)d;
            const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(df, df, e2f));
            const float e = __builtin_fmaf(-df, df, e2f);
            P += (double)(r * (r * e));
            Q += (double)(r * (r * df));
        } else {
            const double r = rcp_f64(__builtin_fma(d, d, e2));
            P = __builtin_fma(r, r * __builtin_fma(-d, d, e2), P);
            Q = __builtin_fma(r, r * d, Q);
        }
    }
    s_p[t] = P;
    s_q[t] = Q;
    __syncthreads();
#pragma unroll
    for (int w = kGridThreads / 2; w > 0; w >>= 1) {
        if (t < w) {
            s_p[t] += s_p[t + w];
            s_q[t] += s_q[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double dre = s_p[0], dim = -2.0 * eta * s_q[0];
        const double a = gre[i], b = gim[i];
        if (grad_q) grad_q[i] = a * dre + b * dim;
        if (eta_term) eta_term[i] = a * dim - b * dre;
    }
}

// Split reduction of (P, Q), then grad_x and the per-query eta term.
__global__ __launch_bounds__(kRedWG) void adj_qry_reduce_kernel(
    const double *__restrict__ part, int n_split, int nq, double eta,
    const double *__restrict__ gre, const double *__restrict__ gim,
    double *__restrict__ grad_q, double *__restrict__ eta_term) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    double P = 0.0, Q = 0.0;
    int s = 0;
    for (; s + kRedAhead <= n_split; s += kRedAhead) {
        double vp[kRedAhead], vq[kRedAhead];
#pragma unroll
        for (int u = 0; u < kRedAhead; ++u) {
            vp[u] = part[((size_t)0 * n_split + s + u) * nq + i];
            vq[u] = part[((size_t)1 * n_split + s + u) * nq + i];
        }
#pragma unroll
        for (int u = 0; u < kRedAhead; ++u) {
            P += vp[u];
            Q += vq[u];
        }
    }
    for (; s < n_split; ++s) {
        P += part[((size_t)0 * n_split + s) * nq + i];
        Q += part[((size_t)1 * n_split + s) * nq + i];
    }
    const double dre = P, dim = -2.0 * eta * Q;
    const double a = gre[i], b = gim[i];
    if (grad_q) grad_q[i] = a * dre + b * dim;
    if (eta_term) eta_term[i] = a * dim - b * dre;
}

// One workgroup: out = scale * sum(v) in a fixed order (strided per-thread
// sums, then a tree); with accumulate != 0 the result is added to *out.
__global__ __launch_bounds__(kSumThreads) void ordered_sum_kernel(const double *__restrict__ v,
                                                                  int n, double scale,
                                                                  double *__restrict__ out,
                                                                  int accumulate) {
    __shared__ double s[kSumThreads];
    const int tid = threadIdx.x;
    double a = 0.0;
    for (int i = tid; i < n; i += kSumThreads) a += v[i];
    s[tid] = a;
    __syncthreads();
    for (int o = kSumThreads / 2; o > 0; o >>= 1) {
        if (tid < o) s[tid] += s[tid + o];
        __syncthreads();
    }
    if (tid == 0) *out = accumulate ? *out + s[0] * scale : s[0] * scale;
}

// ---- El Karoui fit ---------------------------------------------------------

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx crecip(cplx a) {
    const double inv = 1.0 / (a.re * a.re + a.im * a.im);
    return {a.re * inv, -a.im * inv};
}

// Fixed-order sum over the 64 lanes of a wavefront (butterfly; every lane
// ends with the total, and the order is the same on every run).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One wavefront per grid point; lane l takes atoms l, l+64, ...  The input is
// (in_re * g_re, in_im * g_im) and the g-gradient is written scaled by
// (out_re, out_im): 1s for g itself, or (-1/p, 1/p) both ways for the raw
// sums S (g = -conj(S)/p).
// a = 1 - c - c z g, h_k = 1/(t_k a - z), r = g - sum w_k h_k,
// dr/dg = 1 - c z sum w_k t_k h_k^2, dr/dz = -sum w_k h_k^2 (1 + c t_k g).
__global__ __launch_bounds__(64) void ek_point_kernel(
    const double *__restrict__ x, const double *__restrict__ g_re, const double *__restrict__ g_im,
    int m, double eta, double c, const double *__restrict__ t, const double *__restrict__ w, int k,
    double *__restrict__ a_out, double *__restrict__ r_out, double *__restrict__ sq_out,
    double *__restrict__ eta_out, double *__restrict__ grad_g_re,
    double *__restrict__ grad_g_im, double in_re, double in_im, double out_re, double out_im) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const cplx z = {x[i], eta};
    const cplx g = {in_re * g_re[i], in_im * g_im[i]};
    const cplx zg = cmul(z, g);
    const cplx a = {1.0 - c - c * zg.re, -c * zg.im};
    cplx s = {0.0, 0.0}, sd = {0.0, 0.0}, sz = {0.0, 0.0};
    for (int q = lane; q < k; q += 64) {
        const double tq = t[q], wq = w[q];
        const cplx h = crecip({tq * a.re - z.re, tq * a.im - z.im});
        const cplx h2 = cmul(h, h);
        s.re += wq * h.re;
        s.im += wq * h.im;
        sd.re += wq * tq * h2.re;
        sd.im += wq * tq * h2.im;
        const cplx f = cmul(h2, {1.0 + c * tq * g.re, c * tq * g.im});
        sz.re += wq * f.re;
        sz.im += wq * f.im;
    }
    s = {wave_sum(s.re), wave_sum(s.im)};
    sd = {wave_sum(sd.re), wave_sum(sd.im)};
    sz = {wave_sum(sz.re), wave_sum(sz.im)};
    if (lane != 0) return;
    const cplx r = {g.re - s.re, g.im - s.im};
    const cplx zsd = cmul(z, sd);
    const cplx jg = {1.0 - c * zsd.re, -c * zsd.im};
    const double two_m = 2.0 / (double)m;
    // (2/M) r conj(jg)
    if (grad_g_re) grad_g_re[i] = out_re * (two_m * (r.re * jg.re + r.im * jg.im));
    if (grad_g_im) grad_g_im[i] = out_im * (two_m * (r.im * jg.re - r.re * jg.im));
    // (2/M) Re(conj(r) i (-sz)) = (2/M) Im(conj(r) sz)
    eta_out[i] = two_m * (r.re * sz.im - r.im * sz.re);
    sq_out[i] = r.re * r.re + r.im * r.im;
    a_out[2 * i] = a.re;
    a_out[2 * i + 1] = a.im;
    r_out[2 * i] = r.re;
    r_out[2 * i + 1] = r.im;
}

// One wavefront per atom; lane l takes grid points l, l+64, ...:
// dL/dw_k = -(2/M) sum_m Re(conj(r_m) h_mk).
__global__ __launch_bounds__(64) void ek_weight_kernel(const double *__restrict__ x, int m,
                                                       double eta, const double *__restrict__ t,
                                                       const double *__restrict__ a_in,
                                                       const double *__restrict__ r_in,
                                                       double *__restrict__ grad_w) {
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const double tq = t[q];
    double acc = 0.0;
    for (int i = lane; i < m; i += 64) {
        const cplx h = crecip({tq * a_in[2 * i] - x[i], tq * a_in[2 * i + 1] - eta});
        acc += r_in[2 * i] * h.re + r_in[2 * i + 1] * h.im;
    }
    acc = wave_sum(acc);
    if (lane == 0) grad_w[q] = -(2.0 / (double)m) * acc;
}

__global__ void zero_kernel(double *__restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void add_kernel(double *__restrict__ out, const double *__restrict__ in, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += in[i];
}

}  // namespace

int add_into(double *out, const double *in, size_t n, hipStream_t stream) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, out,
                       in, n);
    SHK_HIP(hipGetLastError());
    return 0;
}

int adjoint_ordered_sum(const double *v, size_t n, double *out, int accumulate,
                        hipStream_t stream) {
    if (n > (size_t)INT32_MAX) {
        set_error("ordered sum: n above 2^31-1");
        return 6;
    }
    hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(kSumThreads), 0, stream, v, (int)n, 1.0,
                       out, accumulate);
    SHK_HIP(hipGetLastError());
    return 0;
}

int zero_fill(double *p, size_t n, hipStream_t stream) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
    SHK_HIP(hipGetLastError());
    return 0;
}

int exact_adjoint(const double *q, size_t nq, const double *ev, size_t p, double eta,
                  const double *gre, const double *gim, int f32c, double *grad_ev,
                  double *grad_q, double *grad_eta, hipStream_t stream, int accumulate_eta) {
    if (nq == 0 || p == 0) return 0;
    if (nq > (size_t)INT32_MAX || p > (size_t)INT32_MAX) {
        set_error("adjoint kernel: sizes above 2^31-1 are not supported");
        return 6;
    }
    if (grad_ev) {
        int tiles, ns, per;
        plan(p, nq, &tiles, &ns, &per);
        double *part = nullptr;
        CachedBuf cache(CACHE_ADJ_PART, stream);
        if (int e = cache.get((size_t)ns * p, &part)) return e;
        ProfileMark mk = profile_begin(stream);
        if (f32c)
            hipLaunchKernelGGL(adj_src_f32c_kernel, dim3(tiles, ns), dim3(kAWG), 0, stream, q,
                               (int)nq, gre, gim, ev, (int)p, eta, per, part);
        else
            hipLaunchKernelGGL(adj_src_f64_kernel, dim3(tiles, ns), dim3(kAWG), 0, stream, q,
                               (int)nq, gre, gim, ev, (int)p, eta, per, part);
        profile_end(mk, stream);
        hipError_t le = hipGetLastError();
        if (le == hipSuccess) {
            hipLaunchKernelGGL(adj_src_reduce_kernel, dim3((unsigned)((p + kRedWG - 1) / kRedWG)),
                               dim3(kRedWG), 0, stream, part, ns, (int)p, grad_ev);
            le = hipGetLastError();
        }
        if (le != hipSuccess) {
            set_error(std::string("adjoint (eigenvalue sweep): ") + hipGetErrorString(le));
            return 7;
        }
    }
    if ((grad_q || grad_eta) && grid_queries(nq, p)) {
        double *eta_term = nullptr;
        CachedBuf cache(CACHE_ADJ_PART, stream);
        if (grad_eta)
            if (int e = cache.get(nq, &eta_term)) return e;
        if (f32c)
            hipLaunchKernelGGL(adj_grid_qry_kernel<true>, dim3((unsigned)nq), dim3(kGridThreads),
                               0, stream, q, ev, (int)p, eta, gre, gim, grad_q, eta_term);
        else
            hipLaunchKernelGGL(adj_grid_qry_kernel<false>, dim3((unsigned)nq), dim3(kGridThreads),
                               0, stream, q, ev, (int)p, eta, gre, gim, grad_q, eta_term);
        hipError_t le = hipGetLastError();
        if (le == hipSuccess && grad_eta) {
            hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(kSumThreads), 0, stream,
                               eta_term, (int)nq, 1.0, grad_eta, accumulate_eta);
            le = hipGetLastError();
        }
        if (le != hipSuccess) {
            set_error(std::string("adjoint (query sweep at grid points): ") +
                      hipGetErrorString(le));
            return 7;
        }
    } else if (grad_q || grad_eta) {
        int tiles, ns, per;
        plan(nq, p, &tiles, &ns, &per);
        double *part = nullptr;
        CachedBuf cache(CACHE_ADJ_PART, stream);
        if (int e = cache.get(2 * (size_t)ns * nq + nq, &part)) return e;
        double *eta_term = part + 2 * (size_t)ns * nq;
        ProfileMark mk = profile_begin(stream);
        if (f32c)
            hipLaunchKernelGGL(adj_qry_f32c_kernel, dim3(tiles, ns), dim3(kAWG), 0, stream, q,
                               (int)nq, ev, (int)p, eta, per, part);
        else
            hipLaunchKernelGGL(adj_qry_f64_kernel, dim3(tiles, ns), dim3(kAWG), 0, stream, q,
                               (int)nq, ev, (int)p, eta, per, part);
        profile_end(mk, stream);
        hipError_t le = hipGetLastError();
        if (le == hipSuccess) {
            hipLaunchKernelGGL(adj_qry_reduce_kernel, dim3((unsigned)((nq + kRedWG - 1) / kRedWG)),
                               dim3(kRedWG), 0, stream, part, ns, (int)nq, eta, gre, gim, grad_q,
                               eta_term);
            le = hipGetLastError();
        }
        if (le == hipSuccess && grad_eta) {
            hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(kSumThreads), 0, stream,
                               eta_term, (int)nq, 1.0, grad_eta, accumulate_eta);
            le = hipGetLastError();
        }
        if (le != hipSuccess) {
            set_error(std::string("adjoint (query sweep): ") + hipGetErrorString(le));
            return 7;
        }
    }
    return 0;
}

namespace {

int ek_fit_impl(const double *x, const double *g_re, const double *g_im, size_t m, double eta,
                double c, const double *t, const double *w, size_t k, double *loss,
                double *grad_g_re, double *grad_g_im, double *grad_w, double *grad_eta,
                double in_re, double in_im, double out_re, double out_im, hipStream_t stream) {
    if (m == 0 || k == 0) {
        set_error("El Karoui fit: needs at least one grid point and one atom");
        return 1;
    }
    if (m > ((size_t)1 << 24) || k > ((size_t)1 << 24)) {
        set_error("El Karoui fit: more than 2^24 grid points or atoms is not supported");
        return 6;
    }
    double *scr = nullptr;
    CachedBuf cache(CACHE_EK_SCR, stream);
    if (int e = cache.get(6 * m, &scr)) return e;
    double *a_buf = scr, *r_buf = scr + 2 * m, *sq = scr + 4 * m, *et = scr + 5 * m;
    hipLaunchKernelGGL(ek_point_kernel, dim3((unsigned)m), dim3(64), 0, stream, x, g_re, g_im,
                       (int)m, eta, c, t, w, (int)k, a_buf, r_buf, sq, et, grad_g_re, grad_g_im,
                       in_re, in_im, out_re, out_im);
    hipError_t le = hipGetLastError();
    if (le == hipSuccess && grad_w) {
        hipLaunchKernelGGL(ek_weight_kernel, dim3((unsigned)k), dim3(64), 0, stream, x, (int)m,
                           eta, t, a_buf, r_buf, grad_w);
        le = hipGetLastError();
    }
    if (le == hipSuccess && loss) {
        hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(kSumThreads), 0, stream, sq, (int)m,
                           1.0 / (double)m, loss, 0);
        le = hipGetLastError();
    }
    if (le == hipSuccess && grad_eta) {
        hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(kSumThreads), 0, stream, et, (int)m,
                           1.0, grad_eta, 0);
        le = hipGetLastError();
    }
    if (le != hipSuccess) {
        set_error(std::string("El Karoui fit kernels: ") + hipGetErrorString(le));
        return 7;
    }
    return 0;
}

}  // namespace

int ek_fit(const double *x, const double *g_re, const double *g_im, size_t m, double eta,
           double c, const double *t, const double *w, size_t k, double *loss,
           double *grad_g_re, double *grad_g_im, double *grad_w, double *grad_eta,
           hipStream_t stream) {
    return ek_fit_impl(x, g_re, g_im, m, eta, c, t, w, k, loss, grad_g_re, grad_g_im, grad_w,
                       grad_eta, 1.0, 1.0, 1.0, 1.0, stream);
}

int ek_fit_step(const double *ev, size_t p, const double *x, size_t m, double eta, double c,
                const double *t, const double *w, size_t k, int f32c, double *loss,
                double *grad_ev, double *grad_eta, double *grad_w, hipStream_t stream) {
    if (p == 0) {
        set_error("El Karoui fit step: needs at least one eigenvalue");
        return 1;
    }
    double *scr = nullptr;
    CachedBuf cache(CACHE_EK_STEP, stream);
    if (int e = cache.get(4 * m, &scr)) return e;
    double *s_re = scr, *s_im = scr + m, *g_re = scr + 2 * m, *g_im = scr + 3 * m;
    const double inv_p = 1.0 / (double)p;
    int st;
    if (f32c) {
        st = exact_f32c(x, m, ev, p, eta, 1.0, s_re, s_im, 0, stream);
    } else {
        double *outs[2] = {s_re, s_im};
        st = exact_f64(x, m, ev, p, eta, EXACT_PLAIN, 0.0, 1.0, outs, stream);
    }
    // g = -conj(S)/p in, dL/dS = (-dL/dRe g, dL/dIm g)/p out; eta's direct term first
    if (!st)
        st = ek_fit_impl(x, s_re, s_im, m, eta, c, t, w, k, loss, g_re, g_im, grad_w, grad_eta,
                         -inv_p, inv_p, -inv_p, inv_p, stream);
    // then the part through S, added to it
    if (!st && (grad_ev || grad_eta))
        st = exact_adjoint(x, m, ev, p, eta, g_re, g_im, f32c, grad_ev, nullptr, grad_eta, stream,
                           1);
    return st;
}

}  // namespace shk
