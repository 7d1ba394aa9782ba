// Pair arithmetic shared by the exact, adjoint and batched kernels: the FP64
// reciprocal and the FP32-compensated pair forms.
//
// Eigenvalues are re-expressed relative to a tile centre c and split into
// f32 hi/lo parts; FORM selects how much of the residual enters the pair
// difference (see shk_exact.hip).  Terms are summed in f32 over runs of
// kF32Run sources and folded into f64 accumulators.
#pragma once

#include <hip/hip_runtime.h>

namespace shk {

// 1/x to ~1 ulp from the v_rcp_f64 seed r0.  With e = 1 - x*r0,
// 1/x = r0 (1 + e + e^2 + ...): the default takes the cubic correction
// r0 + r0*(e + e^2) (relative error e^3, three FMAs); the alternative takes
// two Newton-Raphson steps r <- r + r*(1 - x*r) (error e^4, four FMAs).
#ifndef SHK_F64_RCP_NEWTON2
__device__ __forceinline__ double rcp_f64(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, __builtin_fma(e, e, e), r);
}
#else
__device__ __forceinline__ double rcp_f64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
#endif

// 1/t[k] for R (even) positive denominators, two per v_rcp_f64: with
// P = t0 t1 and Q = 1/P (rcp_f64), 1/t0 = t1 Q and 1/t1 = t0 Q.  That holds
// while P is a positive normal number; a lane whose product is not (t0 t1
// overflowing or underflowing, i.e. |x - lam| near 1e154 or eta near 1e-154)
// takes rcp_f64 of each denominator instead.  Every lane's result depends
// on its own two denominators only.  Relative error about two ulp (one
// rounding each for P, Q and the product) against one for rcp_f64; per pair
// half a quarter-rate v_rcp_f64 and two FMAs' worth of full-rate work
// instead of one v_rcp_f64 and three.
template <int R>
__device__ __forceinline__ void rcp_pairs_f64(const double (&t)[R], double (&r)[R]) {
    static_assert(R % 2 == 0, "denominators are paired");
    bool ok[R / 2], all = true;
#pragma unroll
    for (int k = 0; k < R; k += 2) {
        const double p = t[k] * t[k + 1];
        const double q = rcp_f64(p);
        r[k] = t[k + 1] * q;
        r[k + 1] = t[k] * q;
        ok[k / 2] = __builtin_amdgcn_class(p, 0x100);  // a positive normal number
        all = all && ok[k / 2];
    }
    if (!all) {  // one branch for all R/2 products, skipped unless a lane needs it
#pragma unroll
        for (int k = 0; k < R; k += 2) {
            if (!ok[k / 2]) {
                r[k] = rcp_f64(t[k]);
                r[k + 1] = rcp_f64(t[k + 1]);
            }
        }
    }
}

// 1/t[k] for R (even) positive denominators, four per v_rcp_f64 for the
// leading multiple of four and two per v_rcp_f64 (rcp_pairs_f64's
// arithmetic) for a last pair, one fallback branch for all of them.  Four:
// with A = t0 t1, B = t2 t3, P = A B and Q = 1/P (rcp_f64), 1/A = B Q and
// 1/B = A Q, so 1/t0 = t1 (B Q), 1/t1 = t0 (B Q), 1/t2 = t3 (A Q) and
// 1/t3 = t2 (A Q).  That holds while A, B and P are positive normal numbers;
// a lane whose products are not takes rcp_f64 of each of the four
// denominators instead.  Every lane's result depends on its own denominators
// only.  Relative error about 2.5 ulp (five roundings: A, P, Q, B Q and the
// last product, against three for rcp_pairs_f64); per four terms one
// quarter-rate v_rcp_f64, three FMAs and nine multiplies instead of two
// reciprocals, six FMAs and six multiplies.  R = 2 gives rcp_pairs_f64's
// results; SHK_F64_RCP_QUADS = 0 takes pairs throughout.
#ifndef SHK_F64_RCP_QUADS
#define SHK_F64_RCP_QUADS 1
#endif
template <int R>
__device__ __forceinline__ void rcp_many_f64(const double (&t)[R], double (&r)[R]) {
    static_assert(R % 2 == 0, "denominators are paired");
    if constexpr (!SHK_F64_RCP_QUADS || R < 4) {
        rcp_pairs_f64<R>(t, r);
    } else {
        constexpr int Q = R - R % 4;  // denominators taken four at a time
        bool ok[R / 2], all = true;
#pragma unroll
        for (int k = 0; k < Q; k += 4) {
            const double a = t[k] * t[k + 1], b = t[k + 2] * t[k + 3];
            const double p = a * b;
            const double q = rcp_f64(p);
            const double qa = b * q, qb = a * q;
            r[k] = t[k + 1] * qa;
            r[k + 1] = t[k] * qa;
            r[k + 2] = t[k + 3] * qb;
            r[k + 3] = t[k + 2] * qb;
            ok[k / 2] = __builtin_amdgcn_class(a, 0x100) && __builtin_amdgcn_class(b, 0x100) &&
                        __builtin_amdgcn_class(p, 0x100);
            all = all && ok[k / 2];
        }
#pragma unroll
        for (int k = Q; k < R; k += 2) {
            const double p = t[k] * t[k + 1];
            const double q = rcp_f64(p);
            r[k] = t[k + 1] * q;
            r[k + 1] = t[k] * q;
            ok[k / 2] = __builtin_amdgcn_class(p, 0x100);
            all = all && ok[k / 2];
        }
        if (!all) {  // one branch for all groups, skipped unless a lane needs it
#pragma unroll
            for (int k = 0; k < Q; k += 4) {
                if (!ok[k / 2]) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) r[k + u] = rcp_f64(t[k + u]);
                }
            }
#pragma unroll
            for (int k = Q; k < R; k += 2) {
                if (!ok[k / 2]) {
                    r[k] = rcp_f64(t[k]);
                    r[k + 1] = rcp_f64(t[k + 1]);
                }
            }
        }
    }
}

// Tile pair (a, b), a <= b, of the symmetric kernels' workgroup `id` in
// [0, nt(nt+1)/2), in row-major order (row a holds b = a .. nt-1).  A
// one-dimensional grid of exactly these pairs dispatches no workgroup that
// would only exit, as the lower half of an nt x nt grid would.
__device__ __forceinline__ void tile_pair_of(int id, int nt, int &a, int &b) {
    auto start = [nt](int x) { return x * (2 * nt - x + 1) / 2; };  // first id of row x
    const double w = 2.0 * nt + 1.0;
    int r = (int)(0.5 * (w - sqrt(w * w - 8.0 * id)));
    while (r > 0 && start(r) > id) --r;
    while (r + 1 < nt && start(r + 1) <= id) ++r;
    a = r;
    b = r + (id - start(r));
}

constexpr int kF32Run = 32;  // f32 terms per FP64 fold

template <int FORM>
__device__ __forceinline__ float pair_diff(float qh, float ql, float sh, float sl) {
    if constexpr (FORM == 0) return qh - sh;
    else if constexpr (FORM == 1) return (qh - sh) + ql;
    else return (qh - sh) + (ql - sl);
}

// Sweeps `cnt` staged sources for this thread's R queries.
template <int R, int FORM>
__device__ __forceinline__ void f32c_sweep(const float *__restrict__ sh_lds,
                                           const float *__restrict__ sl_lds, int cnt,
                                           const float (&qh)[R], const float (&ql)[R],
                                           float eta2f, double (&re)[R], double (&iv)[R]) {
    int j = 0;
    for (; j + kF32Run <= cnt; j += kF32Run) {
        float ar[R], ai[R];
#pragma unroll
        for (int k = 0; k < R; ++k) ar[k] = ai[k] = 0.0f;
#pragma unroll 8
        for (int u = 0; u < kF32Run; ++u) {
            const float sh = sh_lds[j + u];
            const float sl = (FORM == 2) ? sl_lds[j + u] : 0.0f;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const float d = pair_diff<FORM>(qh[k], ql[k], sh, sl);
                const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2f));
                ar[k] = __builtin_fmaf(d, r, ar[k]);
                ai[k] += r;
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            re[k] += (double)ar[k];
            iv[k] += (double)ai[k];
        }
    }
    if (j < cnt) {
        float ar[R], ai[R];
#pragma unroll
        for (int k = 0; k < R; ++k) ar[k] = ai[k] = 0.0f;
        for (; j < cnt; ++j) {
            const float sh = sh_lds[j];
            const float sl = (FORM == 2) ? sl_lds[j] : 0.0f;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const float d = pair_diff<FORM>(qh[k], ql[k], sh, sl);
                const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2f));
                ar[k] = __builtin_fmaf(d, r, ar[k]);
                ai[k] += r;
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            re[k] += (double)ar[k];
            iv[k] += (double)ai[k];
        }
    }
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}


// Pair-difference form for a tile of half-width half_w (see shk_exact.hip).
__device__ __forceinline__ int f32c_form(double half_w, double eta) {
    return (half_w <= 2.0 * eta) ? 0 : ((half_w <= 32.0 * eta) ? 1 : 2);
}

}  // namespace shk
