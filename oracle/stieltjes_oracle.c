/*
 * CPU oracle for the Stieltjes-transform hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference
 * algorithms of alexisrosuel/shrinkers (Rust) that the MI355X library must
 * reproduce.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or as the timed CPU
 * baseline -- never as part of the product path.  The shipped library
 * (shrinkers_amd) never links, imports or calls anything in oracle/.
 *
 * Conventions (src/stieltjes/term.rs:3-7, src/stieltjes/mod.rs:29-34):
 *   every kernel returns RAW sums S_i = sum_j 1/((x_i - lam_j) - i*eta),
 *   Re S = sum d/(d^2+eta^2), Im S = eta * sum 1/(d^2+eta^2), d = x_i - lam_j,
 *   diagonal term included.  Scaling by 1/p happens in the dispatcher only.
 *
 * Floating point: Rust never contracts a*b+c, and `mul_add` is a true fused
 * multiply-add, so this file is compiled with -ffp-contract=off and every
 * reference `mul_add` is written as fma().  The reference's aarch64 SIMD
 * helpers (src/stieltjes/simd.rs) fuse their lane-wise fma and use a refined
 * reciprocal (<= 1 ulp); here fma() is fused and reciprocals are IEEE
 * divisions (the reference's own non-aarch64 arithmetic, simd.rs:29-31).
 *
 * Pinning (what checks each function against something other than itself;
 * the tests live in tests/test_oracle.py):
 *   orc_term                  closed form term.rs:51-59
 *   orc_naive_sum/autovec/    exact rational brute force (Python fractions)
 *   exact/symmetric/tiled       at p <= 24; closed form [1,2,3] at 2+0.1i
 *                               (deconvolution/mod.rs:646-658); p=1 diagonal
 *                               (cacheblock.rs:1604-1607)
 *   orc_at_points             reference golden master (deconvolution/mod.rs:
 *                               281-397, 1e-14) through the El Karoui algebra
 *   orc_with_deriv            exact rational derivative + central FD
 *                               (autovec.rs:115-148)
 *   orc_scaled_ray            exact rational sums at eta_i = h*lam_i
 *   orc_tiled_f32             reference bound vs f64 (cacheblock.rs:1767-1793)
 *                               -- weak; the f32 path is a compatibility arm
 *   orc_cheb_*                recorded rel-L2 of every f64 preset on the
 *                               bit-reproducible harness spectrum
 *                               (docs/pareto/bench_after.json), the
 *                               reference's own bounds (chebcode.rs:1402-1487)
 *   orc_harness_spectrum etc. integer LCG recomputed with Python ints
 *   orc_adjoint_at_points     central FD of orc_at_points in lam, x and eta;
 *                               translation invariance (sum of lam and x
 *                               gradients is 0); homogeneity of degree -1
 *                               (lam.grad + x.grad + eta.grad = -L); eta part
 *                               against orc_with_deriv (dS/deta = -i dS/dx)
 *   orc_nufft_*               the reference's sizing test (nufft.rs:239-263),
 *                               its accuracy test (nufft.rs:265-318) and its
 *                               recorded errors on its benchmark distribution
 *                               (docs/nufft_algorithms.md:141-147)
 *   orc_adjoint_all_points    grad_ev + grad_q of orc_adjoint_at_points at
 *                               moderate eta; exact rational sums at tiny eta;
 *                               central FD of the all-points loss
 *   orc_ek_residual           a(z) and z/a against the reference golden master
 *                               (population Im m = Im(a g), w_re = Re(z/a));
 *                               r = 0 on the Marchenko-Pastur law (H = delta_1,
 *                               g from the closed-form quadratic)
 *   orc_ek_fit                loss = mean |orc_ek_residual|^2; central FD of the
 *                               loss in Re g, Im g, w and eta
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Generators                                                                */
/* ------------------------------------------------------------------------ */

/* 64-bit LCG uniform in [0,1): src/stieltjes/testutil.rs:36-54,
 * benches/support/mod.rs:17-35. */
EXPORT void orc_lcg_fill(uint64_t seed, int64_t n, double *out) {
    uint64_t s = seed;
    for (int64_t k = 0; k < n; k++) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        out[k] = (double)(s >> 11) / (double)(1ULL << 53);
    }
}

/* ln(1..p): src/stieltjes/testutil.rs:10-12. */
EXPORT void orc_log_spectrum(int64_t p, double *out) {
    for (int64_t i = 0; i < p; i++) out[i] = log((double)i + 1.0);
}

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* The recorded benchmark spectrum "mp_c05_spikes":
 * benches/support/mod.rs:65-82. */
EXPORT void orc_harness_spectrum(int64_t p, double *out) {
    const double c = 0.5;
    double lo = (1.0 - sqrt(c)) * (1.0 - sqrt(c));
    double hi = (1.0 + sqrt(c)) * (1.0 + sqrt(c));
    if (p < 3) {
        orc_lcg_fill(42, p, out);
        for (int64_t k = 0; k < p; k++) out[k] = lo + out[k] * (hi - lo);
        return;
    }
    orc_lcg_fill(42, p - 2, out);
    for (int64_t k = 0; k < p - 2; k++) out[k] = lo + out[k] * (hi - lo);
    out[p - 2] = hi * 2.3;
    out[p - 1] = lo * 0.35;
    qsort(out, (size_t)p, sizeof(double), cmp_double);
}

/* Deterministic Fisher-Yates shuffle: src/stieltjes/testutil.rs:59-68. */
EXPORT void orc_shuffled(const double *v, int64_t n, uint64_t seed, double *out) {
    memcpy(out, v, (size_t)n * sizeof(double));
    uint64_t s = seed;
    for (int64_t i = n - 1; i >= 1; i--) {
        s = s * 6364136223846793005ULL + 1ULL;
        int64_t j = (int64_t)((s >> 33) % (uint64_t)(i + 1));
        double t = out[i];
        out[i] = out[j];
        out[j] = t;
    }
}

/* ------------------------------------------------------------------------ */
/* Term and single-point sums                                                */
/* ------------------------------------------------------------------------ */

/* src/stieltjes/term.rs:26-31. */
EXPORT void orc_term(double li, double lj, double eta, double *re, double *im) {
    double diff = li - lj;
    double denom = diff * diff + eta * eta;
    double inv = 1.0 / denom;
    *re = diff * inv;
    *im = eta * inv;
}

/* src/stieltjes/naive.rs:15-25. */
EXPORT void orc_naive_sum(double x, const double *ev, int64_t p, double eta,
                          double *re, double *im) {
    double sr = 0.0, si = 0.0;
    for (int64_t j = 0; j < p; j++) {
        double r, i;
        orc_term(x, ev[j], eta, &r, &i);
        sr += r;
        si += i;
    }
    *re = sr;
    *im = si;
}

/* src/stieltjes/autovec.rs:19-30 with term.rs:40-45 (eta hoisted). */
EXPORT void orc_autovec_sum(double x, const double *ev, int64_t p, double eta,
                            double *re, double *im) {
    double sr = 0.0, sinv = 0.0, eta2 = eta * eta;
    for (int64_t j = 0; j < p; j++) {
        double diff = x - ev[j];
        double inv = 1.0 / fma(diff, diff, eta2);
        sr += diff * inv;
        sinv += inv;
    }
    *re = sr;
    *im = eta * sinv;
}

/* Brute force all-points: src/stieltjes/testutil.rs:16-30. */
EXPORT void orc_exact_stieltjes(const double *ev, int64_t p, double eta,
                                double *out_re, double *out_im) {
    for (int64_t i = 0; i < p; i++) {
        double sr = 0.0, si = 0.0, li = ev[i];
        for (int64_t j = 0; j < p; j++) {
            double d = li - ev[j];
            double denom = d * d + eta * eta;
            sr += d / denom;
            si += eta / denom;
        }
        out_re[i] = sr;
        out_im[i] = si;
    }
}

/* Single-point sum with optional far-field window (cutoff <= 0 means none):
 * src/stieltjes/cacheblock.rs:124-151 (the window needs sorted input). */
EXPORT void orc_sum_cutoff(double x, const double *ev, int64_t p, double eta,
                           double cutoff, double *re, double *im) {
    if (!(cutoff > 0.0)) {
        orc_autovec_sum(x, ev, p, eta, re, im);
        return;
    }
    double eta2 = eta * eta, window = cutoff * eta;
    int64_t lo = 0, hi = p;
    {   /* partition_point(x < lambda - window) */
        int64_t a = 0, b = p;
        while (a < b) { int64_t m = (a + b) / 2; if (ev[m] < x - window) a = m + 1; else b = m; }
        lo = a;
        a = 0; b = p;
        while (a < b) { int64_t m = (a + b) / 2; if (ev[m] <= x + window) a = m + 1; else b = m; }
        hi = a;
    }
    double sr = 0.0, sinv = 0.0;
    for (int64_t j = lo; j < hi; j++) {
        double diff = x - ev[j];
        double inv = 1.0 / fma(diff, diff, eta2);
        sr += diff * inv;
        sinv += inv;
    }
    *re = sr;
    *im = eta * sinv;
}

/* Value + derivative w.r.t. the real query point:
 * src/stieltjes/autovec.rs:38-59 (all points: :63-76). */
EXPORT void orc_with_deriv(const double *ev, int64_t p, double eta, double *v_re,
                           double *v_im, double *d_re, double *d_im) {
    double eta_sq = eta * eta;
    for (int64_t i = 0; i < p; i++) {
        double li = ev[i], s_re = 0.0, s_inv = 0.0, dr = 0.0, di = 0.0;
        for (int64_t j = 0; j < p; j++) {
            double d = li - ev[j];
            double den = d * d + eta_sq;
            double inv = 1.0 / den;
            s_re += d * inv;
            s_inv += inv;
            dr -= (d * d - eta_sq) * inv * inv;
            di -= (2.0 * d * eta) * inv * inv;
        }
        v_re[i] = s_re;
        v_im[i] = eta * s_inv;
        d_re[i] = dr;
        d_im[i] = di;
    }
}

/* Scaled ray z_i = lam_i (1 + i*scale), eta_i = scale*lam_i, exact per-point
 * sums: src/stieltjes/mod.rs:304-319 (single-point kernel = autovec,
 * mod.rs:244-288, cutoff disabled). */
EXPORT void orc_scaled_ray(const double *ev, int64_t p, double scale,
                           double *out_re, double *out_im) {
    for (int64_t i = 0; i < p; i++)
        orc_autovec_sum(ev[i], ev, p, scale * ev[i], &out_re[i], &out_im[i]);
}

/* ------------------------------------------------------------------------ */
/* Exact all-points kernels                                                  */
/* ------------------------------------------------------------------------ */

/* Sequential symmetric half-work sweep: src/stieltjes/cacheblock.rs:841-954
 * (SoaSink :749-780).  4-row target tiles, row sums in registers, 4-column
 * source quads with a single column flush, diagonal seeded as i/eta. */
EXPORT void orc_symmetric_sweep(const double *evs, int64_t p, double eta,
                                double *out_r, double *out_i) {
    double eta_sq = eta * eta;
    for (int64_t k = 0; k < p; k++) {
        out_r[k] = 0.0;
        out_i[k] = 1.0 / eta;
    }
    int64_t t0 = 0;
    while (t0 < p) {
        int64_t t1 = t0 + 4 < p ? t0 + 4 : p;
        double rr[4] = {0, 0, 0, 0}, ri[4] = {0, 0, 0, 0};
        for (int64_t i = t0; i < t1; i++) {
            double li = evs[i];
            int64_t ii = i - t0;
            for (int64_t j = i + 1; j < t1; j++) {
                double d = li - evs[j];
                double inv = 1.0 / fma(d, d, eta_sq);
                rr[ii] = fma(d, inv, rr[ii]);
                ri[ii] = fma(eta, inv, ri[ii]);
                out_r[j] -= d * inv;
                out_i[j] += eta * inv;
            }
        }
        int64_t j0 = t1;
        while (j0 + 4 <= p) {
            double l[4] = {evs[j0], evs[j0 + 1], evs[j0 + 2], evs[j0 + 3]};
            double cr[4] = {0, 0, 0, 0}, ci[4] = {0, 0, 0, 0};
            for (int64_t i = t0; i < t1; i++) {
                double li = evs[i];
                int64_t ii = i - t0;
                for (int q = 0; q < 4; q++) {
                    double d = li - l[q];
                    double inv = 1.0 / fma(d, d, eta_sq);
                    rr[ii] = fma(d, inv, rr[ii]);
                    ri[ii] = fma(eta, inv, ri[ii]);
                    cr[q] = fma(-d, inv, cr[q]);
                    ci[q] = fma(eta, inv, ci[q]);
                }
            }
            for (int q = 0; q < 4; q++) {
                out_r[j0 + q] += cr[q];
                out_i[j0 + q] += ci[q];
            }
            j0 += 4;
        }
        while (j0 < p) {
            double lj = evs[j0];
            for (int64_t i = t0; i < t1; i++) {
                double d = evs[i] - lj;
                double inv = 1.0 / fma(d, d, eta_sq);
                rr[i - t0] = fma(d, inv, rr[i - t0]);
                ri[i - t0] = fma(eta, inv, ri[i - t0]);
                out_r[j0] -= d * inv;
                out_i[j0] += eta * inv;
            }
            j0++;
        }
        for (int64_t i = t0; i < t1; i++) {
            out_r[i] += rr[i - t0];
            out_i[i] += ri[i - t0];
        }
        t0 = t1;
    }
}

/* One output element of the full-square tiled kernel:
 * src/stieltjes/cacheblock.rs:1286-1424.  Every output element accumulates
 * the sources in ascending j with re = fma(d, inv, re), im = fma(eta, inv, im)
 * whatever the block/quad shape, so the element-wise loop below is the same
 * sequence of operations. */
static inline void tiled_element(double li, const double *ev, int64_t p, double eta,
                                 double eta_sq, double *re, double *im) {
    double r = 0.0, m = 0.0;
    for (int64_t j = 0; j < p; j++) {
        double d = li - ev[j];
        double inv = 1.0 / fma(d, d, eta_sq);
        r = fma(d, inv, r);
        m = fma(eta, inv, m);
    }
    *re = r;
    *im = m;
}

/* Parallel full-square tiled kernel with the reference's span rule:
 * src/stieltjes/cacheblock.rs:1454-1513 (bs = PARALLEL_TILED_BS = 32,
 * span_len = bs * ceil(n_blocks / (4*threads))).  Disjoint output spans,
 * no reduction.  threads <= 0 uses the OpenMP default. */
EXPORT void orc_tiled_parallel(const double *ev, int64_t p, double eta, int threads,
                               double *out_r, double *out_i) {
    if (p == 0) return;
    const int64_t bs = 32;
    double eta_sq = eta * eta;
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#endif
    int64_t n_blocks = (p + bs - 1) / bs;
    int64_t per = (n_blocks + (int64_t)nt * 4 - 1) / ((int64_t)nt * 4);
    int64_t span_len = bs * (per > 1 ? per : 1);
    int64_t n_spans = (p + span_len - 1) / span_len;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t s = 0; s < n_spans; s++) {
        int64_t a = s * span_len, b = a + span_len < p ? a + span_len : p;
        for (int64_t i = a; i < b; i++)
            tiled_element(ev[i], ev, p, eta, eta_sq, &out_r[i], &out_i[i]);
    }
}

/* f32 full-square tiled kernel: src/stieltjes/cacheblock.rs:1525-1572
 * (no cutoff branch: tiled_span_no_cutoff over all outputs). */
EXPORT void orc_tiled_f32(const float *ev, int64_t p, float eta, float *out_r,
                          float *out_i) {
    float eta_sq = eta * eta;
    for (int64_t i = 0; i < p; i++) {
        float li = ev[i], r = 0.0f, m = 0.0f;
        for (int64_t j = 0; j < p; j++) {
            float d = li - ev[j];
            float inv = 1.0f / fmaf(d, d, eta_sq);
            r = fmaf(d, inv, r);
            m = fmaf(eta, inv, m);
        }
        out_r[i] = r;
        out_i[i] = m;
    }
}

/* Query-point kernel: src/stieltjes/cacheblock.rs:167-437 with the cutoff
 * disabled.  Sources in pairs (j, j+1) outer, queries inner, each query
 * accumulating re = fma(d, inv, re), inv_sum += inv, and eta applied once at
 * the end (:218-221).  Per query the source order is ascending j. */
EXPORT void orc_at_points(const double *q, int64_t nq, const double *ev, int64_t p,
                          double eta, double *out_r, double *out_i) {
    double eta_sq = eta * eta;
    for (int64_t i = 0; i < nq; i++) {
        double r = 0.0, s = 0.0, qi = q[i];
        for (int64_t j = 0; j < p; j++) {
            double d = qi - ev[j];
            double inv = 1.0 / fma(d, d, eta_sq);
            r = fma(d, inv, r);
            s += inv;
        }
        out_r[i] = r;
        out_i[i] = s * eta;
    }
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the raw at-points sums                                         */
/* ------------------------------------------------------------------------ */

/* No reference counterpart: the reference differentiates only with respect
 * to the query (autovec.rs:38-59).  With d = x_m - lam_j, D = d^2 + eta^2,
 * u = d/D, v = eta/D, the per-pair derivatives of (Re, Im) of
 * 1/(d - i*eta) = (d + i*eta)/D are
 *   d/dx   : (v^2 - u^2, -2uv)      (the reference's dS/dx, autovec.rs:38-59)
 *   d/dlam : -(d/dx)
 *   d/deta : (-2uv, u^2 - v^2)      (S depends on x - i*eta only: dS/deta = -i dS/dx)
 * For upstream gradients G_re[m], G_im[m] of a real function of (Re S_m, Im S_m):
 *   grad_ev[j] = sum_m G_re (u^2 - v^2) + 2 G_im u v
 *   grad_q[m]  = sum_j G_re (v^2 - u^2) - 2 G_im u v
 *   grad_eta   = sum_m sum_j -2 G_re u v + G_im (u^2 - v^2)
 * Plain loops with IEEE divisions; any output pointer may be NULL. */
EXPORT void orc_adjoint_at_points(const double *q, int64_t nq, const double *ev, int64_t p,
                                  double eta, const double *gre, const double *gim,
                                  double *grad_ev, double *grad_q, double *grad_eta) {
    const double e2 = eta * eta;
    if (grad_ev) {
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < p; j++) {
            double acc = 0.0;
            for (int64_t m = 0; m < nq; m++) {
                const double d = q[m] - ev[j];
                const double den = d * d + e2;
                const double u = d / den, v = eta / den;
                acc += gre[m] * (u * u - v * v) + 2.0 * gim[m] * u * v;
            }
            grad_ev[j] = acc;
        }
    }
    if (grad_q || grad_eta) {
        double ge = 0.0;
        for (int64_t m = 0; m < nq; m++) {
            double a = 0.0, b = 0.0; /* a = sum (v^2 - u^2), b = sum u v */
            for (int64_t j = 0; j < p; j++) {
                const double d = q[m] - ev[j];
                const double den = d * d + e2;
                const double u = d / den, v = eta / den;
                a += v * v - u * u;
                b += u * v;
            }
            const double dre = a, dim = -2.0 * b;
            if (grad_q) grad_q[m] = gre[m] * dre + gim[m] * dim;
            ge += gre[m] * dim - gim[m] * dre;
        }
        if (grad_eta) *grad_eta = ge;
    }
}

/* The all-points form of orc_adjoint_at_points (queries = eigenvalues):
 * grad[i] is the total gradient with respect to lam_i through both of its
 * roles, i.e. grad_ev + grad_q of orc_adjoint_at_points(ev, ev, ...), and
 * grad_eta the same eta gradient.  The pair (i, i) contributes -G_re,i/eta^2
 * to the source role and +G_re,i/eta^2 to the query role; they cancel
 * exactly, so grad[i] sums over j != i only and does not pay the rounding
 * of 1/eta^2 (which swamps the neighbours' terms at small eta).  For j != i,
 * with d = lam_i - lam_j, u = d/D, v = eta/D:
 *   grad[i] += (G_re,j - G_re,i)(u^2 - v^2) - 2uv (G_im,i + G_im,j).
 * grad_eta keeps every ordered pair, the diagonal's -G_im,i/eta^2 included.
 * No reference counterpart (autovec.rs:38-59 differentiates the query only).
 * Either output may be NULL. */
EXPORT void orc_adjoint_all_points(const double *ev, int64_t p, double eta, const double *gre,
                                   const double *gim, double *grad, double *grad_eta) {
    const double e2 = eta * eta;
    if (grad) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < p; i++) {
            double acc = 0.0;
            for (int64_t j = 0; j < p; j++) {
                if (j == i) continue;
                const double d = ev[i] - ev[j];
                const double den = d * d + e2;
                const double u = d / den, v = eta / den;
                acc += (gre[j] - gre[i]) * (u * u - v * v) - 2.0 * u * v * (gim[i] + gim[j]);
            }
            grad[i] = acc;
        }
    }
    if (grad_eta) orc_adjoint_at_points(ev, p, ev, p, eta, gre, gim, NULL, NULL, grad_eta);
}

/* ------------------------------------------------------------------------ */
/* El Karoui bulk fit: residual of the Marchenko-Pastur equation             */
/* ------------------------------------------------------------------------ */

/* The MP equation as written in deconvolution/mod.rs:6,
 *   g(z) = int dH(t) / (t (1 - c - c z g(z)) - z),
 * for a population H = sum_k w_k delta_{t_k} and the sample Stieltjes
 * transform g (convention Im g > 0 for Im z > 0, scaled by 1/p, as the
 * reference's inversion uses it, deconvolution/mod.rs:206-248).  With
 * a_m = 1 - c - c z_m g_m and h_mk = 1/(t_k a_m - z_m), z_m = x_m + i*eta:
 *   r_m = g_m - sum_k w_k h_mk,      loss = (1/M) sum_m |r_m|^2.
 * The fit itself has no reference counterpart (the reference inverts the
 * equation pointwise, deconvolution/mod.rs:116-259).  Outputs: r and a per
 * point (either may be NULL). */
EXPORT void orc_ek_residual(const double *x, const double *g_re, const double *g_im, int64_t m,
                            double eta, double c, const double *t, const double *w, int64_t k,
                            double *r_re, double *r_im, double *a_re, double *a_im) {
    for (int64_t i = 0; i < m; i++) {
        const double complex z = x[i] + eta * I;
        const double complex g = g_re[i] + g_im[i] * I;
        const double complex a = 1.0 - c - c * z * g;
        double complex s = 0.0;
        for (int64_t q = 0; q < k; q++) s += w[q] / (t[q] * a - z);
        const double complex r = g - s;
        if (r_re) r_re[i] = creal(r);
        if (r_im) r_im[i] = cimag(r);
        if (a_re) a_re[i] = creal(a);
        if (a_im) a_im[i] = cimag(a);
    }
}

/* Loss and its gradients.  r is holomorphic in g and in z, so for the real
 * pairs (Re g, Im g) the gradient is (2/M) r conj(dr/dg), and eta enters
 * through z only as d/deta = i d/dz:
 *   dr/dg    = 1 - c z sum_k w_k t_k h^2
 *   dr/dz|g  = -sum_k w_k h^2 (1 + c t_k g)
 *   dL/dw_k  = -(2/M) sum_m Re(conj(r_m) h_mk)
 *   dL/deta  = (2/M) sum_m Re(conj(r_m) i dr/dz)       (the direct part; the
 *                                                        part through g is the
 *                                                        Stieltjes adjoint's)
 * Any gradient pointer may be NULL. */
EXPORT void orc_ek_fit(const double *x, const double *g_re, const double *g_im, int64_t m,
                       double eta, double c, const double *t, const double *w, int64_t k,
                       double *loss, double *grad_g_re, double *grad_g_im, double *grad_w,
                       double *grad_eta) {
    double l = 0.0, ge = 0.0;
    if (grad_w)
        for (int64_t q = 0; q < k; q++) grad_w[q] = 0.0;
    for (int64_t i = 0; i < m; i++) {
        const double complex z = x[i] + eta * I;
        const double complex g = g_re[i] + g_im[i] * I;
        const double complex a = 1.0 - c - c * z * g;
        double complex s = 0.0, sd = 0.0, sz = 0.0;
        for (int64_t q = 0; q < k; q++) {
            const double complex h = 1.0 / (t[q] * a - z);
            s += w[q] * h;
            sd += w[q] * t[q] * h * h;
            sz += w[q] * h * h * (1.0 + c * t[q] * g);
        }
        const double complex r = g - s;
        l += creal(r) * creal(r) + cimag(r) * cimag(r);
        const double complex jg = 1.0 - c * z * sd;
        const double complex gg = (2.0 / (double)m) * r * conj(jg);
        if (grad_g_re) grad_g_re[i] = creal(gg);
        if (grad_g_im) grad_g_im[i] = cimag(gg);
        ge += (2.0 / (double)m) * creal(conj(r) * I * (-sz));
        if (grad_w)
            for (int64_t q = 0; q < k; q++) {
                const double complex h = 1.0 / (t[q] * a - z);
                grad_w[q] -= (2.0 / (double)m) * creal(conj(r) * h);
            }
    }
    if (loss) *loss = l / (double)m;
    if (grad_eta) *grad_eta = ge;
}

/* ------------------------------------------------------------------------ */
/* ChebCode treecode                                                         */
/* ------------------------------------------------------------------------ */

#define MAX_TREE_DEPTH 48          /* chebcode.rs:95 */
/* chebcode.rs:108.  The FAST rows of docs/pareto/bench_after.json were
 * recorded before this rule existed; setting it to 1.0 reproduces them. */
static float OUTSIDE_THETA2 = 1.69f;

EXPORT void orc_set_outside_theta2(float v) { OUTSIDE_THETA2 = v; }

typedef struct {
    int n;
    double theta_sq;
    int64_t p;
    double *sorted;
    int64_t count, cap;
    double *lo, *hi, *accept_hw, *nodes, *w;
    int64_t *lo_idx, *hi_idx;
    int32_t *right;
    /* f32 mirrors, chebcode.rs:551-570 */
    float *lo32, *hi32, *acc32, *panel32;
} orc_cheb_tree;

/* chebcode.rs:242-246 */
static void cheb_nodes_m1(int n, double *s) {
    for (int j = 0; j < n; j++) s[j] = cos((double)j * M_PI / (double)(n - 1));
}

/* chebcode.rs:250-261 */
static void bary_weights_m1(int n, double *lam) {
    for (int j = 0; j < n; j++) {
        double sign = (j % 2 == 0) ? 1.0 : -1.0;
        lam[j] = (j == 0 || j == n - 1) ? 0.5 * sign : sign;
    }
}

/* chebcode.rs:309-337 */
static void barycentric_row(double x, const double *t, const double *lam, double *w,
                            double *v, double mass, int n) {
    double s = 0.0;
    int hit = -1;
    for (int j = 0; j < n; j++) {
        double d = x - t[j];
        if (d == 0.0) { hit = j; break; }
        double q = lam[j] / d;
        v[j] = q;
        s += q;
    }
    if (hit >= 0) { w[hit] += mass; return; }
    double m = mass * (1.0 / s);
    for (int j = 0; j < n; j++) w[j] = fma(m, v[j], w[j]);
}

static void tree_grow(orc_cheb_tree *t) {
    int64_t nc = t->cap ? 2 * t->cap : 64;
    int n = t->n;
    t->lo = realloc(t->lo, nc * sizeof(double));
    t->hi = realloc(t->hi, nc * sizeof(double));
    t->accept_hw = realloc(t->accept_hw, nc * sizeof(double));
    t->lo_idx = realloc(t->lo_idx, nc * sizeof(int64_t));
    t->hi_idx = realloc(t->hi_idx, nc * sizeof(int64_t));
    t->right = realloc(t->right, nc * sizeof(int32_t));
    t->nodes = realloc(t->nodes, nc * n * sizeof(double));
    t->w = realloc(t->w, nc * n * sizeof(double));
    t->cap = nc;
}

/* Recursive topology pass: chebcode.rs:347-388 (positions :269-279). */
static int32_t build_topology(orc_cheb_tree *t, const double *sm1, int64_t lo_idx,
                              int64_t hi_idx, int64_t leaf_cap, int depth) {
    if (t->count == t->cap) tree_grow(t);
    int32_t idx = (int32_t)t->count++;
    int64_t count = hi_idx - lo_idx;
    double lo = t->sorted[lo_idx], hi = t->sorted[hi_idx - 1];
    t->lo[idx] = lo;
    t->hi[idx] = hi;
    double hw = 0.5 * (hi - lo);
    t->accept_hw[idx] = (hw * hw) / t->theta_sq;
    t->lo_idx[idx] = lo_idx;
    t->hi_idx[idx] = hi_idx;
    t->right[idx] = -1;
    int n = t->n;
    double c = 0.5 * (lo + hi), l = 0.5 * (hi - lo);
    for (int j = 0; j < n; j++) {
        t->nodes[(int64_t)idx * n + j] = c + l * sm1[j];
        t->w[(int64_t)idx * n + j] = 0.0;
    }
    if (count <= leaf_cap || depth >= MAX_TREE_DEPTH) return idx;
    double mid = lo + 0.5 * (hi - lo);
    int64_t a = lo_idx, b = hi_idx; /* partition_point(x <= mid) */
    while (a < b) { int64_t m = a + (b - a) / 2; if (t->sorted[m] <= mid) a = m + 1; else b = m; }
    int64_t split = a;
    if (split == lo_idx || split == hi_idx) split = (lo_idx + hi_idx) / 2;
    build_topology(t, sm1, lo_idx, split, leaf_cap, depth + 1);
    int32_t ri = build_topology(t, sm1, split, hi_idx, leaf_cap, depth + 1);
    t->right[idx] = ri;
    return idx;
}

/* FlatChebTree::build: chebcode.rs:509-572 (leaf fill :401-460, merge
 * :474-506).  `sorted` must be ascending. */
EXPORT orc_cheb_tree *orc_cheb_build(const double *sorted, int64_t p, int n, double theta,
                                     int64_t leaf_cap) {
    if (p <= 0 || n < 2) return NULL;
    orc_cheb_tree *t = calloc(1, sizeof(orc_cheb_tree));
    t->n = n;
    t->theta_sq = theta * theta;
    t->p = p;
    t->sorted = malloc(p * sizeof(double));
    memcpy(t->sorted, sorted, p * sizeof(double));
    double *sm1 = malloc(n * sizeof(double)), *lam = malloc(n * sizeof(double));
    double *v = malloc(n * sizeof(double));
    cheb_nodes_m1(n, sm1);
    bary_weights_m1(n, lam);
    build_topology(t, sm1, 0, p, leaf_cap, 0);
    /* leaves */
    for (int64_t idx = 0; idx < t->count; idx++) {
        if (t->right[idx] >= 0) continue;
        double *w = &t->w[idx * n];
        const double *tn = &t->nodes[idx * n];
        for (int j = 0; j < n; j++) w[j] = 0.0;
        for (int64_t k = t->lo_idx[idx]; k < t->hi_idx[idx]; k++)
            barycentric_row(t->sorted[k], tn, lam, w, v, 1.0, n);
    }
    /* parents, reverse pre-order */
    for (int64_t pi = t->count - 1; pi >= 0; pi--) {
        if (t->right[pi] < 0) continue;
        double *pw = &t->w[pi * n];
        const double *tn = &t->nodes[pi * n];
        for (int j = 0; j < n; j++) pw[j] = 0.0;
        int64_t kids[2] = {pi + 1, t->right[pi]};
        for (int k = 0; k < 2; k++) {
            int64_t cb = kids[k] * n;
            for (int ti = 0; ti < n; ti++) {
                double m = t->w[cb + ti];
                if (m == 0.0) continue;
                barycentric_row(t->nodes[cb + ti], tn, lam, pw, v, m, n);
            }
        }
    }
    /* f32 mirrors */
    t->lo32 = malloc(t->count * sizeof(float));
    t->hi32 = malloc(t->count * sizeof(float));
    t->acc32 = malloc(t->count * sizeof(float));
    t->panel32 = malloc(t->count * 2 * n * sizeof(float));
    for (int64_t i = 0; i < t->count; i++) {
        t->lo32[i] = (float)t->lo[i];
        t->hi32[i] = (float)t->hi[i];
        t->acc32[i] = (float)t->accept_hw[i];
        for (int j = 0; j < n; j++) {
            t->panel32[i * 2 * n + j] = (float)t->nodes[i * n + j];
            t->panel32[i * 2 * n + n + j] = (float)t->w[i * n + j];
        }
    }
    free(sm1); free(lam); free(v);
    return t;
}

EXPORT void orc_cheb_free(orc_cheb_tree *t) {
    if (!t) return;
    free(t->sorted); free(t->lo); free(t->hi); free(t->accept_hw); free(t->nodes);
    free(t->w); free(t->lo_idx); free(t->hi_idx); free(t->right); free(t->lo32);
    free(t->hi32); free(t->acc32); free(t->panel32); free(t);
}

EXPORT int64_t orc_cheb_node_count(const orc_cheb_tree *t) { return t ? t->count : 0; }

/* Copies the tree arrays out (any pointer may be NULL). */
EXPORT void orc_cheb_export(const orc_cheb_tree *t, double *lo, double *hi, int64_t *lo_idx,
                            int64_t *hi_idx, int32_t *right, double *nodes, double *w) {
    int64_t c = t->count, n = t->n;
    if (lo) memcpy(lo, t->lo, c * sizeof(double));
    if (hi) memcpy(hi, t->hi, c * sizeof(double));
    if (lo_idx) memcpy(lo_idx, t->lo_idx, c * sizeof(int64_t));
    if (hi_idx) memcpy(hi_idx, t->hi_idx, c * sizeof(int64_t));
    if (right) memcpy(right, t->right, c * sizeof(int32_t));
    if (nodes) memcpy(nodes, t->nodes, c * n * sizeof(double));
    if (w) memcpy(w, t->w, c * n * sizeof(double));
}

/* Exact leaf sum in two f64 lanes + scalar tail: chebcode.rs:622-647. */
static inline void leaf_sum(const double *xs, int64_t len, double x, double eta2,
                            double *lr_out, double *li_out) {
    double ar0 = 0, ar1 = 0, ai0 = 0, ai1 = 0;
    int64_t k = 0;
    while (k + 2 <= len) {
        double d0 = x - xs[k], d1 = x - xs[k + 1];
        double inv0 = 1.0 / fma(d0, d0, eta2), inv1 = 1.0 / fma(d1, d1, eta2);
        ar0 = fma(d0, inv0, ar0);
        ar1 = fma(d1, inv1, ar1);
        ai0 += inv0;
        ai1 += inv1;
        k += 2;
    }
    double lr = ar0 + ar1, li = ai0 + ai1;
    while (k < len) {
        double d = x - xs[k];
        double inv = 1.0 / (d * d + eta2);
        lr += d * inv;
        li += inv;
        k++;
    }
    *lr_out = lr;
    *li_out = li;
}

/* Full-precision traversal: chebcode.rs:607-711.  If `near_pairs` is not
 * NULL it receives the number of exact leaf pair-evaluations. */
EXPORT void orc_cheb_eval_f64(const orc_cheb_tree *t, double x, double eta, double *re_out,
                              double *im_out, int64_t *near_pairs, int64_t *far_panels) {
    int n = t->n;
    int32_t stack[MAX_TREE_DEPTH + 2];
    int top = 0;
    stack[top++] = 0;
    double re = 0.0, im = 0.0, eta2 = eta * eta;
    int64_t np = 0, fp = 0;
    while (top > 0) {
        int32_t ni = stack[--top];
        if (t->right[ni] < 0) {
            double lr, li;
            int64_t a = t->lo_idx[ni], len = t->hi_idx[ni] - a;
            leaf_sum(&t->sorted[a], len, x, eta2, &lr, &li);
            np += len;
            re += lr;
            im += eta * li;
            continue;
        }
        double lo = t->lo[ni], hi = t->hi[ni];
        double cl = fmax(fmax(lo - x, x - hi), 0.0);
        double d_sq = fma(cl, cl, eta2);
        if (t->accept_hw[ni] < d_sq) {
            const double *ts = &t->nodes[(int64_t)ni * n], *ws = &t->w[(int64_t)ni * n];
            double afr0 = 0, afr1 = 0, afi0 = 0, afi1 = 0;
            int j = 0;
            while (j + 2 <= n) {
                double d0 = x - ts[j], d1 = x - ts[j + 1];
                double inv0 = 1.0 / fma(d0, d0, eta2), inv1 = 1.0 / fma(d1, d1, eta2);
                double q0 = ws[j] * inv0, q1 = ws[j + 1] * inv1;
                afr0 = fma(q0, d0, afr0);
                afr1 = fma(q1, d1, afr1);
                afi0 += q0;
                afi1 += q1;
                j += 2;
            }
            double fr = afr0 + afr1, fi = afi0 + afi1;
            while (j < n) {
                double dj = x - ts[j];
                double q = ws[j] / (dj * dj + eta2);
                fr += q * dj;
                fi += q;
                j++;
            }
            re += fr;
            im += eta * fi;
            fp++;
        } else {
            stack[top++] = t->right[ni];
            stack[top++] = ni + 1;
        }
    }
    *re_out = re;
    *im_out = im;
    if (near_pairs) *near_pairs = np;
    if (far_panels) *far_panels = fp;
}

/* FAST traversal, f32 far field + f64 leaves: chebcode.rs:722-828
 * (reciprocal as IEEE division, simd.rs:29-31). */
EXPORT void orc_cheb_eval_fast32(const orc_cheb_tree *t, double xd, double eta, double *re_out,
                                 double *im_out, int64_t *near_pairs, int64_t *far_panels) {
    int n = t->n;
    int32_t stack[MAX_TREE_DEPTH + 2];
    int top = 0;
    stack[top++] = 0;
    float x = (float)xd, e32 = (float)eta, eta2 = e32 * e32;
    float far_re[4] = {0, 0, 0, 0}, far_im[4] = {0, 0, 0, 0};
    double leaf_re = 0.0, leaf_im = 0.0, eta2d = eta * eta;
    int64_t np = 0, fp = 0;
    while (top > 0) {
        int32_t node = stack[--top];
        if (t->right[node] < 0) {
            double lr, li;
            int64_t a = t->lo_idx[node], len = t->hi_idx[node] - a;
            leaf_sum(&t->sorted[a], len, xd, eta2d, &lr, &li);
            np += len;
            leaf_re += lr;
            leaf_im += eta * li;
            continue;
        }
        float cl = fmaxf(fmaxf(t->lo32[node] - x, x - t->hi32[node]), 0.0f);
        float d_sq = fmaf(cl, cl, eta2);
        float outside = (cl == 0.0f) ? d_sq : OUTSIDE_THETA2 * d_sq;
        if (t->acc32[node] < outside) {
            const float *ts = &t->panel32[(int64_t)node * 2 * n];
            const float *ws = ts + n;
            int j = 0;
            while (j + 4 <= n) {
                for (int l = 0; l < 4; l++) {
                    float d = x - ts[j + l];
                    float inv = 1.0f / fmaf(d, d, eta2);
                    float q = ws[j + l] * inv;
                    far_re[l] = fmaf(q, d, far_re[l]);
                    far_im[l] = far_im[l] + q;
                }
                j += 4;
            }
            if (j < n) {
                for (int l = 0; l < 4; l++) {
                    float d = 0.0f, wt = 0.0f;
                    if (j + l < n) { d = x - ts[j + l]; wt = ws[j + l]; }
                    float inv = 1.0f / fmaf(d, d, eta2);
                    float q = wt * inv;
                    far_re[l] = fmaf(q, d, far_re[l]);
                    far_im[l] = far_im[l] + q;
                }
            }
            fp++;
        } else {
            stack[top++] = t->right[node];
            stack[top++] = node + 1;
        }
    }
    float hr = (far_re[0] + far_re[1]) + (far_re[2] + far_re[3]);
    float hi = (far_im[0] + far_im[1]) + (far_im[2] + far_im[3]);
    *re_out = leaf_re + (double)hr;
    *im_out = leaf_im + eta * (double)hi;
    if (near_pairs) *near_pairs = np;
    if (far_panels) *far_panels = fp;
}

/* Two-eta traversal, acceptance at min(eta0, eta1): chebcode.rs:836-915. */
EXPORT void orc_cheb_eval_x2(const orc_cheb_tree *t, double x, double e0, double e1,
                             double *r0, double *i0, double *r1, double *i1) {
    int n = t->n;
    int32_t stack[MAX_TREE_DEPTH + 2];
    int top = 0;
    stack[top++] = 0;
    double e[2] = {e0, e1}, e2[2] = {e0 * e0, e1 * e1};
    double ar[2] = {0, 0}, ai[2] = {0, 0};
    while (top > 0) {
        int32_t ni = stack[--top];
        if (t->right[ni] < 0) {
            for (int64_t k = t->lo_idx[ni]; k < t->hi_idx[ni]; k++) {
                double d = x - t->sorted[k];
                for (int l = 0; l < 2; l++) {
                    double inv = 1.0 / fma(d, d, e2[l]);
                    ar[l] = fma(d, inv, ar[l]);
                    ai[l] = fma(e[l], inv, ai[l]);
                }
            }
            continue;
        }
        double lo = t->lo[ni], hi = t->hi[ni];
        double cl = x < lo ? lo - x : (x > hi ? x - hi : 0.0);
        double emin = e0 < e1 ? e0 : e1;
        double d_sq = cl * cl + emin * emin;
        double half_w = 0.5 * (hi - lo);
        if (half_w * half_w < t->theta_sq * d_sq) {
            for (int j = 0; j < n; j++) {
                double dj = x - t->nodes[(int64_t)ni * n + j];
                double wj = t->w[(int64_t)ni * n + j];
                for (int l = 0; l < 2; l++) {
                    double inv = 1.0 / fma(dj, dj, e2[l]);
                    ar[l] = fma(wj * dj, inv, ar[l]);
                    ai[l] = fma(wj * e[l], inv, ai[l]);
                }
            }
        } else {
            stack[top++] = t->right[ni];
            stack[top++] = ni + 1;
        }
    }
    *r0 = ar[0]; *i0 = ai[0]; *r1 = ar[1]; *i1 = ai[1];
}

static int is_sorted_ascending(const double *v, int64_t p) {
    for (int64_t k = 1; k < p; k++) if (!(v[k - 1] <= v[k])) return 0;
    return 1;
}

/* All-points driver: chebcode.rs:1079-1120 (sort check mod.rs:542-544;
 * Rayon chunks of EVAL_CHUNK = 128 at :156-175 become OpenMP chunks).
 * fast32 != 0 selects the f32 far field.  Queries are the input order. */
EXPORT void orc_chebcode_all(const double *ev, int64_t p, double eta, double theta, int n,
                             int64_t leaf_cap, int fast32, int threads, double *out_r,
                             double *out_i) {
    if (p == 0) return;
    double *sorted = malloc(p * sizeof(double));
    memcpy(sorted, ev, p * sizeof(double));
    if (!is_sorted_ascending(sorted, p)) qsort(sorted, (size_t)p, sizeof(double), cmp_double);
    orc_cheb_tree *t = orc_cheb_build(sorted, p, n, theta, leaf_cap);
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#endif
#pragma omp parallel for schedule(dynamic, 128) num_threads(nt)
    for (int64_t i = 0; i < p; i++) {
        if (fast32)
            orc_cheb_eval_fast32(t, ev[i], eta, &out_r[i], &out_i[i], NULL, NULL);
        else
            orc_cheb_eval_f64(t, ev[i], eta, &out_r[i], &out_i[i], NULL, NULL);
    }
    orc_cheb_free(t);
    free(sorted);
}

/* Evaluate a built tree at arbitrary points: chebcode.rs:1241-1277. */
EXPORT void orc_cheb_eval_points(const orc_cheb_tree *t, const double *q, int64_t nq,
                                 double eta, int fast32, double *out_r, double *out_i) {
    for (int64_t i = 0; i < nq; i++) {
        if (fast32)
            orc_cheb_eval_fast32(t, q[i], eta, &out_r[i], &out_i[i], NULL, NULL);
        else
            orc_cheb_eval_f64(t, q[i], eta, &out_r[i], &out_i[i], NULL, NULL);
    }
}

/* Total traversal work over nq points: near-field pairs and accepted far
 * panels of the f64 (chebcode.rs:607-711) or FAST (:722-828) traversal. */
EXPORT void orc_cheb_count_work(const orc_cheb_tree *t, const double *q, int64_t nq, double eta,
                                int fast32, int64_t *near_total, int64_t *panels_total) {
    int64_t nt = 0, pt = 0;
    for (int64_t i = 0; i < nq; i++) {
        double r, im;
        int64_t np = 0, fp = 0;
        if (fast32)
            orc_cheb_eval_fast32(t, q[i], eta, &r, &im, &np, &fp);
        else
            orc_cheb_eval_f64(t, q[i], eta, &r, &im, &np, &fp);
        nt += np;
        pt += fp;
    }
    *near_total = nt;
    *panels_total = pt;
}

/* Grid leaf capacity L* = n*sqrt(2p/nq) clamped to [preset, 8192]:
 * src/stieltjes/mod.rs:361-370. */
EXPORT int64_t orc_grid_leaf_cap(int64_t n, int64_t p, int64_t nq, int64_t preset_leaf) {
    const int64_t MAX_GRID_LEAF = 8192;
    if (nq == 0 || p == 0) return preset_leaf;
    double optimal = (double)n * sqrt(2.0 * (double)p / (double)nq);
    int64_t o = (int64_t)llround(optimal);
    int64_t hi = MAX_GRID_LEAF > preset_leaf ? MAX_GRID_LEAF : preset_leaf;
    if (o < preset_leaf) o = preset_leaf;
    if (o > hi) o = hi;
    return o;
}

/* Threads the OpenMP runtime would use (for reporting the CPU baseline). */
EXPORT int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Coarse-grid NUFFT (the reference's CPU pick above p = 50 000)             */
/* ------------------------------------------------------------------------ */

/* Grid size rule, src/stieltjes/nufft.rs:104-124 with the defaults of
 * :33-57 (pad 2000 eta or 0.75 range, 5 cells per eta, grid in
 * [8192, 65536], power of two); m_override > 0 forces it (:109-111).
 * Pinned by the reference's own sizing test (nufft.rs:239-263). */
EXPORT int64_t orc_nufft_grid_size(double raw_range, double eta, double pad_eta_mult,
                                   double pad_range_frac, int64_t m_override) {
    const int64_t min_grid = 8192, max_grid_opt = 65536;
    if (m_override > 0) return m_override < 2 ? 2 : m_override;
    double pad = pad_eta_mult * eta;
    if (pad_range_frac * raw_range > pad) pad = pad_range_frac * raw_range;
    const double range = raw_range + 2.0 * pad;
    int64_t max_grid = max_grid_opt > min_grid ? max_grid_opt : min_grid;
    if (max_grid < 2) max_grid = 2;
    double target = 5.0 * range / eta;
    target = isfinite(target) ? (target < (double)max_grid ? target : (double)max_grid)
                              : (double)max_grid;
    int64_t t = (int64_t)ceil(target);
    if (t < min_grid) t = min_grid;
    int64_t m = 1;
    while (m < t) m <<= 1;  /* next_pow2, mod.rs:531-537 */
    return m < max_grid ? m : max_grid;
}

/* Lagrange weights on N uniform nodes at offset t in [0, 1) from node i0
 * (fft5.rs:136-164): prefix/suffix products over (t - x_l), x_l = l - (N/2-1). */
static void nufft_lagrange(int N, double t, double *w) {
    const double q = (double)(N / 2);
    double pref[8], suf[8], acc = 1.0;
    for (int j = 0; j < N; j++) {
        pref[j] = acc;
        acc *= t - (double)j + q - 1.0;
    }
    acc = 1.0;
    for (int j = N - 1; j >= 0; j--) {
        suf[j] = acc;
        acc *= t - (double)j + q - 1.0;
    }
    for (int j = 0; j < N; j++) {
        double den = 1.0;
        for (int l = 0; l < N; l++)
            if (l != j) den *= (double)(j - l);
        w[j] = pref[j] * suf[j] / den;
    }
}

/* In-place complex DFT of length m (a power of two): iterative radix-2,
 * X[k] = sum_j x[j] exp(sign * 2 pi i jk / m), unnormalised -- the
 * convention of rustfft's Forward (sign -1) and Inverse (sign +1)
 * (fftplan.rs:26-28).  Rounding differs from rustfft's mixed-radix plans
 * (~1e-15 relative), far below the method's ~1e-5 error. */
static void nufft_fft(double complex *x, int64_t m, int sign) {
    for (int64_t i = 1, j = 0; i < m; i++) {
        int64_t bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double complex t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (int64_t len = 2; len <= m; len <<= 1) {
        const double ang = sign * 2.0 * M_PI / (double)len;
        for (int64_t k = 0; k < len / 2; k++) {
            const double complex w = cos(ang * (double)k) + I * sin(ang * (double)k);
            for (int64_t s = 0; s < m; s += len) {
                const double complex u = x[s + k], v = x[s + k + len / 2] * w;
                x[s + k] = u + v;
                x[s + k + len / 2] = u - v;
            }
        }
    }
}

/* NUFFT Stieltjes sums at nq query points (raw, as the reference returns
 * them): nufft.rs:162-200 with NufftOptions::default (Cubic stencil, :80-91)
 * over fft5::fft_convolution (fft5.rs:325-433) and interpolate_grid
 * (:440-476): splat the eigenvalues (sorted) onto m cells of the padded
 * range, one forward FFT of density + i*R_odd (R_odd(x) = x/(x^2+eta^2) on
 * signed cell distances), unpack both spectra (fftplan.rs:43-54), the even
 * kernel's spectrum analytically (pi/dx) r^|k| with r = exp(-2 pi eta /
 * (m dx)) by recurrence, pack D K_even + i D R_odd (fftplan.rs:62-76), one
 * inverse FFT, then a 4-point Lagrange interpolation of (Im, Re) / m at each
 * query.  threads > 0: per-thread splat buffers added element-wise (the
 * reference's parallel splat, fft5.rs:259-293) and parallel interpolation;
 * the FFTs stay sequential, as in the reference.  Pinned by the reference's
 * test bound (nufft.rs:265-318: < 1e-4 normwise on the log spectrum at p =
 * 8192, all points and a 200-point grid), its sizing test, and its
 * documented error against the exact sums (~2e-5 at p = 5e4, ~4e-5 at 1e6,
 * nufft.rs:30-32). */
EXPORT void orc_nufft_at_points(const double *ev, int64_t p, const double *q, int64_t nq,
                                double eta, int threads, double *out_re, double *out_im) {
    const int N = 4; /* Order::Cubic */
    if (p == 0 || nq == 0) return;
    const double lo_raw = ev[0], hi_raw = ev[p - 1], raw_range = hi_raw - lo_raw;
    const int64_t m = orc_nufft_grid_size(raw_range, eta, 2000.0, 0.75, 0);
    double pad = 2000.0 * eta;
    if (0.75 * raw_range > pad) pad = 0.75 * raw_range;
    const double lo = lo_raw - pad, hi = hi_raw + pad, range = hi - lo;
    const double dx = range / (double)m;
    const int64_t half = m / 2;
    const int off0 = -N / 2 + 1;
    double *density = calloc((size_t)m, sizeof(double));
#ifdef _OPENMP
    const int nthr = threads > 0 ? threads : 1;
#else
    const int nthr = 1;
#endif
    if (nthr > 1 && p >= 4096) {
        const int64_t chunk = (p + nthr - 1) / nthr;
        double *bufs = calloc((size_t)m * (size_t)nthr, sizeof(double));
#pragma omp parallel for num_threads(nthr) schedule(static)
        for (int w = 0; w < nthr; w++) {
            double *d = bufs + (size_t)w * (size_t)m;
            const int64_t a = (int64_t)w * chunk, b = a + chunk < p ? a + chunk : p;
            for (int64_t s = a; s < b; s++) {
                const double pos = (ev[s] - lo) / dx;
                const int64_t i0 = (int64_t)floor(pos);
                double wts[8];
                nufft_lagrange(N, pos - (double)i0, wts);
                for (int j = 0; j < N; j++) {
                    int64_t cell = i0 + off0 + j;
                    cell = cell < 0 ? 0 : (cell > m - 1 ? m - 1 : cell);
                    d[cell] += wts[j];
                }
            }
        }
        /* reduce in worker order (the reference's reduce order is Rayon's) */
        for (int w = 0; w < nthr; w++)
            for (int64_t i = 0; i < m; i++) density[i] += bufs[(size_t)w * (size_t)m + i];
        free(bufs);
    } else {
        for (int64_t s = 0; s < p; s++) {
            const double pos = (ev[s] - lo) / dx;
            const int64_t i0 = (int64_t)floor(pos);
            double wts[8];
            nufft_lagrange(N, pos - (double)i0, wts);
            for (int j = 0; j < N; j++) {
                int64_t cell = i0 + off0 + j;
                cell = cell < 0 ? 0 : (cell > m - 1 ? m - 1 : cell);
                density[cell] += wts[j];
            }
        }
    }
    double complex *packed = malloc((size_t)m * sizeof(double complex));
    for (int64_t i = 0; i < m; i++) {
        const double sd = i <= half ? (double)i : (double)i - (double)m;
        const double x = sd * dx, den = x * x + eta * eta;
        packed[i] = density[i] + I * (x / den);
    }
    free(density);
    nufft_fft(packed, m, -1);
    /* even kernel spectrum by recurrence (fft5.rs:383-405) */
    double *ke = malloc((size_t)m * sizeof(double));
    const double r = exp(-2.0 * M_PI * eta / ((double)m * dx)), ke0 = M_PI / dx;
    double pw = 1.0;
    for (int64_t k = 0; k <= half; k++) {
        ke[k] = ke0 * pw;
        pw *= r;
    }
    pw = r;
    for (int64_t k = m - 1; k > half; k--) {
        ke[k] = ke0 * pw;
        pw *= r;
    }
    double complex *out = malloc((size_t)m * sizeof(double complex));
    for (int64_t k = 0; k < m; k++) {
        const double complex ck = packed[k], cnk = conj(packed[(m - k) % m]);
        const double complex d = 0.5 * (ck + cnk);
        const double complex ko = -0.5 * I * (ck - cnk);
        const double complex im_hat = d * ke[k], re_hat = d * ko;
        out[k] = (creal(im_hat) - cimag(re_hat)) + I * (cimag(im_hat) + creal(re_hat));
    }
    free(ke);
    free(packed);
    nufft_fft(out, m, +1);
    const double inv_m = 1.0 / (double)m;
#pragma omp parallel for num_threads(nthr) schedule(static) if (nthr > 1 && nq >= 4096)
    for (int64_t k = 0; k < nq; k++) {
        const double pos = (q[k] - lo) / dx;
        const int64_t i0 = (int64_t)floor(pos);
        double wts[8];
        nufft_lagrange(N, pos - (double)i0, wts);
        double re_acc = 0.0, im_acc = 0.0;
        for (int j = 0; j < N; j++) {
            int64_t cell = i0 + off0 + j;
            cell = cell < 0 ? 0 : (cell > m - 1 ? m - 1 : cell);
            im_acc += creal(out[cell]) * wts[j];
            re_acc += cimag(out[cell]) * wts[j];
        }
        out_re[k] = re_acc * inv_m;
        out_im[k] = im_acc * inv_m;
    }
    free(out);
}
