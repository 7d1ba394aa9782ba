// ChebCode Chebyshev treecode on gfx950.
//
// Reference: src/stieltjes/chebcode.rs.  The tree is the reference's
// (midpoint value split with index-midpoint fallback, leaf if count <=
// leaf_cap or depth >= 48, pre-order numbering with left child = node+1,
// chebcode.rs:347-388), built entirely on the GPU:
//
//   1. sort (rocPRIM radix sort of (value, index) pairs) unless the input is
//      already ascending; the permutation scatters results back to input
//      order (ChebCodeBatch::build_mode_parallel, :1186-1217);
//   2. topology: one workgroup walks the tree breadth-first, one thread per
//      frontier node, with a workgroup scan placing the children so the BFS
//      order is deterministic; subtree sizes bottom-up and pre-order numbers
//      top-down follow in the same kernel;
//   3. finalize: pre-order node records, Chebyshev positions (:269-279);
//   4. leaf fill: one thread per leaf, barycentric rows in source order
//      (:286-337, :401-460);
//   5. merge: parents from their children's weights, deepest level first
//      (:474-506);
//   6. f32 mirrors for the FAST traversal (:551-570) and f32 hi/lo splits of
//      the sorted sources for the FAST near field.
//
// Evaluation is one thread per query, queries in ascending order so that the
// 64 lanes of a wavefront walk nearly the same path.  The reference's DFS
// with an explicit stack (push right, then left) visits nodes in pre-order
// with pruned subtrees skipped; the GPU walk reproduces exactly that order
// without a stack: a pruned or finished node continues at its escape index
// (node + subtree size), a descended node at node + 1.  Per query the
// contributions therefore accumulate in the reference's order.
//
//   F64 traversal (DEFAULT/BALANCED/XTREME, :607-711): f64 leaves and far
//   field, two-lane partial sums as in the reference's F64x2 loops.
//   FAST traversal (:722-828): f32 node test with OUTSIDE_THETA2 = 1.69 for
//   panels not containing the query, four-lane f32 far field with the
//   v_rcp_f32 reciprocal, folded to f64 once per query.  The near field
//   (leaves) runs in f32 on hi/lo-split eigenvalues,
//   d = (x_hi - s_hi) + (x_lo - s_lo), with per-leaf f32 sums folded into
//   f64; the reference keeps these leaves in f64, the split keeps the pair
//   difference within one f32 rounding so the leaf error (~1e-7) stays far
//   below the preset's interpolation error (~2e-5).
#include "shk_internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <memory>
#include <vector>

namespace shk {

struct ChebTree {
    size_t p = 0, N = 0;
    int n = 0;
    double theta = 0, theta_sq = 0;
    int64_t leaf_cap = 0;
    int f32_far = 0;
    int n_levels = 0;
    double *sorted = nullptr;  // p ascending
    int32_t *perm = nullptr;   // p, or null when the input was ascending
    float *s_hi = nullptr, *s_lo = nullptr;
    // pre-order records
    double *lo = nullptr, *hi = nullptr, *acc = nullptr;
    int32_t *lo_idx = nullptr, *hi_idx = nullptr, *right = nullptr, *esc = nullptr;
    double *nodes = nullptr, *w = nullptr;  // N*n
    float4 *rec32 = nullptr;                // {lo, hi, acc, unused} per node
    int4 *link = nullptr;                   // {right, esc, lo_idx, hi_idx} per node
    float *panel32 = nullptr;               // N*2n: nodes then weights
    void *arena = nullptr;   // sources (sorted, perm, hi/lo)
    void *arena2 = nullptr;  // node records and panels
    hipEvent_t ready = nullptr;  // recorded when the build's last kernel is enqueued
};

namespace {

constexpr int kMaxDepth = 48;    // chebcode.rs:95
constexpr int kTopoThreads = 1024;
constexpr int kMaxN = 120;       // largest interpolation order (LDS of the build)
constexpr float kOutsideTheta2 = 1.69f;  // chebcode.rs:108
constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI

__device__ __forceinline__ double rcp_f64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__global__ void check_sorted_kernel(const double *__restrict__ v, int p, int *__restrict__ bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i < p && !(v[i - 1] <= v[i])) atomicOr(bad, 1);
}

__global__ void iota_kernel(int32_t *__restrict__ v, int p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p) v[i] = i;
}

// Workgroup exclusive scan of one int per thread (1024 threads).
__device__ int block_exclusive_scan(int v, int *s, int *total) {
    const int tid = threadIdx.x;
    s[tid] = v;
    __syncthreads();
    for (int off = 1; off < kTopoThreads; off <<= 1) {
        int t = (tid >= off) ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += t;
        __syncthreads();
    }
    int incl = s[tid];
    *total = s[kTopoThreads - 1];
    __syncthreads();
    return incl - v;
}

// Breadth-first topology, then subtree sizes and pre-order numbers.
// b_* arrays are indexed by BFS id and hold `cap` entries.  info[0] = node
// count, info[1] = number of levels, info[2] = overflow flag.
__global__ __launch_bounds__(kTopoThreads) void topo_kernel(
    const double *__restrict__ sorted, int p, int leaf_cap, int cap, int *__restrict__ b_lo,
    int *__restrict__ b_hi, int *__restrict__ b_left, int *__restrict__ b_size,
    int *__restrict__ b_pre, int *__restrict__ info) {
    __shared__ int s_scan[kTopoThreads];
    __shared__ int s_level[kMaxDepth + 3];
    const int tid = threadIdx.x;
    if (tid == 0) {
        b_lo[0] = 0;
        b_hi[0] = p;
        s_level[0] = 0;
        s_level[1] = 1;
    }
    __syncthreads();
    int lvl_begin = 0, lvl_end = 1, total = 1, depth = 0, overflow = 0;
    while (lvl_begin < lvl_end && !overflow) {
        for (int base = lvl_begin; base < lvl_end; base += kTopoThreads) {
            const int v = base + tid;
            int a = 0, b = 0, split = 0, nchild = 0;
            if (v < lvl_end) {
                a = b_lo[v];
                b = b_hi[v];
                if (b - a > leaf_cap && depth < kMaxDepth) {
                    const double lo = sorted[a], hi = sorted[b - 1];
                    const double mid = lo + 0.5 * (hi - lo);
                    int l = a, r = b;  // partition_point(x <= mid)
                    while (l < r) {
                        const int m = l + (r - l) / 2;
                        if (sorted[m] <= mid) l = m + 1;
                        else r = m;
                    }
                    split = (l == a || l == b) ? (a + b) / 2 : l;
                    nchild = 2;
                }
            }
            int added;
            const int off = block_exclusive_scan(nchild, s_scan, &added);
            if (total + added > cap) {
                overflow = 1;
                break;
            }
            if (v < lvl_end) {
                if (nchild) {
                    const int c = total + off;
                    b_left[v] = c;
                    b_lo[c] = a;
                    b_hi[c] = split;
                    b_lo[c + 1] = split;
                    b_hi[c + 1] = b;
                } else {
                    b_left[v] = -1;
                }
            }
            total += added;
        }
        __syncthreads();
        lvl_begin = lvl_end;
        lvl_end = total;
        ++depth;
        if (tid == 0) s_level[depth + 1] = total;
        __syncthreads();
    }
    const int n_levels = depth;  // levels 0..n_levels-1
    // subtree sizes, deepest level first
    for (int d = n_levels - 1; d >= 0 && !overflow; --d) {
        for (int v = s_level[d] + tid; v < s_level[d + 1]; v += kTopoThreads) {
            const int l = b_left[v];
            b_size[v] = (l < 0) ? 1 : 1 + b_size[l] + b_size[l + 1];
        }
        __syncthreads();
    }
    // pre-order numbers, root first
    if (tid == 0) b_pre[0] = 0;
    __syncthreads();
    for (int d = 0; d < n_levels && !overflow; ++d) {
        for (int v = s_level[d] + tid; v < s_level[d + 1]; v += kTopoThreads) {
            const int l = b_left[v];
            if (l >= 0) {
                b_pre[l] = b_pre[v] + 1;
                b_pre[l + 1] = b_pre[v] + 1 + b_size[l];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        info[0] = total;
        info[1] = n_levels;
        info[2] = overflow;
    }
    // level offsets for the merge pass
    for (int d = tid; d <= n_levels; d += kTopoThreads) info[4 + d] = s_level[d];
}

struct TreeArrays {
    double *lo, *hi, *acc, *nodes, *w;
    int32_t *lo_idx, *hi_idx, *right, *esc;
};

__global__ void finalize_kernel(const double *__restrict__ sorted, int N, int n, double theta_sq,
                                const int *__restrict__ b_lo, const int *__restrict__ b_hi,
                                const int *__restrict__ b_left, const int *__restrict__ b_size,
                                const int *__restrict__ b_pre, const double *__restrict__ sm1,
                                TreeArrays t) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int k = b_pre[v];
    const int a = b_lo[v], b = b_hi[v];
    const double lo = sorted[a], hi = sorted[b - 1];
    t.lo[k] = lo;
    t.hi[k] = hi;
    const double hw = 0.5 * (hi - lo);
    t.acc[k] = (hw * hw) / theta_sq;
    t.lo_idx[k] = a;
    t.hi_idx[k] = b;
    const int l = b_left[v];
    t.right[k] = (l < 0) ? -1 : b_pre[l + 1];
    t.esc[k] = k + b_size[v];
    const double c = 0.5 * (lo + hi), hl = 0.5 * (hi - lo);
    for (int j = 0; j < n; ++j) {
        t.nodes[(size_t)k * n + j] = c + hl * sm1[j];
        t.w[(size_t)k * n + j] = 0.0;
    }
}

// Barycentric accumulation of `nrows` rows (x_r, mass_r) into one panel's
// weights w[0..n) (chebcode.rs:309-337), one wavefront per panel.
//
// Lanes first evaluate 64 rows at a time: the quotients lam_j/(x_r - t_j)
// go to LDS together with m_r = mass_r * (1/s_r), or the node index hit by
// an exact coincidence.  Then lane j adds the rows into w_j in row order,
// w_j = fma(m_r, q_rj, w_j), which is the reference's per-row update in the
// reference's order, so the weights equal a sequential build bit for bit.
// Rows with zero mass are skipped (merge_all_weights, chebcode.rs:496-499).
// `row(r, x, mass)` supplies row r; w must hold zeros on entry.
constexpr int kBaryLanes = 64;
constexpr int kWaccPerLane = (kMaxN + kBaryLanes - 1) / kBaryLanes;

template <class RowFn>
__device__ void bary_accumulate(RowFn row, int nrows, const double *__restrict__ t,
                                const double *__restrict__ lam, double *__restrict__ w, int n,
                                double *__restrict__ sq, double *__restrict__ sm,
                                int *__restrict__ shit) {
    const int lane = threadIdx.x;
    double wacc[kWaccPerLane];
#pragma unroll
    for (int u = 0; u < kWaccPerLane; ++u) wacc[u] = 0.0;
    for (int base = 0; base < nrows; base += kBaryLanes) {
        const int r = base + lane;
        int hit = -2;  // -2: no row or zero mass; -1: regular row; >= 0: exact hit
        double m = 0.0;
        if (r < nrows) {
            double x, mass;
            row(r, x, mass);
            if (mass != 0.0) {
                double s = 0.0;
                hit = -1;
                for (int j = 0; j < n; ++j) {
                    const double d = x - t[j];
                    if (d == 0.0) {
                        hit = j;
                        break;
                    }
                    const double q = lam[j] / d;
                    sq[lane * n + j] = q;
                    s += q;
                }
                m = (hit >= 0) ? mass : mass * (1.0 / s);
            }
        }
        sm[lane] = m;
        shit[lane] = hit;
        __syncthreads();
        const int rows = min(kBaryLanes, nrows - base);
#pragma unroll
        for (int u = 0; u < kWaccPerLane; ++u) {
            const int j = lane + u * kBaryLanes;
            if (j < n) {
                double acc = wacc[u];
                for (int rr = 0; rr < rows; ++rr) {
                    const int h = shit[rr];
                    if (h == -1) acc = __builtin_fma(sm[rr], sq[rr * n + j], acc);
                    else if (h == j) acc += sm[rr];
                }
                wacc[u] = acc;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kWaccPerLane; ++u) {
        const int j = lane + u * kBaryLanes;
        if (j < n) w[j] = wacc[u];
    }
}

// LDS layout of the two build kernels below: sq[64*n] doubles, sm[64]
// doubles, shit[64] ints.
inline size_t bary_lds_bytes(int n) {
    return sizeof(double) * ((size_t)kBaryLanes * n + kBaryLanes) + sizeof(int) * kBaryLanes;
}

// One wavefront per node; internal nodes exit at once.
__global__ __launch_bounds__(kBaryLanes) void leaf_fill_kernel(const double *__restrict__ sorted,
                                                               int n,
                                                               const double *__restrict__ lam,
                                                               TreeArrays t) {
    extern __shared__ double lds[];
    const int k = blockIdx.x;
    if (t.right[k] >= 0) return;
    double *sq = lds;
    double *sm = lds + kBaryLanes * n;
    int *shit = reinterpret_cast<int *>(sm + kBaryLanes);
    const double *xs = sorted + t.lo_idx[k];
    auto row = [xs](int r, double &x, double &mass) {
        x = xs[r];
        mass = 1.0;
    };
    bary_accumulate(row, t.hi_idx[k] - t.lo_idx[k], t.nodes + (size_t)k * n, lam,
                    t.w + (size_t)k * n, n, sq, sm, shit);
}

// Parents of one tree level (BFS ids [v0, v0 + gridDim.x)), one wavefront
// each.  The 2n rows are the children's Chebyshev nodes with their weights
// as masses, left child first (merge_all_weights, chebcode.rs:474-506).
__global__ __launch_bounds__(kBaryLanes) void merge_level_kernel(int n, int v0,
                                                                 const int *__restrict__ b_left,
                                                                 const int *__restrict__ b_pre,
                                                                 const double *__restrict__ lam,
                                                                 TreeArrays t) {
    extern __shared__ double lds[];
    const int v = v0 + blockIdx.x;
    if (b_left[v] < 0) return;
    const int k = b_pre[v];
    double *sq = lds;
    double *sm = lds + kBaryLanes * n;
    int *shit = reinterpret_cast<int *>(sm + kBaryLanes);
    const size_t lb = (size_t)(k + 1) * n, rb = (size_t)t.right[k] * n;
    const double *nodes = t.nodes;
    const double *w = t.w;
    auto row = [=](int r, double &x, double &mass) {
        const size_t o = (r < n) ? lb + r : rb + (r - n);
        x = nodes[o];
        mass = w[o];
    };
    bary_accumulate(row, 2 * n, t.nodes + (size_t)k * n, lam, t.w + (size_t)k * n, n, sq, sm,
                    shit);
}

__global__ void mirror_kernel(int N, int n, TreeArrays t, float4 *__restrict__ rec32,
                              int4 *__restrict__ link, float *__restrict__ panel32) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    rec32[k] = make_float4((float)t.lo[k], (float)t.hi[k], (float)t.acc[k], 0.0f);
    link[k] = make_int4(t.right[k], t.esc[k], t.lo_idx[k], t.hi_idx[k]);
    for (int j = 0; j < n; ++j) {
        panel32[(size_t)k * 2 * n + j] = (float)t.nodes[(size_t)k * n + j];
        panel32[(size_t)k * 2 * n + n + j] = (float)t.w[(size_t)k * n + j];
    }
}

__global__ void split_kernel(const double *__restrict__ v, int p, float *__restrict__ hi,
                             float *__restrict__ lo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p) return;
    const float h = (float)v[i];
    hi[i] = h;
    lo[i] = (float)(v[i] - (double)h);
}

// ---- evaluation --------------------------------------------------------------

struct EvalView {
    const double *sorted;
    const float *s_hi, *s_lo;
    const double *lo, *hi, *acc, *nodes, *w;
    const int4 *link;
    const float4 *rec32;
    const float *panel32;
    const int32_t *perm;
    int N, n;
};

// Exact f64 leaf sum in the reference's two-lane form (chebcode.rs:622-647).
__device__ __forceinline__ void leaf_f64(const double *__restrict__ xs, int len, double x,
                                         double eta2, double &lr, double &li) {
    double ar0 = 0.0, ar1 = 0.0, ai0 = 0.0, ai1 = 0.0;
    int k = 0;
    // even sources to (ar0, ai0), odd to (ar1, ai1), eight per step
    for (; k + 8 <= len; k += 8) {
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            const double d0 = x - xs[k + u], d1 = x - xs[k + u + 1];
            const double i0 = rcp_f64(__builtin_fma(d0, d0, eta2));
            const double i1 = rcp_f64(__builtin_fma(d1, d1, eta2));
            ar0 = __builtin_fma(d0, i0, ar0);
            ar1 = __builtin_fma(d1, i1, ar1);
            ai0 += i0;
            ai1 += i1;
        }
    }
    for (; k + 2 <= len; k += 2) {
        const double d0 = x - xs[k], d1 = x - xs[k + 1];
        const double i0 = rcp_f64(__builtin_fma(d0, d0, eta2));
        const double i1 = rcp_f64(__builtin_fma(d1, d1, eta2));
        ar0 = __builtin_fma(d0, i0, ar0);
        ar1 = __builtin_fma(d1, i1, ar1);
        ai0 += i0;
        ai1 += i1;
    }
    lr = ar0 + ar1;
    li = ai0 + ai1;
    if (k < len) {
        const double d = x - xs[k];
        const double inv = rcp_f64(d * d + eta2);
        lr += d * inv;
        li += inv;
    }
}

// Far-field acceptance of an internal node for query x: the f64 test of
// chebcode.rs:648-657 (accept_hw < cl^2 + eta^2, cl the clamped distance) and
// the FAST f32 test with the OUTSIDE_THETA2 factor for panels that do not
// contain the query (chebcode.rs:776-783).  Shared by the evaluation and the
// work-counting kernels so both make the same decisions.
__device__ __forceinline__ bool accept_f64(const EvalView &v, int node, double x, double eta2) {
    const double lo = v.lo[node], hi = v.hi[node];
    const double cl = fmax(fmax(lo - x, x - hi), 0.0);
    const double d_sq = __builtin_fma(cl, cl, eta2);
    return v.acc[node] < d_sq;
}

__device__ __forceinline__ bool accept_fast(const float4 r, float x, float eta2) {
    // r = (lo, hi, accept_hw)
    const float cl = fmaxf(fmaxf(r.x - x, x - r.y), 0.0f);
    const float d_sq = __builtin_fmaf(cl, cl, eta2);
    const float outside = (cl == 0.0f) ? d_sq : kOutsideTheta2 * d_sq;
    return r.z < outside;
}

// Traversal, one query per lane.  Every lane follows the reference's stack
// DFS in pre-order (a pruned or finished node continues at its escape index,
// a descended node at node + 1).  Two walks give bitwise the same per-lane
// results:
//
// kWave = true: the wavefront walks together with one node pointer.  Lane L
// is active at node u unless it accepted an ancestor of u, recorded as
// skip_L = escape(ancestor) > u.  At an internal node the active lanes that
// accept evaluate the panel and set skip; the wavefront descends if any
// active lane rejects, and otherwise jumps to the escape index (every
// inactive lane's skip is at least that far).  Each lane processes exactly
// the nodes of its own traversal, in the same order; node records, panels and
// leaf sources are read at wavefront-uniform addresses.  This suits dense
// sorted queries, whose neighbours share their near fields.
//
// kWave = false: each lane walks on its own (sparse queries, such as a grid
// spanning the spectrum, whose near fields differ).
template <bool kWave>
__global__ __launch_bounds__(256) void eval_f64_kernel(EvalView v, const double *__restrict__ query,
                                                       int nq, int queries_are_sources, double eta,
                                                       double scale, double *__restrict__ out_re,
                                                       double *__restrict__ out_im) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nq;
    if (!kWave && !valid) return;
    const double x = valid ? (queries_are_sources ? v.sorted[i] : query[i]) : 0.0;
    const double eta2 = eta * eta;
    const int n = v.n;
    double re = 0.0, im = 0.0;
    int skip = valid ? 0 : v.N;
    int node = 0;
    while (node < v.N) {
        if constexpr (kWave) node = __builtin_amdgcn_readfirstlane(node);
        const int4 lk = v.link[node];  // right, esc, lo_idx, hi_idx
        const bool active = !kWave || node >= skip;
        if (lk.x < 0) {
            if (active) {
                double lr, li;
                leaf_f64(v.sorted + lk.z, lk.w - lk.z, x, eta2, lr, li);
                re += lr;
                im += eta * li;
            }
            node = lk.y;
            continue;
        }
        const bool accept = accept_f64(v, node, x, eta2);
        if (active && accept) {
            const double *ts = v.nodes + (size_t)node * n;
            const double *ws = v.w + (size_t)node * n;
            double afr0 = 0.0, afr1 = 0.0, afi0 = 0.0, afi1 = 0.0;
            int j = 0;
            for (; j + 2 <= n; j += 2) {
                const double d0 = x - ts[j], d1 = x - ts[j + 1];
                const double q0 = ws[j] * rcp_f64(__builtin_fma(d0, d0, eta2));
                const double q1 = ws[j + 1] * rcp_f64(__builtin_fma(d1, d1, eta2));
                afr0 = __builtin_fma(q0, d0, afr0);
                afr1 = __builtin_fma(q1, d1, afr1);
                afi0 += q0;
                afi1 += q1;
            }
            double fr = afr0 + afr1, fi = afi0 + afi1;
            if (j < n) {
                const double dj = x - ts[j];
                const double q = ws[j] * rcp_f64(dj * dj + eta2);
                fr += q * dj;
                fi += q;
            }
            re += fr;
            im += eta * fi;
            skip = lk.y;
        }
        if constexpr (kWave)
            node = __any(active && !accept) ? node + 1 : lk.y;
        else
            node = accept ? lk.y : node + 1;
    }
    if (!valid) return;
    const int o = (queries_are_sources && v.perm) ? v.perm[i] : i;
    out_re[o] = re * scale;
    out_im[o] = im * scale;
}

template <bool kWave>
__global__ __launch_bounds__(256) void eval_fast_kernel(EvalView v, const double *__restrict__ query,
                                                        int nq, int queries_are_sources, double eta,
                                                        double scale, double *__restrict__ out_re,
                                                        double *__restrict__ out_im) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nq;
    if (!kWave && !valid) return;
    const double xd = valid ? (queries_are_sources ? v.sorted[i] : query[i]) : 0.0;
    const float x = (float)xd;
    const float x_lo = (float)(xd - (double)x);
    const float e32 = (float)eta;
    const float eta2 = e32 * e32;
    const float eta2n = (float)(eta * eta);
    const int n = v.n;
    float far_re[4] = {0.f, 0.f, 0.f, 0.f}, far_im[4] = {0.f, 0.f, 0.f, 0.f};
    double leaf_re = 0.0, leaf_inv = 0.0;
    int skip = valid ? 0 : v.N;
    int node = 0;
    while (node < v.N) {
        if constexpr (kWave) node = __builtin_amdgcn_readfirstlane(node);
        const int4 lk = v.link[node];
        const bool active = !kWave || node >= skip;
        if (lk.x < 0) {
            if (active) {
                float ar0 = 0.f, ar1 = 0.f, ai0 = 0.f, ai1 = 0.f;
                int k = lk.z;
                // even sources to (ar0, ai0), odd to (ar1, ai1), eight per step
                for (; k + 8 <= lk.w; k += 8) {
#pragma unroll
                    for (int u = 0; u < 8; u += 2) {
                        const float d0 = (x - v.s_hi[k + u]) + (x_lo - v.s_lo[k + u]);
                        const float d1 = (x - v.s_hi[k + u + 1]) + (x_lo - v.s_lo[k + u + 1]);
                        const float r0 = __builtin_amdgcn_rcpf(__builtin_fmaf(d0, d0, eta2n));
                        const float r1 = __builtin_amdgcn_rcpf(__builtin_fmaf(d1, d1, eta2n));
                        ar0 = __builtin_fmaf(d0, r0, ar0);
                        ar1 = __builtin_fmaf(d1, r1, ar1);
                        ai0 += r0;
                        ai1 += r1;
                    }
                }
                for (; k + 2 <= lk.w; k += 2) {
                    const float d0 = (x - v.s_hi[k]) + (x_lo - v.s_lo[k]);
                    const float d1 = (x - v.s_hi[k + 1]) + (x_lo - v.s_lo[k + 1]);
                    const float r0 = __builtin_amdgcn_rcpf(__builtin_fmaf(d0, d0, eta2n));
                    const float r1 = __builtin_amdgcn_rcpf(__builtin_fmaf(d1, d1, eta2n));
                    ar0 = __builtin_fmaf(d0, r0, ar0);
                    ar1 = __builtin_fmaf(d1, r1, ar1);
                    ai0 += r0;
                    ai1 += r1;
                }
                if (k < lk.w) {
                    const float d0 = (x - v.s_hi[k]) + (x_lo - v.s_lo[k]);
                    const float r0 = __builtin_amdgcn_rcpf(__builtin_fmaf(d0, d0, eta2n));
                    ar0 = __builtin_fmaf(d0, r0, ar0);
                    ai0 += r0;
                }
                leaf_re += (double)(ar0 + ar1);
                leaf_inv += (double)(ai0 + ai1);
            }
            node = lk.y;
            continue;
        }
        const bool accept = accept_fast(v.rec32[node], x, eta2);
        if (active && accept) {
            const float *ts = v.panel32 + (size_t)node * 2 * n;
            const float *ws = ts + n;
            // term j goes to accumulator j mod 4 (static indices)
            int j = 0;
            for (; j + 4 <= n; j += 4) {
#pragma unroll
                for (int l = 0; l < 4; ++l) {
                    const float d = x - ts[j + l];
                    const float q = ws[j + l] * __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2));
                    far_re[l] = __builtin_fmaf(q, d, far_re[l]);
                    far_im[l] = far_im[l] + q;
                }
            }
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                if (j + l < n) {
                    const float d = x - ts[j + l];
                    const float q = ws[j + l] * __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2));
                    far_re[l] = __builtin_fmaf(q, d, far_re[l]);
                    far_im[l] = far_im[l] + q;
                }
            }
            skip = lk.y;
        }
        if constexpr (kWave)
            node = __any(active && !accept) ? node + 1 : lk.y;
        else
            node = accept ? lk.y : node + 1;
    }
    if (!valid) return;
    const float hr = (far_re[0] + far_re[1]) + (far_re[2] + far_re[3]);
    const float hi = (far_im[0] + far_im[1]) + (far_im[2] + far_im[3]);
    const double re = leaf_re + (double)hr;
    const double im = eta * leaf_inv + eta * (double)hi;
    const int o = (queries_are_sources && v.perm) ? v.perm[i] : i;
    out_re[o] = re * scale;
    out_im[o] = im * scale;
}

// Work of the traversal, without arithmetic: near-field pair-evaluations
// (sum of the sizes of the leaves a query visits) and far-field panel terms
// (n per accepted panel), with the evaluation kernels' decisions; one query
// per lane, sums reduced per wavefront and added to counts[0..1].
template <bool kFast>
__global__ __launch_bounds__(256) void count_kernel(EvalView v, const double *__restrict__ query,
                                                    int nq, int queries_are_sources, double eta,
                                                    unsigned long long *__restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long near = 0, far = 0;
    if (i < nq) {
        const double xd = queries_are_sources ? v.sorted[i] : query[i];
        const float xf = (float)xd;
        const float e32 = (float)eta;
        const float eta2f = e32 * e32;
        const double eta2 = eta * eta;
        int node = 0;
        while (node < v.N) {
            const int4 lk = v.link[node];
            if (lk.x < 0) {
                near += (unsigned long long)(lk.w - lk.z);
                node = lk.y;
                continue;
            }
            const bool accept =
                kFast ? accept_fast(v.rec32[node], xf, eta2f) : accept_f64(v, node, xd, eta2);
            if (accept) {
                far += (unsigned long long)v.n;
                node = lk.y;
            } else {
                node = node + 1;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        near += __shfl_xor(near, o, 64);
        far += __shfl_xor(far, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counts[0], near);
        atomicAdd(&counts[1], far);
    }
}

__global__ void clear_counts_kernel(unsigned long long *counts) {
    if (threadIdx.x < 2) counts[threadIdx.x] = 0ull;
}

// Two-eta traversal (chebcode.rs:836-915); queries are the tree's sources.
__global__ __launch_bounds__(256) void eval_x2_kernel(EvalView v, int nq, double e0, double e1,
                                                      double theta_sq, double *__restrict__ re0,
                                                      double *__restrict__ im0,
                                                      double *__restrict__ re1,
                                                      double *__restrict__ im1) {
    // Queries are the tree's sorted sources (dense): the wavefront walks
    // together as in eval_f64_kernel<true>; each lane's visits and arithmetic
    // are those of its own traversal.
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nq;
    const double x = valid ? v.sorted[i] : 0.0;
    const double e[2] = {e0, e1}, e2[2] = {e0 * e0, e1 * e1};
    const double emin = e0 < e1 ? e0 : e1;
    double ar[2] = {0.0, 0.0}, ai[2] = {0.0, 0.0};
    const int n = v.n;
    int skip = valid ? 0 : v.N;
    int node = 0;
    while (node < v.N) {
        node = __builtin_amdgcn_readfirstlane(node);
        const int4 lk = v.link[node];
        const bool active = node >= skip;
        if (lk.x < 0) {
            if (active) {
                for (int k = lk.z; k < lk.w; ++k) {
                    const double d = x - v.sorted[k];
#pragma unroll
                    for (int l = 0; l < 2; ++l) {
                        const double inv = rcp_f64(__builtin_fma(d, d, e2[l]));
                        ar[l] = __builtin_fma(d, inv, ar[l]);
                        ai[l] = __builtin_fma(e[l], inv, ai[l]);
                    }
                }
            }
            node = lk.y;
            continue;
        }
        const double lo = v.lo[node], hi = v.hi[node];
        const double cl = x < lo ? lo - x : (x > hi ? x - hi : 0.0);
        const double d_sq = cl * cl + emin * emin;
        const double half_w = 0.5 * (hi - lo);
        const bool accept = half_w * half_w < theta_sq * d_sq;
        if (active && accept) {
            const double *ts = v.nodes + (size_t)node * n;
            const double *ws = v.w + (size_t)node * n;
            for (int j = 0; j < n; ++j) {
                const double dj = x - ts[j], wj = ws[j];
#pragma unroll
                for (int l = 0; l < 2; ++l) {
                    const double inv = rcp_f64(__builtin_fma(dj, dj, e2[l]));
                    ar[l] = __builtin_fma(wj * dj, inv, ar[l]);
                    ai[l] = __builtin_fma(wj * e[l], inv, ai[l]);
                }
            }
            skip = lk.y;
        }
        node = __any(active && !accept) ? node + 1 : lk.y;
    }
    if (!valid) return;
    const int o = v.perm ? v.perm[i] : i;
    re0[o] = ar[0];
    im0[o] = ai[0];
    re1[o] = ar[1];
    im1[o] = ai[1];
}

EvalView view_of(const ChebTree *t) {
    EvalView v;
    v.sorted = t->sorted;
    v.s_hi = t->s_hi;
    v.s_lo = t->s_lo;
    v.lo = t->lo;
    v.hi = t->hi;
    v.acc = t->acc;
    v.nodes = t->nodes;
    v.w = t->w;
    v.link = t->link;
    v.rec32 = t->rec32;
    v.panel32 = t->panel32;
    v.perm = t->perm;
    v.N = (int)t->N;
    v.n = t->n;
    return v;
}

inline unsigned blocks_for(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void cheb_destroy(ChebTree *t) {
    if (!t) return;
    if (t->arena || t->arena2) {
        // Stream-ordered pool memory: wait until no stream can still read the
        // tree (the synchronisation hipFree would do), then return it.
        (void)hipDeviceSynchronize();
        if (t->arena) (void)hipFreeAsync(t->arena, nullptr);
        if (t->arena2) (void)hipFreeAsync(t->arena2, nullptr);
    }
    if (t->ready) (void)hipEventDestroy(t->ready);
    delete t;
}

namespace {

// Carves 256-byte aligned sub-buffers out of one allocation.
struct Carver {
    size_t off = 0;
    char *base = nullptr;
    template <class T>
    T *take(size_t count) {
        off = (off + 255) / 256 * 256;
        T *p = reinterpret_cast<T *>(base ? base + off : nullptr);
        off += sizeof(T) * count;
        return p;
    }
};

// Stream-ordered scratch freed when it goes out of scope.
struct Scratch {
    void *p = nullptr;
    hipStream_t s;
    explicit Scratch(hipStream_t st) : s(st) {}
    ~Scratch() {
        if (p) (void)hipFreeAsync(p, s);
    }
};

}  // namespace

int cheb_build(const double *d_ev, size_t p, const ChebParams &prm, ChebTree **out,
               hipStream_t stream) {
    *out = nullptr;
    if (p == 0 || p > (size_t)INT32_MAX / 2) {
        set_error("chebcode build: p must be in [1, 2^30)");
        return 6;
    }
    if (prm.n < 2 || prm.n > kMaxN) {
        set_error("chebcode build: n must be in [2, 120]");
        return 6;
    }
    if (!(prm.theta > 0.0) || prm.leaf_cap < 1) {
        set_error("chebcode build: theta must be > 0 and leaf_cap >= 1");
        return 6;
    }
    std::unique_ptr<ChebTree, void (*)(ChebTree *)> t(new ChebTree(), &cheb_destroy);
    t->p = p;
    t->n = prm.n;
    t->theta = prm.theta;
    t->theta_sq = prm.theta * prm.theta;
    t->leaf_cap = prm.leaf_cap;
    t->f32_far = prm.f32_far;
    const int ip = (int)p;
    const int n = prm.n;
    const int leaf_cap = (int)std::min<int64_t>(prm.leaf_cap, INT32_MAX);
    const size_t cap_a = 2 * p - 1;
    const size_t cap_b = 2 * (size_t)(kMaxDepth + 1) * ((p + leaf_cap) / (leaf_cap + 1)) + 1;
    const int cap = (int)std::min(cap_a, cap_b);

    // Scratch: info words, BFS arrays, node/weight tables, sort temporaries.
    const int n_info = 4 + kMaxDepth + 3;
    Scratch scratch(stream);
    Carver sc;
    (void)sc.take<int>(n_info);
    (void)sc.take<int>(5 * (size_t)cap);
    (void)sc.take<double>(2 * (size_t)n);
    (void)sc.take<int32_t>(p);
    size_t sort_bytes = 0;
    SHK_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, d_ev, (double *)nullptr,
                                      (const int32_t *)nullptr, (int32_t *)nullptr, p, 0, 64,
                                      stream));
    (void)sc.take<char>(sort_bytes);
    SHK_HIP(hipMallocAsync(&scratch.p, sc.off, stream));
    sc.base = static_cast<char *>(scratch.p);
    sc.off = 0;
    int *d_info = sc.take<int>(n_info);
    int *b_buf = sc.take<int>(5 * (size_t)cap);
    double *d_tab = sc.take<double>(2 * (size_t)n);
    int32_t *d_iota = sc.take<int32_t>(p);
    void *d_sort_tmp = sc.take<char>(sort_bytes);
    int *b_lo = b_buf, *b_hi = b_buf + cap, *b_left = b_buf + 2 * (size_t)cap,
        *b_size = b_buf + 3 * (size_t)cap, *b_pre = b_buf + 4 * (size_t)cap;

    // Arena part 1 (sizes known before the topology): sources.
    Carver ar;
    (void)ar.take<double>(p);
    (void)ar.take<int32_t>(p);
    (void)ar.take<float>(p);
    (void)ar.take<float>(p);
    const size_t src_bytes = ar.off;

    // 1. sortedness, then the sorted copy (+ permutation)
    std::vector<int> info(n_info, 0);
    // The topology needs the sorted values; its node count sizes the rest of
    // the arena, so the arena is allocated after a first pass with the
    // sources in scratch-independent storage: allocate the source part now.
    void *src_arena = nullptr;
    SHK_HIP(hipMallocAsync(&src_arena, src_bytes, stream));
    ar.base = static_cast<char *>(src_arena);
    ar.off = 0;
    t->sorted = ar.take<double>(p);
    int32_t *perm_buf = ar.take<int32_t>(p);
    t->s_hi = ar.take<float>(p);
    t->s_lo = ar.take<float>(p);
    t->arena = src_arena;  // owned from here on (freed by cheb_destroy)

    int zero4[4] = {0, 0, 0, 0};
    SHK_HIP(hipMemcpyAsync(d_info, zero4, sizeof(zero4), hipMemcpyHostToDevice, stream));
    if (p > 1) {
        hipLaunchKernelGGL(check_sorted_kernel, dim3(blocks_for(p, 256)), dim3(256), 0, stream,
                           d_ev, ip, d_info + 3);
        SHK_HIP(hipGetLastError());
    }
    SHK_HIP(hipMemcpyAsync(info.data(), d_info, sizeof(int) * 4, hipMemcpyDeviceToHost, stream));
    SHK_HIP(hipStreamSynchronize(stream));
    if (info[3] == 0) {
        SHK_HIP(hipMemcpyAsync(t->sorted, d_ev, sizeof(double) * p, hipMemcpyDeviceToDevice,
                               stream));
    } else {
        t->perm = perm_buf;
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(p, 256)), dim3(256), 0, stream, d_iota, ip);
        SHK_HIP(hipGetLastError());
        SHK_HIP(rocprim::radix_sort_pairs(d_sort_tmp, sort_bytes, d_ev, t->sorted, d_iota, t->perm,
                                          p, 0, 64, stream));
    }

    // 2. topology
    hipLaunchKernelGGL(topo_kernel, dim3(1), dim3(kTopoThreads), 0, stream, t->sorted, ip, leaf_cap,
                       cap, b_lo, b_hi, b_left, b_size, b_pre, d_info);
    SHK_HIP(hipGetLastError());
    SHK_HIP(hipMemcpyAsync(info.data(), d_info, sizeof(int) * n_info, hipMemcpyDeviceToHost,
                           stream));
    SHK_HIP(hipStreamSynchronize(stream));
    if (info[2]) {
        set_error("chebcode build: node capacity exceeded");
        return 6;
    }
    const int N = info[0];
    t->N = (size_t)N;
    t->n_levels = info[1];

    // Arena part 2: node records and panels.
    Carver nr;
    (void)nr.take<double>(N);
    (void)nr.take<double>(N);
    (void)nr.take<double>(N);
    for (int k = 0; k < 4; ++k) (void)nr.take<int32_t>(N);
    (void)nr.take<double>((size_t)N * n);
    (void)nr.take<double>((size_t)N * n);
    (void)nr.take<float4>(N);
    (void)nr.take<int4>(N);
    (void)nr.take<float>((size_t)N * 2 * n);
    void *node_arena = nullptr;
    SHK_HIP(hipMallocAsync(&node_arena, nr.off, stream));
    t->arena2 = node_arena;
    nr.base = static_cast<char *>(node_arena);
    nr.off = 0;
    t->lo = nr.take<double>(N);
    t->hi = nr.take<double>(N);
    t->acc = nr.take<double>(N);
    t->lo_idx = nr.take<int32_t>(N);
    t->hi_idx = nr.take<int32_t>(N);
    t->right = nr.take<int32_t>(N);
    t->esc = nr.take<int32_t>(N);
    t->nodes = nr.take<double>((size_t)N * n);
    t->w = nr.take<double>((size_t)N * n);
    t->rec32 = nr.take<float4>(N);
    t->link = nr.take<int4>(N);
    t->panel32 = nr.take<float>((size_t)N * 2 * n);

    // 3. pre-order arrays and positions.  Chebyshev nodes of the second kind
    // s_j = cos(j*pi/(n-1)) and barycentric weights (-1)^j halved at both
    // ends (chebcode.rs:242-261).
    std::vector<double> tab(2 * n);
    for (int j = 0; j < n; ++j) {
        tab[j] = std::cos((double)j * kPi / (double)(n - 1));
        const double sign = (j % 2 == 0) ? 1.0 : -1.0;
        tab[n + j] = (j == 0 || j == n - 1) ? 0.5 * sign : sign;
    }
    SHK_HIP(hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice,
                           stream));
    TreeArrays ta{t->lo, t->hi, t->acc, t->nodes, t->w, t->lo_idx, t->hi_idx, t->right, t->esc};
    hipLaunchKernelGGL(finalize_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, stream, t->sorted, N,
                       n, t->theta_sq, b_lo, b_hi, b_left, b_size, b_pre, d_tab, ta);
    SHK_HIP(hipGetLastError());
    // 4. leaves: one wavefront per node
    const size_t lds = bary_lds_bytes(n);
    hipLaunchKernelGGL(leaf_fill_kernel, dim3(N), dim3(kBaryLanes), lds, stream, t->sorted, n,
                       d_tab + n, ta);
    SHK_HIP(hipGetLastError());
    // 5. parents, deepest level first: one launch per level, one wavefront per node
    for (int d = t->n_levels - 2; d >= 0; --d) {
        const int v0 = info[4 + d], v1 = info[4 + d + 1];
        if (v1 <= v0) continue;
        hipLaunchKernelGGL(merge_level_kernel, dim3(v1 - v0), dim3(kBaryLanes), lds, stream, n, v0,
                           b_left, b_pre, d_tab + n, ta);
        SHK_HIP(hipGetLastError());
    }
    // 6. mirrors
    hipLaunchKernelGGL(mirror_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, stream, N, n, ta,
                       t->rec32, t->link, t->panel32);
    SHK_HIP(hipGetLastError());
    hipLaunchKernelGGL(split_kernel, dim3(blocks_for(p, 256)), dim3(256), 0, stream, t->sorted, ip,
                       t->s_hi, t->s_lo);
    SHK_HIP(hipGetLastError());
    // Evaluations on any stream wait for this event (cheb_eval*).
    SHK_HIP(hipEventCreateWithFlags(&t->ready, hipEventDisableTiming));
    SHK_HIP(hipEventRecord(t->ready, stream));
    *out = t.release();
    return 0;
}

size_t cheb_node_count(const ChebTree *t) { return t ? t->N : 0; }
size_t cheb_p(const ChebTree *t) { return t ? t->p : 0; }

int cheb_params(const ChebTree *t, ChebParams *prm) {
    prm->theta = t->theta;
    prm->n = t->n;
    prm->leaf_cap = t->leaf_cap;
    prm->f32_far = t->f32_far;
    return 0;
}

int cheb_export(const ChebTree *t, double *lo, double *hi, int64_t *lo_idx, int64_t *hi_idx,
                int32_t *right, double *nodes, double *w, hipStream_t stream) {
    const size_t N = t->N, n = t->n;
    SHK_HIP(hipStreamWaitEvent(stream, t->ready, 0));
    if (lo) SHK_HIP(hipMemcpyAsync(lo, t->lo, sizeof(double) * N, hipMemcpyDeviceToHost, stream));
    if (hi) SHK_HIP(hipMemcpyAsync(hi, t->hi, sizeof(double) * N, hipMemcpyDeviceToHost, stream));
    if (right)
        SHK_HIP(hipMemcpyAsync(right, t->right, sizeof(int32_t) * N, hipMemcpyDeviceToHost, stream));
    if (nodes)
        SHK_HIP(hipMemcpyAsync(nodes, t->nodes, sizeof(double) * N * n, hipMemcpyDeviceToHost,
                               stream));
    if (w) SHK_HIP(hipMemcpyAsync(w, t->w, sizeof(double) * N * n, hipMemcpyDeviceToHost, stream));
    if (lo_idx || hi_idx) {
        std::vector<int32_t> a(N), b(N);
        SHK_HIP(hipMemcpyAsync(a.data(), t->lo_idx, sizeof(int32_t) * N, hipMemcpyDeviceToHost,
                               stream));
        SHK_HIP(hipMemcpyAsync(b.data(), t->hi_idx, sizeof(int32_t) * N, hipMemcpyDeviceToHost,
                               stream));
        SHK_HIP(hipStreamSynchronize(stream));
        for (size_t k = 0; k < N; ++k) {
            if (lo_idx) lo_idx[k] = a[k];
            if (hi_idx) hi_idx[k] = b[k];
        }
    }
    SHK_HIP(hipStreamSynchronize(stream));
    return 0;
}

int cheb_eval(const ChebTree *t, const double *d_query, size_t nq, double eta, double scale,
              int queries_are_sources, double *d_out_re, double *d_out_im, hipStream_t stream) {
    if (nq == 0) return 0;
    if (nq > (size_t)INT32_MAX) {
        set_error("chebcode eval: too many query points");
        return 6;
    }
    SHK_HIP(hipStreamWaitEvent(stream, t->ready, 0));
    EvalView v = view_of(t);
    // Dense queries (the tree's own sources, or at least as many points as
    // sources) share near fields within a wavefront: walk together.
    const bool wave = queries_are_sources || nq >= t->p;
    const dim3 grid(blocks_for(nq, 256)), block(256);
    ProfileMark m = profile_begin(stream);
    if (t->f32_far && wave)
        hipLaunchKernelGGL(eval_fast_kernel<true>, grid, block, 0, stream, v, d_query, (int)nq,
                           queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (t->f32_far)
        hipLaunchKernelGGL(eval_fast_kernel<false>, grid, block, 0, stream, v, d_query, (int)nq,
                           queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (wave)
        hipLaunchKernelGGL(eval_f64_kernel<true>, grid, block, 0, stream, v, d_query, (int)nq,
                           queries_are_sources, eta, scale, d_out_re, d_out_im);
    else
        hipLaunchKernelGGL(eval_f64_kernel<false>, grid, block, 0, stream, v, d_query, (int)nq,
                           queries_are_sources, eta, scale, d_out_re, d_out_im);
    profile_end(m, stream);
    SHK_HIP(hipGetLastError());
    return 0;
}

int cheb_count(const ChebTree *t, const double *d_query, size_t nq, double eta,
               int64_t *near_pairs, int64_t *far_terms, hipStream_t stream) {
    const int queries_are_sources = d_query == nullptr;
    if (queries_are_sources) nq = t->p;
    *near_pairs = *far_terms = 0;
    if (nq == 0) return 0;
    if (nq > (size_t)INT32_MAX) {
        set_error("chebcode count: too many query points");
        return 6;
    }
    SHK_HIP(hipStreamWaitEvent(stream, t->ready, 0));
    unsigned long long *d_counts = nullptr;
    SHK_HIP(hipMallocAsync((void **)&d_counts, 2 * sizeof(unsigned long long), stream));
    EvalView v = view_of(t);
    hipLaunchKernelGGL(clear_counts_kernel, dim3(1), dim3(64), 0, stream, d_counts);
    if (t->f32_far)
        hipLaunchKernelGGL(count_kernel<true>, dim3(blocks_for(nq, 256)), dim3(256), 0, stream, v,
                           d_query, (int)nq, queries_are_sources, eta, d_counts);
    else
        hipLaunchKernelGGL(count_kernel<false>, dim3(blocks_for(nq, 256)), dim3(256), 0, stream, v,
                           d_query, (int)nq, queries_are_sources, eta, d_counts);
    hipError_t e = hipGetLastError();
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, d_counts, sizeof(h), hipMemcpyDeviceToHost, stream);
    const hipError_t fe = hipFreeAsync(d_counts, stream);
    if (e == hipSuccess) e = fe;
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_error(std::string("chebcode count: ") + hipGetErrorString(e));
        return 7;
    }
    *near_pairs = (int64_t)h[0];
    *far_terms = (int64_t)h[1];
    return 0;
}

int cheb_eval_x2(const ChebTree *t, double eta0, double eta1, double *d_re0, double *d_im0,
                 double *d_re1, double *d_im1, hipStream_t stream) {
    SHK_HIP(hipStreamWaitEvent(stream, t->ready, 0));
    EvalView v = view_of(t);
    const size_t nq = t->p;
    hipLaunchKernelGGL(eval_x2_kernel, dim3(blocks_for(nq, 256)), dim3(256), 0, stream, v, (int)nq,
                       eta0, eta1, t->theta_sq, d_re0, d_im0, d_re1, d_im1);
    SHK_HIP(hipGetLastError());
    return 0;
}

}  // namespace shk
