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
//   2. topology: one workgroup walks the top of the tree breadth-first until
//      a level holds 512 nodes, then one workgroup per node of that level
//      builds its subtree; a partition search narrowed by a value-bucket
//      table finds each split; subtree sizes and pre-order numbers follow,
//      with per-depth node lists for the merge;
//   3. node records, one wavefront per node: pre-order arrays, Chebyshev
//      positions (:269-279), the f32 mirrors for the FAST traversal
//      (:551-570), and for a leaf its weights, barycentric rows in source
//      order (:286-337, :401-460);
//   4. merge: parents from their children's weights, deepest depth first
//      (:474-506): one workgroup per frontier subtree merges that subtree's
//      depths, then the top levels (one launch per large level, one
//      workgroup for the small ones);
//   5. the f32 hi/lo splits of the sorted sources for the FAST near field
//      are written with the partition hint.
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
#include "shk_pair.hpp"

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
    // transient trees: arena / arena2 are the thread's cached buffers
    // (CACHE_CHEB_SRC / CACHE_CHEB_NODES), released at destruction on `stream`
    bool cached_src = false, cached_nodes = false;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;  // recorded after the build's last kernel (none if transient)
    // a tree built in a caller's workspace with its sizes on the device
    // (cheb_oneshot_ws): the build's info words, which the walks read
    const int *info_dev = nullptr;
};

namespace {

constexpr int kMaxDepth = 48;    // chebcode.rs:95
constexpr int kTopoThreads = 1024;
constexpr int kMaxN = 120;       // largest interpolation order (LDS of the build)
constexpr float kOutsideTheta2 = 1.69f;  // chebcode.rs:108
constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI

__global__ void iota_kernel(int32_t *__restrict__ v, int p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p) v[i] = i;
}

// Workgroup barrier that orders LDS only: it does not wait for this
// workgroup's global stores (s_waitcnt lgkmcnt(0); s_barrier).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Makes a wavefront's LDS writes visible to its own lanes' reads (for LDS
// that only this wavefront uses at the time).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup exclusive scan of one int per thread (THREADS threads):
// wavefront scans by shuffles, then the wavefront totals through LDS.
// LDS_ONLY: the barriers order LDS only (callers with global stores in
// flight that no other thread reads before the kernel ends).
template <int THREADS, bool LDS_ONLY = false>
__device__ int block_exclusive_scan(int v, int *s_wave, int *total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wid] = x;
    if (LDS_ONLY) lds_barrier();
    else __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w) {
        const int t = s_wave[w];
        before += (w < wid) ? t : 0;
        all += t;
    }
    if (LDS_ONLY) lds_barrier();  // s_wave is reused by the next call
    else __syncthreads();
    *total = all;
    return before + x - v;
}

// Bucket hint for the partition search.  bucket(v) = floor((v - x0) * R /
// (x_last - x0)) clamped to [0, R-1] is monotone in v (rounded subtraction,
// multiplication by a non-negative constant and truncation all are), so for
// k = bucket(mid) every element in a lower bucket is < mid and every element
// in a higher bucket is > mid: upper_bound(mid) lies in [bstart[k],
// bstart[k+1]], where bstart[k] is the first index whose bucket is >= k.
struct Hint {
    double x0, inv;
    int R;
};

__device__ __forceinline__ Hint hint_of(const double *__restrict__ sorted, int p, int R) {
    const double x0 = sorted[0];
    const double range = sorted[p - 1] - x0;
    double inv = (range > 0.0) ? (double)R / range : 0.0;
    if (!(inv <= 1.0e300)) inv = 0.0;  // a range too small to scale
    return {x0, inv, R};
}

__device__ __forceinline__ int hint_bucket(double v, const Hint &h) {
    if (h.inv == 0.0) return 0;  // also a range that overflowed
    const double f = (v - h.x0) * h.inv;
    if (!(f < (double)h.R)) return h.R - 1;
    return (f > 0.0) ? (int)f : 0;
}

// gap table: gap[0] = number of long gaps recorded, then kMaxGaps entries
// int4 {first bucket, last bucket, index of the first source after the gap}
// from gap + 4.
constexpr int kMaxGaps = 64;
constexpr int kInfoUnsortedFlag = 3;  // info[] word set by ingest_kernel
constexpr int kGapInts = 4 + 4 * kMaxGaps;

// Also clears the info words the checks below set (count, overflow flag,
// unsortedness flag).  Thread i clears bstart[4i, 4i + 4) with one 16-byte
// store (bstart is 256-byte aligned), threads past them the last n % 4
// entries: hint_clear_threads(n) threads in all.
__host__ __device__ inline int hint_clear_threads(int n) { return n / 4 + (n & 3) + 4; }
__global__ void hint_clear_kernel(int *__restrict__ bstart, int n, int *__restrict__ gap,
                                  int *__restrict__ info) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n4 = n / 4;
    if (i < n4)
        reinterpret_cast<int4 *>(bstart)[i] = make_int4(-1, -1, -1, -1);
    else if (i < n4 + (n & 3))
        bstart[4 * n4 + (i - n4)] = -1;
    if (i == 0) gap[0] = 0;
    if (i < 4) info[i] = 0;
}

// bstart[0, R]: element i fills the buckets (bucket(x[i-1]), bucket(x[i])]
// with i, the last element also (bucket(x[p-1]), R] with p.  Runs longer
// than kHintGap buckets (gaps in the spectrum) stay -1 and go to the gap
// table instead (first bucket, last bucket, i): a midpoint in a gap splits
// at i, since every source before the gap is below it and every source from
// i on above it.  Gaps past the table's capacity fall back to a search of
// the node's own range.
constexpr int kHintGap = 64;
// Buckets per source.  Measured on MI355X (FAST build + evaluate, ms, at
// p = 5e4 / 2e5 / 1e6): 1 per source 0.271 / 0.295 / 0.439, 2: 0.270 /
// 0.290 / 0.435, 4: 0.268 / 0.290 / 0.437, 8: 0.266 / 0.290 / 0.457.
#ifndef SHK_HINT_PER_SOURCE
#define SHK_HINT_PER_SOURCE 2
#endif
constexpr int kHintPerSource = SHK_HINT_PER_SOURCE;
constexpr int64_t kHintMaxBuckets = int64_t(1) << 26;  // past p = 2^25, one per source

__device__ __forceinline__ void record_gap(int *__restrict__ gap, int lo, int hi, int idx) {
    const int g = atomicAdd(gap, 1);
    if (g < kMaxGaps) reinterpret_cast<int4 *>(gap + 4)[g] = make_int4(lo, hi, idx, 0);
}

// Source i of the hint: its buckets (or its gap), and the f32 hi/lo split
// of x = v[i] for the FAST near field (hi = f32(x), lo = f32(x - hi)); xp is
// v[i - 1] for i > 0.
__device__ __forceinline__ void hint_source(int i, double x, double xp, const Hint &h, int p,
                                            int R, int *__restrict__ bstart,
                                            int *__restrict__ gap, float *__restrict__ s_hi,
                                            float *__restrict__ s_lo) {
    const float xh = (float)x;
    s_hi[i] = xh;
    s_lo[i] = (float)(x - (double)xh);
    const int bi = hint_bucket(x, h);
    const int bp = (i == 0) ? -1 : hint_bucket(xp, h);
    if (bi - bp <= kHintGap)
        for (int k = bp + 1; k <= bi; ++k) bstart[k] = i;
    else
        record_gap(gap, bp + 1, bi, i);
    if (i == p - 1) {
        if (R - bi <= kHintGap)
            for (int k = bi + 1; k <= R; ++k) bstart[k] = p;
        else
            record_gap(gap, bi + 1, R, p);
    }
}

__global__ void hint_scatter_kernel(const double *__restrict__ sorted, int p, int R,
                                    int *__restrict__ bstart, int *__restrict__ gap,
                                    float *__restrict__ s_hi, float *__restrict__ s_lo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p) return;
    const double xp = (i == 0) ? 0.0 : sorted[i - 1];
    hint_source(i, sorted[i], xp, hint_of(sorted, p, R), p, R, bstart, gap, s_hi, s_lo);
}

// Chebyshev nodes and barycentric weights (node_kernel and the merges read
// them from device memory), passed by value.
struct ChebTab {
    double v[2 * kMaxN];
};

// First pass over the input ev, taken to be ascending: the tree's copy of
// the sources, the ascending check (info[kInfoUnsorted]), the hint and the
// hi/lo split, as hint_scatter_kernel does on the copy; threads below 2n
// also store the table.
__global__ void ingest_kernel(const double *__restrict__ ev, int p, int R,
                              double *__restrict__ sorted, int *__restrict__ bstart,
                              int *__restrict__ gap, float *__restrict__ s_hi,
                              float *__restrict__ s_lo, int *__restrict__ info, ChebTab tab,
                              int n_tab, double *__restrict__ d_tab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_tab) d_tab[i] = tab.v[i];
    if (i >= p) return;
    const double x = ev[i];
    sorted[i] = x;
    const double xp = (i == 0) ? 0.0 : ev[i - 1];
    if (i > 0 && !(xp <= x)) atomicOr(info + kInfoUnsortedFlag, 1);
    hint_source(i, x, xp, hint_of(ev, p, R), p, R, bstart, gap, s_hi, s_lo);
}

// ---- topology ---------------------------------------------------------------
//
// The tree is built in node "slots"; the output (pre-order records) does not
// depend on how slots are numbered, only on the split rule, so any schedule
// gives the same tree.  Four kernels:
//   A  one workgroup walks breadth-first from the root until a level holds
//      kFrontier nodes or more (or the tree ends); that level is the
//      frontier.  Top slots are BFS ids [0, T).
//   B  one workgroup per frontier node builds its subtree breadth-first into
//      slots [T + 2a, T + 2b) (a subtree over sources [a, b) has at most
//      2(b - a) - 1 nodes, and the frontier ranges are disjoint), then the
//      subtree sizes bottom-up.
//   C  one workgroup: sizes of the top levels, pre-order numbers of the top
//      levels and the frontier roots, and each subtree's block of the node
//      list (the list names every node slot once: the top slots, then the
//      subtrees' nodes below their roots, subtree by subtree).
//   D  one workgroup per frontier node: pre-order numbers inside its subtree
//      and its nodes' entries of the node list.
// Every split is the reference's (midpoint value split with index-midpoint
// fallback, leaf if count <= leaf_cap or depth >= 48, chebcode.rs:347-388).

constexpr int kTopoK = 4;        // nodes per thread per pass (interleaved searches)
constexpr int kSubThreads = 256; // workgroup of the subtree kernels
constexpr int kSubCap = 512;     // subtree level held in LDS (2 nodes per thread)
constexpr int kSubStaged = 5120; // sources of a subtree staged in LDS (its searches run there)
constexpr int kSubStageMin = 256; // smaller subtrees (a level or two) search the global array
constexpr int kFrontier = 512;   // top walk stops at a level this large
constexpr int kMaxFrontier = 2 * kFrontier;  // the frontier level is smaller
constexpr int kLv = kMaxDepth + 2;           // levels per subtree (root level 0)

// info[] layout (ints), copied to the host after the topology.
constexpr int kInfoN = 0, kInfoOverflow = 2, kInfoUnsorted = kInfoUnsortedFlag;
constexpr int kInfoT = 4, kInfoF = 5, kInfoD = 6, kInfoFront = 7;
constexpr int kInfoTopLevel = 8;                          // kMaxDepth + 3 entries
constexpr int kInfoInts = kInfoTopLevel + kMaxDepth + 3;
static_assert(kMailData + kInfoInts <= kMailInts, "the info words fit in the host mailbox");

// A build whose later kernels read its sizes from the info words: it failed
// when its input was not ascending or its node slots ran out, and then has
// no nodes.
__device__ __forceinline__ bool dev_build_failed(const int *info) {
    return info[kInfoUnsorted] != 0 || info[kInfoOverflow] != 0;
}
__device__ __forceinline__ int dev_node_count(const int *info) {
    return dev_build_failed(info) ? 0 : info[kInfoN];
}

struct TopoCtx {
    const double *sorted;
    int p, leaf_cap;
    const int *bstart;
    int R;   // hint buckets (bstart holds R + 1 entries)
    Hint h;  // set on the device at kernel entry (hint_of)
    const int *gap;  // gap table (hint_scatter_kernel)
    int ng;          // its usable entries, set on the device at kernel entry
    int *b_lo, *b_hi, *b_left, *b_size, *b_pre;
};

// Splits the nodes of slots [lb, le) at depth `depth`: nodes with more than
// leaf_cap sources (and depth < kMaxDepth) get two children, placed in node
// order at slots alloc, alloc + 1, ...; leaves get b_left = -1.  Returns
// the number of child slots used (uniform), or sets *overflow when they
// would pass `limit`.  Loads are unconditional at clamped, in-bounds indices
// and masked afterwards, so the kTopoK nodes of a thread issue them together.
template <int THREADS>
__device__ int expand_level(const TopoCtx &c, int lb, int le, int depth, int alloc, int limit,
                            int *s_wave, int *overflow) {
    const int tid = threadIdx.x;
    int used = 0;
    for (int base = lb; base < le; base += THREADS * kTopoK) {
        const int v0 = base + tid * kTopoK;
        int a[kTopoK], b[kTopoK], l[kTopoK], r[kTopoK];
        bool need[kTopoK];
        double mid[kTopoK];
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            const int v = min(v0 + k, le - 1);
            a[k] = c.b_lo[v];
            b[k] = c.b_hi[v];
        }
#pragma unroll
        for (int k = 0; k < kTopoK; ++k)
            need[k] = (v0 + k < le) && (b[k] - a[k] > c.leaf_cap) && (depth < kMaxDepth);
        double xa[kTopoK], xb[kTopoK];
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            xa[k] = c.sorted[need[k] ? a[k] : 0];
            xb[k] = c.sorted[need[k] ? b[k] - 1 : 0];
        }
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) mid[k] = xa[k] + 0.5 * (xb[k] - xa[k]);
        int hl[kTopoK], hr[kTopoK], kb[kTopoK];
        bool gap_lookup = false;
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            kb[k] = hint_bucket(mid[k], c.h);  // in [0, R-1]
            hl[k] = c.bstart[kb[k]];
            hr[k] = c.bstart[kb[k] + 1];
            gap_lookup |= need[k] && (hl[k] < 0 || hr[k] < 0);
        }
        if (gap_lookup) {  // a bucket in a long gap: the gap table gives its split
            const int4 *gt = reinterpret_cast<const int4 *>(c.gap + 4);
            for (int g = 0; g < c.ng; ++g) {
                const int4 e = gt[g];
#pragma unroll
                for (int k = 0; k < kTopoK; ++k) {
                    if (hl[k] < 0 && e.x <= kb[k] && kb[k] <= e.y) hl[k] = e.z;
                    if (hr[k] < 0 && e.x <= kb[k] + 1 && kb[k] + 1 <= e.y) hr[k] = e.z;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            const bool ok = hl[k] >= 0 && hr[k] >= 0;
            // upper_bound(mid) in [hl, hr], or partition_point in [a, b)
            l[k] = need[k] ? (ok ? hl[k] : a[k]) : 0;
            r[k] = need[k] ? (ok ? hr[k] : b[k]) : 0;
        }
        // partition_point(x <= mid), the kTopoK searches in lockstep
        bool busy = false;
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) busy |= l[k] < r[k];
        while (busy) {
            int m[kTopoK];
            double xm[kTopoK];
#pragma unroll
            for (int k = 0; k < kTopoK; ++k) {
                m[k] = l[k] + (r[k] - l[k]) / 2;
                xm[k] = c.sorted[min(m[k], c.p - 1)];
            }
            busy = false;
#pragma unroll
            for (int k = 0; k < kTopoK; ++k) {
                const bool act = l[k] < r[k];
                const bool le_mid = xm[k] <= mid[k];
                l[k] = (act && le_mid) ? m[k] + 1 : l[k];
                r[k] = (act && !le_mid) ? m[k] : r[k];
                busy |= l[k] < r[k];
            }
        }
        int split[kTopoK], mine = 0;
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            split[k] = 0;
            if (need[k]) {
                const int u = min(max(l[k], a[k]), b[k]);
                split[k] = (u == a[k] || u == b[k]) ? (a[k] + b[k]) / 2 : u;
                mine += 2;
            }
        }
        int added;
        const int off = block_exclusive_scan<THREADS>(mine, s_wave, &added);
        if (alloc + used + added > limit) {
            *overflow = 1;
            return used;
        }
        int ch = alloc + used + off;
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            const int v = v0 + k;
            if (v < le) {
                if (need[k]) {
                    c.b_left[v] = ch;
                    c.b_lo[ch] = a[k];
                    c.b_hi[ch] = split[k];
                    c.b_lo[ch + 1] = split[k];
                    c.b_hi[ch + 1] = b[k];
                    ch += 2;
                } else {
                    c.b_left[v] = -1;
                }
            }
        }
        used += added;
    }
    return used;
}

// b_size of the nodes in slots [lb, le) from their children's sizes.
template <int THREADS>
__device__ void sizes_level(const TopoCtx &c, int lb, int le) {
    for (int v0 = lb + (int)threadIdx.x * kTopoK; v0 < le; v0 += THREADS * kTopoK) {
        int lf[kTopoK], s0[kTopoK], s1[kTopoK];
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            lf[k] = c.b_left[min(v0 + k, le - 1)];
            if (v0 + k >= le) lf[k] = -1;
        }
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            s0[k] = c.b_size[lf[k] >= 0 ? lf[k] : 0];
            s1[k] = c.b_size[lf[k] >= 0 ? lf[k] + 1 : 0];
        }
#pragma unroll
        for (int k = 0; k < kTopoK; ++k)
            if (v0 + k < le) c.b_size[v0 + k] = (lf[k] < 0) ? 1 : 1 + s0[k] + s1[k];
    }
}

// Pre-order numbers of the children of the nodes in slots [lb, le).
template <int THREADS>
__device__ void preorder_level(const TopoCtx &c, int lb, int le) {
    for (int v0 = lb + (int)threadIdx.x * kTopoK; v0 < le; v0 += THREADS * kTopoK) {
        int lf[kTopoK], pv[kTopoK], sl[kTopoK];
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            const int v = min(v0 + k, le - 1);
            lf[k] = c.b_left[v];
            pv[k] = c.b_pre[v];
            if (v0 + k >= le) lf[k] = -1;
        }
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) sl[k] = c.b_size[lf[k] >= 0 ? lf[k] : 0];
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            if (lf[k] >= 0) {
                c.b_pre[lf[k]] = pv[k] + 1;
                c.b_pre[lf[k] + 1] = pv[k] + 1 + sl[k];
            }
        }
    }
}

// One breadth-first level whose nodes' ranges and end values sit in LDS
// (arrays ca/cb/cxa/cxb, level-local index; thread t holds nodes tK..tK+K-1):
// a node's midpoint is known at once, the split search then fetches the
// children's end values sorted[split - 1] and sorted[split], and the
// children go to the slot arrays at alloc, alloc + 1, ... in node order and,
// when they number at most `cap`, to the next level's LDS arrays (*fits).
// The splits are expand_level's.  Returns the children count (uniform), or
// sets *overflow when they would pass `limit`.  Ends on an LDS-only barrier:
// a caller that reads the slot arrays next must first wait for the stores.
template <int THREADS, int K>
__device__ int lds_level(const TopoCtx &c, int lb, int cnt, int depth, int alloc, int limit,
                         int cap, const int *ca, const int *cb, const double *cxa,
                         const double *cxb, int *na, int *nb, double *nxa, double *nxb,
                         int *s_wave, int *overflow, bool *fits) {
    const int tid = threadIdx.x;
    const int4 *gt = reinterpret_cast<const int4 *>(c.gap + 4);
    int a[K], b[K], split[K];
    bool need[K];
    double xa[K], xb[K], xs0[K], xs1[K];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid * K + k;
        const bool node = i < cnt;
        a[k] = node ? ca[i] : 0;
        b[k] = node ? cb[i] : 0;
        xa[k] = node ? cxa[i] : 0.0;
        xb[k] = node ? cxb[i] : 0.0;
        need[k] = node && (b[k] - a[k] > c.leaf_cap) && (depth < kMaxDepth);
        mine += need[k] ? 2 : 0;
    }
    int l[K], r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double mid = xa[k] + 0.5 * (xb[k] - xa[k]);
        const int kb = hint_bucket(mid, c.h);  // in [0, R-1]
        int hl = need[k] ? c.bstart[kb] : 0, hr = need[k] ? c.bstart[kb + 1] : 0;
        if (need[k] && (hl < 0 || hr < 0)) {  // a bucket in a long gap
            for (int g = 0; g < c.ng; ++g) {
                const int4 e = gt[g];
                if (hl < 0 && e.x <= kb && kb <= e.y) hl = e.z;
                if (hr < 0 && e.x <= kb + 1 && kb + 1 <= e.y) hr = e.z;
            }
        }
        const bool ok = hl >= 0 && hr >= 0;
        // upper_bound(mid) in [hl, hr], or partition_point in [a, b)
        l[k] = need[k] ? (ok ? hl : a[k]) : 0;
        r[k] = need[k] ? (ok ? hr : b[k]) : 0;
    }
    // partition_point(x <= mid), the K searches in lockstep
    bool busy = false;
#pragma unroll
    for (int k = 0; k < K; ++k) busy |= l[k] < r[k];
    while (busy) {
        int m[K];
        double xm[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            m[k] = l[k] + (r[k] - l[k]) / 2;
            xm[k] = c.sorted[min(m[k], c.p - 1)];
        }
        busy = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double mid = xa[k] + 0.5 * (xb[k] - xa[k]);
            const bool act = l[k] < r[k];
            const bool le_mid = xm[k] <= mid;
            l[k] = (act && le_mid) ? m[k] + 1 : l[k];
            r[k] = (act && !le_mid) ? m[k] : r[k];
            busy |= l[k] < r[k];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        split[k] = 0;
        xs0[k] = xs1[k] = 0.0;
        if (need[k]) {
            const int u = min(max(l[k], a[k]), b[k]);
            split[k] = (u == a[k] || u == b[k]) ? (a[k] + b[k]) / 2 : u;
            xs0[k] = c.sorted[split[k] - 1];  // a < split < b
            xs1[k] = c.sorted[split[k]];
        }
    }
    int added;
    const int off = block_exclusive_scan<THREADS, true>(mine, s_wave, &added);
    if (alloc + added > limit) {
        *overflow = 1;  // uniform: every thread sees the same total
        return 0;
    }
    *fits = added <= cap;
    int o = off;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid * K + k;
        if (i < cnt) c.b_left[lb + i] = need[k] ? alloc + o : -1;
        if (need[k]) {
            const int ch = alloc + o;
            c.b_lo[ch] = a[k];
            c.b_hi[ch] = split[k];
            c.b_lo[ch + 1] = split[k];
            c.b_hi[ch + 1] = b[k];
            if (*fits) {
                na[o] = a[k];
                nb[o] = split[k];
                nxa[o] = xa[k];
                nxb[o] = xs0[k];
                na[o + 1] = split[k];
                nb[o + 1] = b[k];
                nxa[o + 1] = xs1[k];
                nxb[o + 1] = xb[k];
            }
            o += 2;
        }
    }
    lds_barrier();
    return added;
}

// Subtree levels whose sources sorted[base, base + staged) sit in LDS (src):
// the partition point of x <= mid is searched in the staged values, in the
// bracket a local bucket index gives (StagedIndex), and the children's end
// values are read there.  The search is over sorted values, so the
// partition point in the node's range [a, b) is the clamp into [a, b) of the
// subtree's one, which is what lds_level and expand_level compute from the
// global array (the hint range always holds the global partition point):
// the splits are theirs.
//
// StagedIndex: bucket(v) = min(floor((v - x0) * scale), nb - 1), v >= x0,
// is monotone in v, so with kb = bucket(mid) every staged value of a lower
// bucket is < mid and every one of a higher bucket is > mid: the partition
// point lies in [first[kb], first[kb + 1]], first[k] being the first local
// index whose bucket is >= k (first[nb] = the staged count).
constexpr int kSubBuckets = 2048;
struct StagedIndex {
    const int *first;
    double x0, scale;
};
__device__ __forceinline__ int staged_bucket(double v, const StagedIndex &ix) {
    const double t = (v - ix.x0) * ix.scale;
    return (t >= (double)(kSubBuckets - 1)) ? kSubBuckets - 1 : (t > 0.0 ? (int)t : 0);
}

template <int K>
__device__ __forceinline__ void staged_splits(const double *src, int base, const StagedIndex &ix,
                                              const int (&a)[K], const int (&b)[K],
                                              const bool (&need)[K], const double (&mid)[K],
                                              int (&split)[K]) {
    int l[K], r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int kb = staged_bucket(mid[k], ix);
        const int lo = ix.first[kb] + base, hi = ix.first[kb + 1] + base;
        // the bracket [lo, hi] clamped into [a, b]: the search below then
        // returns the clamp of the subtree's partition point
        l[k] = need[k] ? min(max(lo, a[k]), b[k]) : 0;
        r[k] = need[k] ? max(min(hi, b[k]), l[k]) : 0;
    }
    bool busy = false;
#pragma unroll
    for (int k = 0; k < K; ++k) busy |= l[k] < r[k];
    while (busy) {
        int m[K];
        double xm[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            m[k] = l[k] + (r[k] - l[k]) / 2;
            xm[k] = src[(l[k] < r[k]) ? m[k] - base : 0];
        }
        busy = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const bool act = l[k] < r[k];
            const bool le_mid = xm[k] <= mid[k];
            l[k] = (act && le_mid) ? m[k] + 1 : l[k];
            r[k] = (act && !le_mid) ? m[k] : r[k];
            busy |= l[k] < r[k];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        split[k] = 0;
        if (need[k]) {
            const int u = l[k];  // in [a, b]
            split[k] = (u == a[k] || u == b[k]) ? (a[k] + b[k]) / 2 : u;
        }
    }
}

// lds_level with the subtree's sources staged in LDS (no hint loads; the
// children's end values come from the staged sources).
template <int THREADS, int K>
__device__ __forceinline__ int lds_level_staged(const TopoCtx &c, int lb, int cnt, int depth,
                                                int alloc, int limit, int cap, const int *ca,
                                                const int *cb, const double *cxa,
                                                const double *cxb, int *na, int *nb, double *nxa,
                                                double *nxb, int *s_wave, int *overflow,
                                                bool *fits, const double *src, int base,
                                                const StagedIndex &ix) {
    const int tid = threadIdx.x;
    int a[K], b[K], split[K];
    bool need[K];
    double xa[K], xb[K], mid[K];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid * K + k;
        const bool node = i < cnt;
        a[k] = node ? ca[i] : 0;
        b[k] = node ? cb[i] : 0;
        xa[k] = node ? cxa[i] : 0.0;
        xb[k] = node ? cxb[i] : 0.0;
        need[k] = node && (b[k] - a[k] > c.leaf_cap) && (depth < kMaxDepth);
        mid[k] = xa[k] + 0.5 * (xb[k] - xa[k]);
        mine += need[k] ? 2 : 0;
    }
    staged_splits<K>(src, base, ix, a, b, need, mid, split);
    int added;
    const int off = block_exclusive_scan<THREADS, true>(mine, s_wave, &added);
    if (alloc + added > limit) {
        *overflow = 1;  // uniform: every thread sees the same total
        return 0;
    }
    *fits = added <= cap;
    int o = off;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid * K + k;
        if (i < cnt) c.b_left[lb + i] = need[k] ? alloc + o : -1;
        if (need[k]) {
            const int ch = alloc + o;
            c.b_lo[ch] = a[k];
            c.b_hi[ch] = split[k];
            c.b_lo[ch + 1] = split[k];
            c.b_hi[ch + 1] = b[k];
            if (*fits) {
                na[o] = a[k];
                nb[o] = split[k];
                nxa[o] = xa[k];
                nxb[o] = src[split[k] - 1 - base];  // a < split < b
                na[o + 1] = split[k];
                nb[o + 1] = b[k];
                nxa[o + 1] = src[split[k] - base];
                nxb[o + 1] = xb[k];
            }
            o += 2;
        }
    }
    lds_barrier();
    return added;
}

// expand_level with the subtree's sources staged in LDS.
template <int THREADS>
__device__ __forceinline__ int expand_level_staged(const TopoCtx &c, int lb, int le, int depth,
                                                   int alloc, int limit, int *s_wave,
                                                   int *overflow, const double *src, int base,
                                                   const StagedIndex &ix) {
    const int tid = threadIdx.x;
    int used = 0;
    for (int vb = lb; vb < le; vb += THREADS * kTopoK) {
        const int v0 = vb + tid * kTopoK;
        int a[kTopoK], b[kTopoK], split[kTopoK];
        bool need[kTopoK];
        double mid[kTopoK];
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            const int v = min(v0 + k, le - 1);
            a[k] = c.b_lo[v];
            b[k] = c.b_hi[v];
        }
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            need[k] = (v0 + k < le) && (b[k] - a[k] > c.leaf_cap) && (depth < kMaxDepth);
            const double xa = need[k] ? src[a[k] - base] : 0.0;
            const double xb = need[k] ? src[b[k] - 1 - base] : 0.0;
            mid[k] = xa + 0.5 * (xb - xa);
        }
        staged_splits<kTopoK>(src, base, ix, a, b, need, mid, split);
        int mine = 0;
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) mine += need[k] ? 2 : 0;
        int added;
        const int off = block_exclusive_scan<THREADS>(mine, s_wave, &added);
        if (alloc + used + added > limit) {
            *overflow = 1;
            return used;
        }
        int ch = alloc + used + off;
#pragma unroll
        for (int k = 0; k < kTopoK; ++k) {
            const int v = v0 + k;
            if (v < le) {
                if (need[k]) {
                    c.b_left[v] = ch;
                    c.b_lo[ch] = a[k];
                    c.b_hi[ch] = split[k];
                    c.b_lo[ch + 1] = split[k];
                    c.b_hi[ch + 1] = b[k];
                    ch += 2;
                } else {
                    c.b_left[v] = -1;
                }
            }
        }
        used += added;
    }
    return used;
}

// One breadth-first level of at most 64 nodes on wavefront 0 alone (its
// lanes only), G = 64 / 2^ceil(log2 cnt) lanes per node.  With G >= 4 the
// binary search stops once [l, r) holds at most G - 2 entries; the node's
// lanes then load the window sorted[l - 1 .. r] together, the count of its
// entries <= mid completes the partition point (the search is over sorted
// values, so the predicate holds on a prefix), and the window already holds
// the children's end values sorted[split - 1] and sorted[split].  A level
// then costs the hint load and one window load.  The splits, slots and LDS
// arrays are lds_level's, and the children of a level of at most 64 nodes
// always fit the next level's LDS arrays.  Returns the children count
// (uniform), or sets *overflow when they would pass `limit`.  Ends with the
// wavefront's LDS writes visible to its own lanes.
__device__ int narrow_level(const TopoCtx &c, int lb, int cnt, int depth, int alloc, int limit,
                            const int *ca, const int *cb, const double *cxa, const double *cxb,
                            int *na, int *nb, double *nxa, double *nxb, int *overflow) {
    const int lane = threadIdx.x & 63;
    int lg = 0;
    while ((1 << lg) < cnt) ++lg;  // cnt <= 64
    const int G = 64 >> lg, i = lane >> (6 - lg), g = lane & (G - 1), gb = i * G;
    const bool valid = i < cnt;
    const int a = valid ? ca[i] : 0, b = valid ? cb[i] : 0;
    const double xa = valid ? cxa[i] : 0.0, xb = valid ? cxb[i] : 0.0;
    const bool need = valid && (b - a > c.leaf_cap) && (depth < kMaxDepth);
    const double mid = xa + 0.5 * (xb - xa);
    const int kb = hint_bucket(mid, c.h);  // in [0, R-1]
    int hl = need ? c.bstart[kb] : 0, hr = need ? c.bstart[kb + 1] : 0;
    if (need && (hl < 0 || hr < 0)) {  // a bucket in a long gap
        const int4 *gt = reinterpret_cast<const int4 *>(c.gap + 4);
        for (int e = 0; e < c.ng; ++e) {
            const int4 t = gt[e];
            if (hl < 0 && t.x <= kb && kb <= t.y) hl = t.z;
            if (hr < 0 && t.x <= kb + 1 && kb + 1 <= t.y) hr = t.z;
        }
    }
    const bool ok = hl >= 0 && hr >= 0;
    int l = need ? (ok ? hl : a) : 0, r = need ? (ok ? hr : b) : 0;
    const int W = (G >= 4) ? G - 2 : 0;  // window capacity
    while (__any(r - l > W)) {
        if (r - l > W) {
            const int m = l + (r - l) / 2;
            const double xm = c.sorted[min(m, c.p - 1)];
            if (xm <= mid) l = m + 1;
            else r = m;
        }
    }
    double wv = 0.0;
    int u = l;
    if (W > 0) {
        const bool in = need && g <= r - l + 1;  // the window is [l - 1, r]
        wv = in ? c.sorted[min(max(l - 1 + g, 0), c.p - 1)] : 0.0;
        const bool pred = need && g >= 1 && g <= r - l && wv <= mid;
        const uint64_t gm = (G == 64) ? ~0ull : (((1ull << G) - 1) << gb);
        u += __popcll(__ballot(pred) & gm);
    }
    int split = 0;
    bool from_window = false;
    if (need) {
        const int uu = min(max(u, a), b);
        split = (uu == a || uu == b) ? (a + b) / 2 : uu;
        from_window = W > 0 && split == u;
    }
    const int g0 = from_window ? split - l : 0;  // sorted[split - 1] sits at window entry g0
    const double w0 = __shfl(wv, gb + g0, 64), w1 = __shfl(wv, gb + min(g0 + 1, G - 1), 64);
    double xs0 = 0.0, xs1 = 0.0;
    if (need) {
        if (from_window) {
            xs0 = w0;
            xs1 = w1;
        } else {
            xs0 = c.sorted[split - 1];  // a < split < b
            xs1 = c.sorted[split];
        }
    }
    const bool lead = g == 0;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t mneed = __ballot(lead && need);
    const int added = 2 * __popcll(mneed);
    if (alloc + added > limit) {
        *overflow = 1;  // uniform
        return 0;
    }
    if (lead && valid) {
        const int o = 2 * __popcll(mneed & below);
        c.b_left[lb + i] = need ? alloc + o : -1;
        if (need) {
            const int ch = alloc + o;
            c.b_lo[ch] = a;
            c.b_hi[ch] = split;
            c.b_lo[ch + 1] = split;
            c.b_hi[ch + 1] = b;
            na[o] = a;
            nb[o] = split;
            nxa[o] = xa;
            nxb[o] = xs0;
            na[o + 1] = split;
            nb[o + 1] = b;
            nxa[o + 1] = xs1;
            nxb[o + 1] = xb;
        }
    }
    wave_sync();
    return added;
}

// A: breadth-first from the root while levels are smaller than kFrontier,
// each level in LDS, double-buffered: levels of at most 64 nodes on
// wavefront 0 (narrow_level), then one node per thread (lds_level).
constexpr int kTopCap = 2 * kFrontier;  // children of a level below kFrontier

__global__ __launch_bounds__(kTopoThreads) void topo_top_kernel(TopoCtx c, int cap_top,
                                                                int *__restrict__ info) {
    static_assert(kFrontier <= kTopoThreads, "one node of a top level per thread");
    __shared__ int s_wave[kTopoThreads / 64];
    __shared__ int s_a[2][kTopCap], s_b[2][kTopCap];
    __shared__ double s_xa[2][kTopCap], s_xb[2][kTopCap];
    const int tid = threadIdx.x;
    // An input that ingest_kernel found not ascending is sorted and its
    // topology built again; this pass then publishes an empty top walk (no
    // frontier: the subtree kernels return at once) instead of walking the
    // unsorted copy.
    if (info[kInfoUnsorted] != 0) {
        if (tid == 0) {
            info[kInfoOverflow] = 0;
            info[kInfoT] = 0;
            info[kInfoF] = 0;
            info[kInfoD] = 0;
            info[kInfoFront] = 0;
        }
        return;
    }
    c.h = hint_of(c.sorted, c.p, c.R);
    c.ng = min(c.gap[0], kMaxGaps);
    if (tid == 0) {
        c.b_lo[0] = 0;
        c.b_hi[0] = c.p;
        s_a[0][0] = 0;
        s_b[0][0] = c.p;
        s_xa[0][0] = c.sorted[0];
        s_xb[0][0] = c.sorted[c.p - 1];
        info[kInfoTopLevel + 0] = 0;
        info[kInfoTopLevel + 1] = 1;
    }
    lds_barrier();
    __shared__ int s_state[4];
    int lb = 0, le = 1, depth = 0, overflow = 0, cur = 0;
    if (tid < 64) {
        while (lb < le && le - lb <= 64 && !overflow) {
            const int nxt = cur ^ 1;
            const int added = narrow_level(c, lb, le - lb, depth, le, cap_top, s_a[cur], s_b[cur],
                                           s_xa[cur], s_xb[cur], s_a[nxt], s_b[nxt], s_xa[nxt],
                                           s_xb[nxt], &overflow);
            if (overflow) break;
            lb = le;
            le += added;
            ++depth;
            cur = nxt;
            if (tid == 0) info[kInfoTopLevel + depth + 1] = le;
        }
        if (tid == 0) {
            s_state[0] = lb;
            s_state[1] = le;
            s_state[2] = depth;
            s_state[3] = overflow | (cur << 1);
        }
    }
    lds_barrier();
    lb = s_state[0];
    le = s_state[1];
    depth = s_state[2];
    overflow = s_state[3] & 1;
    cur = s_state[3] >> 1;
    bool fits = true;
    while (lb < le && le - lb < kFrontier && !overflow) {
        const int nxt = cur ^ 1;
        const int added = lds_level<kTopoThreads, 1>(
            c, lb, le - lb, depth, le, cap_top, kTopCap, s_a[cur], s_b[cur], s_xa[cur],
            s_xb[cur], s_a[nxt], s_b[nxt], s_xa[nxt], s_xb[nxt], s_wave, &overflow, &fits);
        if (overflow) break;
        lb = le;
        le += added;
        ++depth;
        cur = nxt;
        if (tid == 0) info[kInfoTopLevel + depth + 1] = le;
    }
    if (tid == 0) {
        info[kInfoOverflow] = overflow;
        info[kInfoT] = le;
        info[kInfoF] = overflow ? 0 : le - lb;  // 0 when the tree ended within the top walk
        info[kInfoD] = depth;
        info[kInfoFront] = lb;
    }
}

// B: the subtree of frontier node blockIdx.x, then its sizes.
__global__ __launch_bounds__(kSubThreads) void topo_sub_kernel(TopoCtx c,
                                                               const int *__restrict__ info,
                                                               int *__restrict__ lv_start,
                                                               int *__restrict__ lv_cnt,
                                                               int *__restrict__ sub_nodes) {
    __shared__ int s_wave[kSubThreads / 64];
    __shared__ int s_lb[kLv], s_le[kLv];
    __shared__ int s_a[2][kSubCap], s_b[2][kSubCap];
    __shared__ double s_xa[2][kSubCap], s_xb[2][kSubCap];
    __shared__ double s_src[kSubStaged];
    __shared__ int s_first[kSubBuckets + 1];
    const int f = blockIdx.x;
    if (f >= info[kInfoF]) return;
    c.h = hint_of(c.sorted, c.p, c.R);
    c.ng = min(c.gap[0], kMaxGaps);
    const int T = info[kInfoT], D = info[kInfoD];
    const int v = info[kInfoFront] + f;
    const int a = c.b_lo[v], b = c.b_hi[v];
    int lb = v, le = v + 1, alloc = T + 2 * a, lev = 0, overflow = 0, cur = 0;
    const int limit = T + 2 * b;
    // a subtree of kSubStageMin to kSubStaged sources searches its staged copy
    const bool staged = b - a >= kSubStageMin && b - a <= kSubStaged;
    StagedIndex ix{s_first, 0.0, 0.0};
    if (staged) {
        for (int i = threadIdx.x; i < b - a; i += kSubThreads) s_src[i] = c.sorted[a + i];
        lds_barrier();
        ix.x0 = s_src[0];
        const double span = s_src[b - a - 1] - ix.x0;
        ix.scale = (span > 0.0) ? (double)kSubBuckets / span : 0.0;
        // first[k]: the first index of each non-empty bucket, then a suffix
        // minimum carries it into the empty buckets below (and n past the last)
        const int n = b - a;
        for (int k = threadIdx.x; k <= kSubBuckets; k += kSubThreads) s_first[k] = n;
        lds_barrier();
        for (int i = threadIdx.x; i < n; i += kSubThreads) {
            const int bi = staged_bucket(s_src[i], ix);
            if (i == 0 || staged_bucket(s_src[i - 1], ix) < bi) s_first[bi] = i;
        }
        lds_barrier();
        constexpr int C = (kSubBuckets + kSubThreads) / kSubThreads;  // entries per thread
        const int k0 = threadIdx.x * C, k1 = min(k0 + C, kSubBuckets + 1);
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        int m = n;
        for (int k = k0; k < k1; ++k) m = min(m, s_first[k]);
        int x = m;  // becomes the minimum over this wavefront's lanes >= lane
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_down(x, o, 64);
            if (lane + o < 64) x = min(x, y);
        }
        if (lane == 0) s_wave[wid] = x;
        const int next = __shfl_down(x, 1, 64);
        lds_barrier();
        int run = (lane < 63) ? next : n;  // minimum over the chunks after this thread's
        for (int w = wid + 1; w < kSubThreads / 64; ++w) run = min(run, s_wave[w]);
        for (int k = k1 - 1; k >= k0; --k) {
            run = min(run, s_first[k]);
            s_first[k] = run;
        }
        lds_barrier();  // s_wave is reused by the levels' scans
    }
    if (threadIdx.x == 0) {
        s_lb[0] = lb;
        s_le[0] = le;
        s_a[0][0] = a;
        s_b[0][0] = b;
        s_xa[0][0] = c.sorted[a];
        s_xb[0][0] = c.sorted[b - 1];
    }
    lds_barrier();
    // levels in LDS while they fit (kSubCap nodes), then on the slot arrays
    bool in_lds = true;
    while (true) {
        int used;
        if (in_lds) {
            const int nxt = cur ^ 1;
            bool fits = true;
            if (staged)
                used = lds_level_staged<kSubThreads, kSubCap / kSubThreads>(
                    c, lb, le - lb, D + lev, alloc, limit, kSubCap, s_a[cur], s_b[cur], s_xa[cur],
                    s_xb[cur], s_a[nxt], s_b[nxt], s_xa[nxt], s_xb[nxt], s_wave, &overflow, &fits,
                    s_src, a, ix);
            else
                used = lds_level<kSubThreads, kSubCap / kSubThreads>(
                    c, lb, le - lb, D + lev, alloc, limit, kSubCap, s_a[cur], s_b[cur], s_xa[cur],
                    s_xb[cur], s_a[nxt], s_b[nxt], s_xa[nxt], s_xb[nxt], s_wave, &overflow, &fits);
            cur = nxt;
            if (!fits) {
                in_lds = false;
                __syncthreads();  // the next level reads the slot arrays
            }
        } else {
            if (staged)
                used = expand_level_staged<kSubThreads>(c, lb, le, D + lev, alloc, limit, s_wave,
                                                        &overflow, s_src, a, ix);
            else
                used = expand_level<kSubThreads>(c, lb, le, D + lev, alloc, limit, s_wave,
                                                 &overflow);
            __syncthreads();
        }
        if (used == 0 || overflow) break;
        lb = alloc;
        le = alloc + used;
        alloc += used;
        ++lev;
        if (threadIdx.x == 0) {
            s_lb[lev] = lb;
            s_le[lev] = le;
        }
    }
    __syncthreads();  // also orders this workgroup's slot-array stores before the loads below
    const int n_lv = lev + 1;
    // Sizes, deepest level first.  The levels below the root fill the slots
    // [s0, alloc) contiguously; when they fit, their left-child slots and
    // sizes sit in LDS (in the staged sources' space, no longer needed).
    const int s0 = T + 2 * a, ns = alloc - s0;
    if (ns <= kSubStaged) {
        int *l_left = reinterpret_cast<int *>(s_src), *l_size = l_left + kSubStaged;
        for (int i = threadIdx.x; i < ns; i += kSubThreads) l_left[i] = c.b_left[s0 + i];
        lds_barrier();
        for (int l = n_lv - 1; l >= 1; --l) {
            for (int u = s_lb[l] + (int)threadIdx.x; u < s_le[l]; u += kSubThreads) {
                const int lf = l_left[u - s0];
                l_size[u - s0] = (lf < 0) ? 1 : 1 + l_size[lf - s0] + l_size[lf + 1 - s0];
            }
            lds_barrier();
        }
        for (int i = threadIdx.x; i < ns; i += kSubThreads) c.b_size[s0 + i] = l_size[i];
        if (threadIdx.x == 0) {
            const int lf = c.b_left[v];
            c.b_size[v] = (lf < 0) ? 1 : 1 + l_size[lf - s0] + l_size[lf + 1 - s0];
        }
    } else {
        for (int l = n_lv - 1; l >= 0; --l) {
            sizes_level<kSubThreads>(c, s_lb[l], s_le[l]);
            __syncthreads();
        }
    }
    for (int l = threadIdx.x; l < kLv; l += kSubThreads) {
        lv_start[f * kLv + l] = (l < n_lv) ? s_lb[l] : 0;
        lv_cnt[f * kLv + l] = (l < n_lv) ? s_le[l] - s_lb[l] : 0;
    }
    // the levels below the root fill slots [s_lb[1], s_le[n_lv - 1]) contiguously
    if (threadIdx.x == 0) sub_nodes[f] = (n_lv > 1) ? s_le[n_lv - 1] - s_lb[1] : 0;
}

// C: top sizes and pre-order numbers, then the node list's blocks.  With at
// most kMidLds top slots both sweeps run in LDS (a top slot's children are
// top slots; the frontier roots' sizes come from B), otherwise on the slot
// arrays in global memory.
constexpr int kMidLds = 4096;

__global__ __launch_bounds__(kTopoThreads) void topo_mid_kernel(
    TopoCtx c, int *__restrict__ info, const int *__restrict__ sub_nodes, int *__restrict__ lv_off,
    int *__restrict__ list, int *__restrict__ mail, int seq) {
    __shared__ int s_wave[kTopoThreads / 64];
    __shared__ int s_left[kMidLds], s_size[kMidLds], s_pre[kMidLds];
    const int tid = threadIdx.x;
    // the info words to the host mailbox, then its sequence number (a build
    // whose later kernels read the info words themselves passes no mailbox)
    auto publish = [&]() {
        if (!mail) return;
        __syncthreads();
        for (int i = tid; i < kInfoInts; i += kTopoThreads) mail[kMailData + i] = info[i];
        __threadfence_system();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(mail, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    };
    if (info[kInfoUnsorted] != 0) {  // the host sorts and builds again (topo_top_kernel)
        publish();
        return;
    }
    const int T = info[kInfoT], F = info[kInfoF], D = info[kInfoD];
    const int *lvl = info + kInfoTopLevel;  // top level d = BFS slots [lvl[d], lvl[d+1])
    // Expanded top levels are 0..D-1; with a frontier (F > 0) level D holds
    // the subtree roots, whose sizes come from B.
    if (T <= kMidLds) {
        const int front = (F > 0) ? info[kInfoFront] : T;
        for (int v = tid; v < T; v += kTopoThreads) {
            s_left[v] = c.b_left[v];
            s_size[v] = (v >= front) ? c.b_size[v] : 0;
        }
        __syncthreads();
        // The top levels [0, dn) hold at most 64 slots each: wavefront 0
        // sweeps them alone, one slot per lane, ordered by wavefront-scope
        // LDS fences instead of workgroup barriers.
        int dn = 0;
        while (dn < D && lvl[dn + 1] - lvl[dn] <= 64) ++dn;
        for (int d = D - 1; d >= dn; --d) {
            for (int v = lvl[d] + tid; v < lvl[d + 1]; v += kTopoThreads) {
                const int l = s_left[v];
                s_size[v] = (l < 0) ? 1 : 1 + s_size[l] + s_size[l + 1];
            }
            __syncthreads();
        }
        if (tid < 64) {
            for (int d = dn - 1; d >= 0; --d) {
                const int v = lvl[d] + tid;
                if (v < lvl[d + 1]) {
                    const int l = s_left[v];
                    s_size[v] = (l < 0) ? 1 : 1 + s_size[l] + s_size[l + 1];
                }
                wave_sync();
            }
            if (tid == 0) s_pre[0] = 0;
            wave_sync();
            for (int d = 0; d < dn; ++d) {
                const int v = lvl[d] + tid;
                if (v < lvl[d + 1]) {
                    const int l = s_left[v];
                    if (l >= 0) {
                        s_pre[l] = s_pre[v] + 1;
                        s_pre[l + 1] = s_pre[v] + 1 + s_size[l];
                    }
                }
                wave_sync();
            }
        }
        __syncthreads();
        for (int d = dn; d < D; ++d) {
            for (int v = lvl[d] + tid; v < lvl[d + 1]; v += kTopoThreads) {
                const int l = s_left[v];
                if (l >= 0) {
                    s_pre[l] = s_pre[v] + 1;
                    s_pre[l + 1] = s_pre[v] + 1 + s_size[l];
                }
            }
            __syncthreads();
        }
        for (int v = tid; v < T; v += kTopoThreads) {
            c.b_size[v] = s_size[v];
            c.b_pre[v] = s_pre[v];
        }
    } else {
        for (int d = D - 1; d >= 0; --d) {
            sizes_level<kTopoThreads>(c, lvl[d], lvl[d + 1]);
            __syncthreads();
        }
        if (tid == 0) c.b_pre[0] = 0;
        __syncthreads();
        for (int d = 0; d < D; ++d) {
            preorder_level<kTopoThreads>(c, lvl[d], lvl[d + 1]);
            __syncthreads();
        }
    }
    // subtree f's nodes below its root take list[lv_off[f * kLv], + sub_nodes[f])
    const int mine = (tid < F) ? sub_nodes[tid] : 0;
    int tot;
    const int off = block_exclusive_scan<kTopoThreads>(mine, s_wave, &tot);
    if (tid < F) lv_off[tid * kLv] = T + off;
    if (tid == 0) info[kInfoN] = T + tot;
    publish();
    // the top slots' node-list entries (read by later kernels, not the host)
    for (int i = tid; i < T; i += kTopoThreads) list[i] = i;
}

// D: pre-order numbers inside the subtree and its entries of the node list.
__global__ __launch_bounds__(kSubThreads) void topo_fin_kernel(TopoCtx c,
                                                               const int *__restrict__ info,
                                                               const int *__restrict__ lv_start,
                                                               const int *__restrict__ lv_cnt,
                                                               const int *__restrict__ lv_off,
                                                               int *__restrict__ list) {
    const int f = blockIdx.x;
    if (f >= info[kInfoF]) return;
    for (int l = 0; l + 1 < kLv && lv_cnt[f * kLv + l + 1] > 0; ++l) {
        const int lb = lv_start[f * kLv + l];
        preorder_level<kSubThreads>(c, lb, lb + lv_cnt[f * kLv + l]);
        __syncthreads();
    }
    int o = lv_off[f * kLv];
    for (int l = 1; l < kLv && lv_cnt[f * kLv + l] > 0; ++l) {
        const int lb = lv_start[f * kLv + l], n = lv_cnt[f * kLv + l];
        for (int i = threadIdx.x; i < n; i += kSubThreads) list[o + i] = lb + i;
        o += n;
    }
}

struct TreeArrays {
    double *lo, *hi, *acc, *nodes, *w;
    int32_t *lo_idx, *hi_idx, *right, *esc;
};

// Barycentric accumulation of `nrows` rows (x_r, mass_r) into one panel's
// weights w[0..n) (chebcode.rs:309-337), one wavefront per panel; w32
// receives the f32 copy of the weights (the FAST panel).
//
// Lanes first evaluate 64 rows at a time: the quotients lam_j/(x_r - t_j)
// go to LDS together with m_r = mass_r * (1/s_r), or the node index hit by
// an exact coincidence.  Then lane j adds the rows into w_j in row order,
// w_j = fma(m_r, q_rj, w_j), which is the reference's per-row update in the
// reference's order, so the weights equal a sequential build bit for bit.
// Rows with zero mass are skipped (merge_all_weights, chebcode.rs:496-499).
// `row(r, x, mass)` supplies row r; `t` (the panel's Chebyshev positions)
// may be in LDS or global memory.
//
// NT > 0 fixes n = NT (the presets' orders): positions and weights sit in
// registers and every quotient of a row is formed before the first exact
// hit is picked, so the loads and divisions of a row issue together.  A
// row without a hit gets the same quotients and the same sum s in the same
// order as the sequential loop; a row with a hit uses only its mass and
// the hit's index, as there.
constexpr int kBaryLanes = 64;
constexpr int kWaccPerLane = (kMaxN + kBaryLanes - 1) / kBaryLanes;

// Adds rows [r0, r0 + cnt) of the tables into acc in row order: w_j =
// fma(m_r, q_rj, w_j) for a regular row (hit -1), w_j += m_r for a row that
// hit node j, nothing otherwise.  Eight rows' entries are read before they
// are added, and the update is a selection between the candidates (quotients
// of other rows are never used), so the row loop has no branches and no
// per-row LDS round trip; the arithmetic and its order are the sequential
// ones.
__device__ __forceinline__ double bary_rows(double acc, int r0, int cnt, int j, int n,
                                            const double *__restrict__ sq,
                                            const double *__restrict__ sm,
                                            const int *__restrict__ shit) {
    int v = 0;
    for (; v + 8 <= cnt; v += 8) {
        int h[8];
        double m[8], q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            h[u] = shit[r0 + v + u];
            m[u] = sm[r0 + v + u];
            q[u] = sq[(r0 + v + u) * n + j];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double f = __builtin_fma(m[u], q[u], acc), a = acc + m[u];
            acc = (h[u] == -1) ? f : ((h[u] == j) ? a : acc);
        }
    }
    for (; v < cnt; ++v) {
        const int h = shit[r0 + v];
        const double m = sm[r0 + v], q = sq[(r0 + v) * n + j];
        const double f = __builtin_fma(m, q, acc), a = acc + m;
        acc = (h == -1) ? f : ((h == j) ? a : acc);
    }
    return acc;
}

template <int NT, class RowFn>
__device__ void bary_accumulate(RowFn row, int nrows, const double *__restrict__ t,
                                const double *__restrict__ lam, double *__restrict__ w,
                                float *__restrict__ w32, int n_rt, double *__restrict__ sq,
                                double *__restrict__ sm, int *__restrict__ shit) {
    const int n = (NT > 0) ? NT : n_rt;
    const int lane = threadIdx.x & (kBaryLanes - 1);
    constexpr int kReg = (NT > 0) ? NT : 1;
    double tr[kReg], lr[kReg];
    if (NT > 0) {
#pragma unroll
        for (int j = 0; j < kReg; ++j) {
            tr[j] = t[j];
            lr[j] = lam[j];
        }
    }
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
                if (NT > 0) {
#pragma unroll
                    for (int j = 0; j < kReg; ++j) {
                        const double d = x - tr[j];
                        const double q = lr[j] / d;
                        sq[lane * kReg + j] = q;
                        s += q;
                        hit = (hit < 0 && d == 0.0) ? j : hit;
                    }
                } else {
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
                }
                m = (hit >= 0) ? mass : mass * (1.0 / s);
            }
        }
        sm[lane] = m;
        shit[lane] = hit;
        wave_sync();
        const int rows = min(kBaryLanes, nrows - base);
#pragma unroll
        for (int u = 0; u < kWaccPerLane; ++u) {
            const int j = lane + u * kBaryLanes;
            if (j < n) wacc[u] = bary_rows(wacc[u], 0, rows, j, n, sq, sm, shit);
        }
        wave_sync();  // the next rows overwrite sq, sm, shit
    }
#pragma unroll
    for (int u = 0; u < kWaccPerLane; ++u) {
        const int j = lane + u * kBaryLanes;
        if (j < n) {
            w[j] = wacc[u];
            w32[j] = (float)wacc[u];
        }
    }
}

// LDS of one wavefront's accumulation: sq[64*n] doubles, sm[64] doubles,
// shit[64] ints (counted as 32 doubles).
__host__ __device__ inline int bary_lds_doubles(int n) {
    return kBaryLanes * n + kBaryLanes + kBaryLanes / 2;
}
inline size_t bary_lds_bytes(int n) { return sizeof(double) * (size_t)bary_lds_doubles(n); }

// bary_accumulate with the register form for the presets' orders.
template <class RowFn>
__device__ __forceinline__ void bary_any(RowFn row, int nrows, const double *__restrict__ t,
                                         const double *__restrict__ lam, double *__restrict__ w,
                                         float *__restrict__ w32, int n, double *__restrict__ sq,
                                         double *__restrict__ sm, int *__restrict__ shit) {
    switch (n) {
    case 8: bary_accumulate<8>(row, nrows, t, lam, w, w32, n, sq, sm, shit); break;
    case 11: bary_accumulate<11>(row, nrows, t, lam, w, w32, n, sq, sm, shit); break;
    case 12: bary_accumulate<12>(row, nrows, t, lam, w, w32, n, sq, sm, shit); break;
    default: bary_accumulate<0>(row, nrows, t, lam, w, w32, n, sq, sm, shit); break;
    }
}

struct BaryLds {
    double *sq, *sm;
    int *shit;
};
__device__ __forceinline__ BaryLds bary_lds(double *base, int n) {
    BaryLds b;
    b.sq = base;
    b.sm = base + kBaryLanes * n;
    b.shit = reinterpret_cast<int *>(b.sm + kBaryLanes);
    return b;
}

// Node records, one wavefront per node (node slots list[0, N)): pre-order
// arrays, Chebyshev positions (chebcode.rs:269-279), the slot's merge record
// (pre-order number, right child) and the f32 mirrors (chebcode.rs:551-570);
// a leaf also gets its weights from its sources (chebcode.rs:401-460).
// Internal nodes get theirs from the merge kernels.
__device__ __forceinline__ void node_one(
    int slot, const double *__restrict__ sorted, int n, double theta_sq,
    const int *__restrict__ list, const int *__restrict__ b_lo, const int *__restrict__ b_hi,
    const int *__restrict__ b_left, const int *__restrict__ b_size, const int *__restrict__ b_pre,
    const double *__restrict__ tab, const TreeArrays &t, int2 *__restrict__ mrec,
    float4 *__restrict__ rec32, int4 *__restrict__ link, float *__restrict__ panel32,
    double *lds) {
    const int lane = threadIdx.x;
    const int v = list[slot];
    const int k = b_pre[v], a = b_lo[v], b = b_hi[v], l = b_left[v], size = b_size[v];
    const int right = (l < 0) ? -1 : b_pre[l + 1];
    const double lo = sorted[a], hi = sorted[b - 1];
    const double hw = 0.5 * (hi - lo);
    const double acc = (hw * hw) / theta_sq;
    if (lane == 0) {
        t.lo[k] = lo;
        t.hi[k] = hi;
        t.acc[k] = acc;
        t.lo_idx[k] = a;
        t.hi_idx[k] = b;
        t.right[k] = right;
        t.esc[k] = k + size;
        mrec[v] = make_int2(k, right);
        rec32[k] = make_float4((float)lo, (float)hi, (float)acc, 0.0f);
        link[k] = make_int4(right, k + size, a, b);
    }
    const BaryLds bl = bary_lds(lds, n);
    double *pos = lds + bary_lds_doubles(n);  // n doubles
    const double c = 0.5 * (lo + hi), hl = 0.5 * (hi - lo);
    for (int j = lane; j < n; j += kBaryLanes) {
        const double x = c + hl * tab[j];
        t.nodes[(size_t)k * n + j] = x;
        panel32[(size_t)k * 2 * n + j] = (float)x;
        pos[j] = x;
    }
    if (right >= 0) return;
    wave_sync();
    const double *xs = sorted + a;
    auto row = [xs](int r, double &x, double &mass) {
        x = xs[r];
        mass = 1.0;
    };
    bary_any(row, b - a, pos, tab + n, t.w + (size_t)k * n,
                    panel32 + (size_t)k * 2 * n + n, n, bl.sq, bl.sm, bl.shit);
}

// The node count is the grid, or the info words' count when `info` is set
// (a build whose sizes stay on the device): the workgroups then take slots
// blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ __launch_bounds__(kBaryLanes) void node_kernel(
    const double *__restrict__ sorted, int n, double theta_sq, const int *__restrict__ list,
    const int *__restrict__ b_lo, const int *__restrict__ b_hi, const int *__restrict__ b_left,
    const int *__restrict__ b_size, const int *__restrict__ b_pre, const double *__restrict__ tab,
    TreeArrays t, int2 *__restrict__ mrec, float4 *__restrict__ rec32, int4 *__restrict__ link,
    float *__restrict__ panel32, const int *__restrict__ info) {
    extern __shared__ double lds[];
    if (!info) {
        node_one(blockIdx.x, sorted, n, theta_sq, list, b_lo, b_hi, b_left, b_size, b_pre, tab, t,
                 mrec, rec32, link, panel32, lds);
        return;
    }
    const int count = dev_node_count(info);
    for (int s = blockIdx.x; s < count; s += gridDim.x) {
        wave_sync();  // the previous node's LDS reads are done
        node_one(s, sorted, n, theta_sq, list, b_lo, b_hi, b_left, b_size, b_pre, tab, t, mrec,
                 rec32, link, panel32, lds);
    }
}

// node_kernel for the presets' orders with kNodeGroup nodes per wavefront:
// lane group g (kGroupLanes lanes) writes node list[G*blockIdx + g]'s
// records and positions and, for a leaf of up to kGroupRowsMax sources,
// forms its rows kGroupLanes at a time; lane g*kGroupLanes + j then adds
// w_j over those rows in row order.  A larger leaf is weighted by the whole
// wavefront afterwards.  Per node the arithmetic and its order are those of
// node_kernel, so records and weights are the same bit for bit; the gain is
// a quarter of the workgroups, each one dependent-load chain.
constexpr int kGroupLanes = 16;
constexpr int kNodeGroup = kBaryLanes / kGroupLanes;
constexpr int kGroupRowsMax = 2 * kGroupLanes;

__host__ __device__ inline int node_group_lds_doubles(int n) {
    return bary_lds_doubles(n) + kNodeGroup * n;
}

template <int NT>
__device__ __forceinline__ void node_group_one(
    int grp, const double *__restrict__ sorted, int count, double theta_sq,
    const int *__restrict__ list, const int *__restrict__ b_lo, const int *__restrict__ b_hi,
    const int *__restrict__ b_left, const int *__restrict__ b_size, const int *__restrict__ b_pre,
    const double *__restrict__ tab, const TreeArrays &t, int2 *__restrict__ mrec,
    float4 *__restrict__ rec32, int4 *__restrict__ link, float *__restrict__ panel32,
    double *lds) {
    static_assert(NT <= kGroupLanes, "one lane per weight in a lane group");
    const int lane = threadIdx.x;
    const int g = lane / kGroupLanes, r = lane - g * kGroupLanes;
    const int slot = grp * kNodeGroup + g;
    const BaryLds bl = bary_lds(lds, NT);
    double *pos = lds + bary_lds_doubles(NT);  // kNodeGroup x NT positions
    // The loads that do not depend on the node go out first, and a leaf's
    // rows go out with its range ends: three dependent round trips per
    // wavefront (list, slot arrays, sources).
    double lr[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) lr[j] = tab[NT + j];
    const double tab_r = (r < NT) ? tab[r] : 0.0;
    int k = 0, a = 0, cnt = 0;
    bool leaf = false;
    double xrow0 = 0.0, xrow1 = 0.0;  // this lane's rows of a leaf: a + r, a + kGroupLanes + r
    if (slot < count) {
        const int v = list[slot];
        k = b_pre[v];
        a = b_lo[v];
        const int b = b_hi[v], l = b_left[v], size = b_size[v];
        if (l < 0) {
            if (r < b - a) xrow0 = sorted[a + r];
            if (kGroupLanes + r < b - a) xrow1 = sorted[a + kGroupLanes + r];
        }
        const int right = (l < 0) ? -1 : b_pre[l + 1];
        const double lo = sorted[a], hi = sorted[b - 1];
        const double hw = 0.5 * (hi - lo);
        const double acc = (hw * hw) / theta_sq;
        if (r == 0) {
            t.lo[k] = lo;
            t.hi[k] = hi;
            t.acc[k] = acc;
            t.lo_idx[k] = a;
            t.hi_idx[k] = b;
            t.right[k] = right;
            t.esc[k] = k + size;
            mrec[v] = make_int2(k, right);
            rec32[k] = make_float4((float)lo, (float)hi, (float)acc, 0.0f);
            link[k] = make_int4(right, k + size, a, b);
        }
        const double c = 0.5 * (lo + hi), hl = 0.5 * (hi - lo);
        if (r < NT) {
            const double x = c + hl * tab_r;
            t.nodes[(size_t)k * NT + r] = x;
            panel32[(size_t)k * 2 * NT + r] = (float)x;
            pos[g * NT + r] = x;
        }
        leaf = right < 0;
        cnt = b - a;
    }
    const bool small = leaf && cnt <= kGroupRowsMax;
    const unsigned long long any_small = __ballot(small);
    const unsigned long long big = __ballot(leaf && !small);
    if (!any_small && !big) return;
    wave_sync();
    if (any_small) {
        double tr[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) tr[j] = pos[g * NT + j];
        const int chunks = __ballot(small && cnt > kGroupLanes) ? 2 : 1;
        double wacc = 0.0;
        for (int ch = 0; ch < chunks; ++ch) {
            const int rr = ch * kGroupLanes + r;
            int hit = -2;  // as in bary_accumulate
            double m = 0.0;
            if (small && rr < cnt) {
                const double x = (ch == 0) ? xrow0 : xrow1;
                const double mass = 1.0;
                double s = 0.0;
                hit = -1;
#pragma unroll
                for (int j = 0; j < NT; ++j) {
                    const double d = x - tr[j];
                    const double q = lr[j] / d;
                    bl.sq[lane * NT + j] = q;
                    s += q;
                    hit = (hit < 0 && d == 0.0) ? j : hit;
                }
                m = (hit >= 0) ? mass : mass * (1.0 / s);
            }
            bl.sm[lane] = m;
            bl.shit[lane] = hit;
            wave_sync();
            const int rows = small ? min(kGroupLanes, cnt - ch * kGroupLanes) : 0;
            if (r < NT && rows > 0)
                wacc = bary_rows(wacc, g * kGroupLanes, rows, r, NT, bl.sq, bl.sm, bl.shit);
            wave_sync();  // the next chunk overwrites sq, sm, shit
        }
        if (small && r < NT) {
            t.w[(size_t)k * NT + r] = wacc;
            panel32[(size_t)k * 2 * NT + NT + r] = (float)wacc;
        }
    }
    // Leaves above kGroupRowsMax sources: the whole wavefront, one at a time.
    for (int gg = 0; gg < kNodeGroup; ++gg) {
        if (!((big >> (gg * kGroupLanes)) & 1ull)) continue;
        const int src = gg * kGroupLanes;
        const int kg = __shfl(k, src, kBaryLanes), ag = __shfl(a, src, kBaryLanes);
        const int cg = __shfl(cnt, src, kBaryLanes);
        const double *xs = sorted + ag;
        auto row = [xs](int rr, double &x, double &mass) {
            x = xs[rr];
            mass = 1.0;
        };
        bary_accumulate<NT>(row, cg, pos + gg * NT, tab + NT, t.w + (size_t)kg * NT,
                            panel32 + (size_t)kg * 2 * NT + NT, NT, bl.sq, bl.sm, bl.shit);
        wave_sync();  // the next leaf overwrites sq, sm, shit
    }
}

// The node count is `count`, or the info words' count when `info` is set
// (a build whose sizes stay on the device): the workgroups then take groups
// blockIdx.x, blockIdx.x + gridDim.x, ...
template <int NT>
__global__ __launch_bounds__(kBaryLanes) void node_group_kernel(
    const double *__restrict__ sorted, int count, double theta_sq, const int *__restrict__ list,
    const int *__restrict__ b_lo, const int *__restrict__ b_hi, const int *__restrict__ b_left,
    const int *__restrict__ b_size, const int *__restrict__ b_pre, const double *__restrict__ tab,
    TreeArrays t, int2 *__restrict__ mrec, float4 *__restrict__ rec32, int4 *__restrict__ link,
    float *__restrict__ panel32, const int *__restrict__ info) {
    extern __shared__ double lds[];
    if (!info) {
        node_group_one<NT>(blockIdx.x, sorted, count, theta_sq, list, b_lo, b_hi, b_left, b_size,
                           b_pre, tab, t, mrec, rec32, link, panel32, lds);
        return;
    }
    const int n_nodes = dev_node_count(info);
    for (int grp = blockIdx.x; grp * kNodeGroup < n_nodes; grp += gridDim.x) {
        wave_sync();  // the previous group's LDS reads are done
        node_group_one<NT>(grp, sorted, n_nodes, theta_sq, list, b_lo, b_hi, b_left, b_size,
                           b_pre, tab, t, mrec, rec32, link, panel32, lds);
    }
}

// Weights of the node in slot v from its children's: the 2n rows are the
// children's Chebyshev nodes with their weights as masses, left child first
// (merge_all_weights, chebcode.rs:474-506).  One wavefront; leaves return.
__device__ __forceinline__ void merge_node(int v, int n, const int2 *__restrict__ mrec,
                                           const double *__restrict__ lam, const TreeArrays &t,
                                           float *__restrict__ panel32, double *lds_wave) {
    const int2 m = mrec[v];
    if (m.y < 0) return;
    const int k = m.x;
    const BaryLds bl = bary_lds(lds_wave, n);
    const size_t lb = (size_t)(k + 1) * n, rb = (size_t)m.y * n;
    const double *nodes = t.nodes;
    const double *w = t.w;
    auto row = [=](int r, double &x, double &mass) {
        const size_t o = (r < n) ? lb + r : rb + (r - n);
        x = nodes[o];
        mass = w[o];
    };
    bary_any(row, 2 * n, t.nodes + (size_t)k * n, lam, t.w + (size_t)k * n,
                    panel32 + (size_t)k * 2 * n + n, n, bl.sq, bl.sm, bl.shit);
}

// Merges of the nodes in slots lb + [base, base + G) that are below lb + cnt,
// G = 64 / 2NT nodes per wavefront: lane group g (2NT lanes) forms node g's
// rows, then lane g*NT + j adds w_j of node g over that node's rows in row
// order.  Per node this is the arithmetic of the one-node accumulation in
// the same order, so the weights are the same bit for bit.
template <int NT>
__device__ __forceinline__ void merge_group(int lb, int base, int cnt,
                                            const int2 *__restrict__ mrec,
                                            const double *__restrict__ lam, const TreeArrays &t,
                                            float *__restrict__ panel32, double *lds_wave) {
    constexpr int R = 2 * NT, G = kBaryLanes / R;
    const int lane = threadIdx.x & (kBaryLanes - 1);
    const int g = lane / R, r = lane - g * R;
    const BaryLds bl = bary_lds(lds_wave, NT);
    int2 m = make_int2(0, -1);
    if (g < G && base + g < cnt) m = mrec[lb + base + g];
    int hit = -2;  // as in bary_accumulate
    double mm = 0.0;
    if (m.y >= 0) {
        const double *tp = t.nodes + (size_t)m.x * NT;
        double tr[NT], lr[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            tr[j] = tp[j];
            lr[j] = lam[j];
        }
        const size_t o = (r < NT) ? (size_t)(m.x + 1) * NT + r : (size_t)m.y * NT + (r - NT);
        const double x = t.nodes[o], mass = t.w[o];
        if (mass != 0.0) {
            double sum = 0.0;
            hit = -1;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const double d = x - tr[j];
                const double q = lr[j] / d;
                bl.sq[lane * NT + j] = q;
                sum += q;
                hit = (hit < 0 && d == 0.0) ? j : hit;
            }
            mm = (hit >= 0) ? mass : mass * (1.0 / sum);
        }
    }
    bl.sm[lane] = mm;
    bl.shit[lane] = hit;
    wave_sync();
    const int ga = lane / NT, j = lane - ga * NT;
    const int src = min(ga, G - 1) * R;  // first lane of group ga
    const int kx = __shfl(m.x, src, kBaryLanes), ky = __shfl(m.y, src, kBaryLanes);
    if (ga < G && ky >= 0) {
        const double acc = bary_rows(0.0, ga * R, R, j, NT, bl.sq, bl.sm, bl.shit);
        t.w[(size_t)kx * NT + j] = acc;
        panel32[(size_t)kx * 2 * NT + NT + j] = (float)acc;
    }
    wave_sync();  // the next group overwrites sq, sm, shit
}

// Nodes merged at once by one wavefront for order n (1 outside the presets).
__host__ __device__ inline int merge_group_size(int n) {
    return (n == 8 || n == 11 || n == 12) ? kBaryLanes / (2 * n) : 1;
}

// The nodes in slots [lb, lb + cnt), wavefront wv of `waves` taking groups
// wv, wv + waves, ...
template <int NT>
__device__ __forceinline__ void merge_level_t(int lb, int cnt, int wv, int waves,
                                              const int2 *__restrict__ mrec,
                                              const double *__restrict__ lam, const TreeArrays &t,
                                              float *__restrict__ panel32, double *lds_wave) {
    constexpr int G = kBaryLanes / (2 * NT);
    for (int c = wv * G; c < cnt; c += waves * G)
        merge_group<NT>(lb, c, cnt, mrec, lam, t, panel32, lds_wave);
}

__device__ __forceinline__ void merge_level(int n, int lb, int cnt, int wv, int waves,
                                            const int2 *__restrict__ mrec,
                                            const double *__restrict__ lam, const TreeArrays &t,
                                            float *__restrict__ panel32, double *lds_wave) {
    switch (n) {
    case 8: merge_level_t<8>(lb, cnt, wv, waves, mrec, lam, t, panel32, lds_wave); break;
    case 11: merge_level_t<11>(lb, cnt, wv, waves, mrec, lam, t, panel32, lds_wave); break;
    case 12: merge_level_t<12>(lb, cnt, wv, waves, mrec, lam, t, panel32, lds_wave); break;
    default:
        for (int i = wv; i < cnt; i += waves) merge_node(lb + i, n, mrec, lam, t, panel32, lds_wave);
        break;
    }
}

constexpr int kMergeMaxThreads = 1024;
constexpr int kMergeSubWaves = 8;   // wavefronts per subtree workgroup
constexpr size_t kSubLdsMax = 81920;  // LDS of a subtree workgroup (two per CU)
constexpr int kMergeTopRounds = 1;  // a top level of at most this many groups per wavefront
                                    // merges in the one-workgroup kernel

// One top level (slots [lb, lb + cnt)), one wavefront per group of nodes.
__global__ __launch_bounds__(kBaryLanes) void merge_range_kernel(int n, int lb, int cnt,
                                                                 const int2 *__restrict__ mrec,
                                                                 const double *__restrict__ lam,
                                                                 TreeArrays t,
                                                                 float *__restrict__ panel32) {
    extern __shared__ double lds[];
    merge_level(n, lb, cnt, blockIdx.x, gridDim.x, mrec, lam, t, panel32, lds);
}

// The subtree of frontier node blockIdx.x, its depths bottom-up (level l
// holds slots lv_start[l] + [0, lv_cnt[l])); the workgroup barrier orders
// the levels.
__global__ __launch_bounds__(kMergeMaxThreads) void merge_sub_kernel(
    int n, const int *__restrict__ lv_start, const int *__restrict__ lv_cnt,
    const int2 *__restrict__ mrec, const double *__restrict__ lam, TreeArrays t,
    float *__restrict__ panel32) {
    extern __shared__ double lds[];
    const int f = blockIdx.x;
    const int waves = blockDim.x / kBaryLanes, wv = threadIdx.x / kBaryLanes;
    double *mine = lds + (size_t)wv * bary_lds_doubles(n);
    int n_lv = 1;
    while (n_lv < kLv && lv_cnt[f * kLv + n_lv] > 0) ++n_lv;
    for (int l = n_lv - 2; l >= 0; --l) {
        merge_level(n, lv_start[f * kLv + l], lv_cnt[f * kLv + l], wv, waves, mrec, lam, t,
                    panel32, mine);
        __syncthreads();
    }
}

// Top levels [0, levels) bottom-up in one workgroup (level d holds the BFS
// slots [lvl[d], lvl[d+1]) recorded by the top walk).
__global__ __launch_bounds__(kMergeMaxThreads) void merge_top_kernel(
    int n, int levels, const int *__restrict__ info, const int2 *__restrict__ mrec,
    const double *__restrict__ lam, TreeArrays t, float *__restrict__ panel32) {
    extern __shared__ double lds[];
    const int waves = blockDim.x / kBaryLanes, wv = threadIdx.x / kBaryLanes;
    double *mine = lds + (size_t)wv * bary_lds_doubles(n);
    const int *lvl = info + kInfoTopLevel;
    for (int d = levels - 1; d >= 0; --d) {
        merge_level(n, lvl[d], lvl[d + 1] - lvl[d], wv, waves, mrec, lam, t, panel32, mine);
        __syncthreads();
    }
}

// Quotient tables of a merge level (used by the LDS forms of the merges).  A
// merge evaluates, for each of its 2n rows (a child's Chebyshev point x_r,
// with that child's weight as mass), the quotients q_rj = lam_j / (x_r - t_j)
// against the node's positions t_j, their sum s_r and the first exact hit,
// then adds the rows into the weights in row order: w_j = fma(mass_r *
// (1 / s_r), q_rj, w_j), or w_hit += mass_r, skipping rows of zero mass
// (merge_group).  Only the row sums depend on the level below.  So the whole
// workgroup first builds the tables of P nodes (slots [u0, u0 + P), one row
// per thread): q at ((u - u0) * 2n + r) * n + j, 1 / s_r and the hit at
// (u - u0) * 2n + r.  Each node's weights are then the 2n-long fused
// multiply-add chain alone, n lanes per node (lane j owns w_j).  The
// operations and their order are merge_group's: the weights are the same
// bit for bit.
__host__ __device__ constexpr int quot_table_doubles(int P, int n) {
    return P * 2 * n * n + P * 2 * n + (P * 2 * n + 1) / 2;
}
// Nodes whose tables fit `doubles` and whose chains take one pass of `waves`.
__host__ __device__ constexpr int quot_table_cap(int doubles, int n, int waves) {
    int P = 1;
    while (quot_table_doubles(P + 1, n) <= doubles && P + 1 <= waves * (kBaryLanes / n)) ++P;
    return P;
}

template <int NT>
struct QuotTables {
    double *q, *inv;
    int *hit;
    int u0;
};

template <int NT>
__device__ __forceinline__ QuotTables<NT> quot_tables(double *scratch, int u0, int P,
                                                      const int *s_left, const double *s_pos,
                                                      const double (&lr)[NT]) {
    constexpr int R = 2 * NT;
    QuotTables<NT> tb{scratch, scratch + (size_t)P * R * NT, nullptr, u0};
    tb.hit = reinterpret_cast<int *>(tb.inv + (size_t)P * R);
    for (int e = threadIdx.x; e < P * R; e += blockDim.x) {
        const int u = u0 + e / R, r = e % R;
        const int left = s_left[u];
        if (left < 0) continue;
        const double x = s_pos[(left + (r < NT ? 0 : 1)) * NT + (r < NT ? r : r - NT)];
        double sum = 0.0;
        int hit = -1;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const double d = x - s_pos[u * NT + j];
            const double q = lr[j] / d;
            tb.q[(size_t)e * NT + j] = q;
            sum += q;
            hit = (hit < 0 && d == 0.0) ? j : hit;
        }
        tb.inv[e] = 1.0 / sum;
        tb.hit[e] = hit;
    }
    return tb;
}

// The row-sum chains of the nodes in slots [lb, lb + cnt) (tables built),
// one pass: cnt <= wavefronts * (kBaryLanes / n).
template <int NT>
__device__ __forceinline__ void quot_chains(const QuotTables<NT> &tb, int lb, int cnt,
                                            const int *s_left, const int *s_k, double *s_w,
                                            const TreeArrays &t, float *__restrict__ panel32) {
    constexpr int R = 2 * NT, PG = kBaryLanes / NT;
    const int lane = threadIdx.x & (kBaryLanes - 1), pg = lane / NT, j = lane - pg * NT;
    const int i = (int)(threadIdx.x / kBaryLanes) * PG + pg;
    const int u = lb + i;
    const int left = (pg < PG && i < cnt) ? s_left[u] : -1;
    if (left < 0) return;
    const int e0 = (u - tb.u0) * R;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double mass = s_w[(left + (r < NT ? 0 : 1)) * NT + (r < NT ? r : r - NT)];
        const int h = tb.hit[e0 + r];
        const double f = __builtin_fma(mass * tb.inv[e0 + r], tb.q[(size_t)(e0 + r) * NT + j], acc),
                     a = acc + mass;
        acc = (mass == 0.0) ? acc : ((h == -1) ? f : ((h == j) ? a : acc));
    }
    s_w[u * NT + j] = acc;
    const int k = s_k[u];
    t.w[(size_t)k * NT + j] = acc;
    panel32[(size_t)k * 2 * NT + NT + j] = (float)acc;
}

// One level (slots [lb, lb + cnt)), pcap nodes per round of tables and chains.
template <int NT>
__device__ __forceinline__ void merge_level_quot(int lb, int cnt, int pcap, const int *s_left,
                                                 const int *s_k, const double *s_pos, double *s_w,
                                                 const double (&lr)[NT], const TreeArrays &t,
                                                 float *__restrict__ panel32, double *scratch) {
    for (int c0 = 0; c0 < cnt; c0 += pcap) {
        const int P = min(pcap, cnt - c0);
        const QuotTables<NT> tb = quot_tables<NT>(scratch, lb + c0, P, s_left, s_pos, lr);
        lds_barrier();
        quot_chains<NT>(tb, lb + c0, P, s_left, s_k, s_w, t, panel32);
        lds_barrier();  // the next round overwrites the tables
    }
}

// LDS form of the top merges for the presets' orders.  The slots of levels
// [0, levels] (level `levels` holds finished children) are staged in LDS once
// with their left-child slots, pre-order numbers, positions and weights;
// each level then reads and writes LDS only (weights also go to the tree and
// the f32 panel) behind an LDS-only barrier, merged through quotient tables
// (merge_level_quot).  The narrow top levels (the first kQNarrow slots, one
// or a few nodes per level, where the levels are a chain of latencies on one
// CU) get their tables in one pass.
constexpr int kTopLdsWaves = 16;          // the top-level merge workgroup
constexpr size_t kTopLdsWgMax = 163840;   // LDS of the one top-level merge workgroup
constexpr int kQNarrow = 40;              // top slots with quotient tables

__host__ __device__ inline size_t top_lds_bytes(int slots, int n, int waves) {
    return (size_t)slots * (2 * sizeof(int) + 2 * (size_t)n * sizeof(double)) +
           (size_t)waves * sizeof(double) * (size_t)bary_lds_doubles(n);
}

// One group of nodes merged from LDS (merge_sub_lds_kernel): merge_group's
// arithmetic with positions and weights staged in LDS.
template <int NT>
__device__ __forceinline__ void merge_group_lds(int lb, int base, int cnt,
                                                const int *__restrict__ s_left,
                                                const int *__restrict__ s_k, const double *s_pos,
                                                double *s_w, const double (&lr)[NT],
                                                const TreeArrays &t, float *__restrict__ panel32,
                                                double *lds_wave) {
    constexpr int R = 2 * NT, G = kBaryLanes / R;
    const int lane = threadIdx.x & (kBaryLanes - 1);
    const int g = lane / R, r = lane - g * R;
    const BaryLds bl = bary_lds(lds_wave, NT);
    const int u = lb + base + g;
    const bool ok = g < G && base + g < cnt && s_left[u] >= 0;
    int hit = -2;  // as in bary_accumulate
    double mm = 0.0;
    if (ok) {
        double tr[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) tr[j] = s_pos[u * NT + j];
        const int ch = s_left[u] + (r < NT ? 0 : 1), jj = (r < NT) ? r : r - NT;
        const double x = s_pos[ch * NT + jj], mass = s_w[ch * NT + jj];
        if (mass != 0.0) {
            double sum = 0.0;
            hit = -1;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const double d = x - tr[j];
                const double q = lr[j] / d;
                bl.sq[lane * NT + j] = q;
                sum += q;
                hit = (hit < 0 && d == 0.0) ? j : hit;
            }
            mm = (hit >= 0) ? mass : mass * (1.0 / sum);
        }
    }
    bl.sm[lane] = mm;
    bl.shit[lane] = hit;
    wave_sync();
    const int ga = lane / NT, j = lane - ga * NT;
    const int ua = lb + base + ga;
    if (ga < G && base + ga < cnt && s_left[ua] >= 0) {
        const double acc = bary_rows(0.0, ga * R, R, j, NT, bl.sq, bl.sm, bl.shit);
        s_w[ua * NT + j] = acc;
        const int k = s_k[ua];
        t.w[(size_t)k * NT + j] = acc;
        panel32[(size_t)k * 2 * NT + NT + j] = (float)acc;
    }
    wave_sync();  // the next group overwrites sq, sm, shit
}

// The top levels with internal nodes (top) and how many leading ones the
// one-workgroup LDS merge takes (small), from the info words: cheb_build's
// host rule, for a build whose sizes stay on the device.  Levels [small, top)
// are the wide ones, merged one level per launch.
__device__ __forceinline__ void top_split(const int *info, int n, int &top, int &small) {
    const int F = info[kInfoF], D = info[kInfoD];
    top = dev_build_failed(info) ? 0 : ((F > 0) ? D : max(D - 1, 0));
    const int *lvl = info + kInfoTopLevel;
    const int G = merge_group_size(n);
    small = 0;
    while (small < top && top_lds_bytes(lvl[small + 2], n, kTopLdsWaves) <= kTopLdsWgMax &&
           (lvl[small + 1] - lvl[small] + G - 1) / G <= kMergeTopRounds * kTopLdsWaves)
        ++small;
}

// Wide top levels of a build whose sizes stay on the device: launch k merges
// level top - 1 - k if it is a wide one (launches k = 0, 1, ... in stream
// order take the deepest first), kWideWaves wavefronts per workgroup striding
// over its groups as merge_range_kernel's workgroups do.  kWideLaunches such
// launches run; wide levels beyond them go to the one-workgroup kernel's
// global form.
constexpr int kWideWaves = 4;
constexpr int kWideLaunches = 5;

__global__ __launch_bounds__(kWideWaves * kBaryLanes) void merge_wide_kernel(
    int n, int k, const int *__restrict__ info, const int2 *__restrict__ mrec,
    const double *__restrict__ lam, TreeArrays t, float *__restrict__ panel32) {
    extern __shared__ double lds[];
    int top, small;
    top_split(info, n, top, small);
    const int d = top - 1 - k;
    if (d < small) return;  // uniform
    const int *lvl = info + kInfoTopLevel;
    const int waves = blockDim.x / kBaryLanes, wv = threadIdx.x / kBaryLanes;
    merge_level(n, lvl[d], lvl[d + 1] - lvl[d], blockIdx.x * waves + wv, gridDim.x * waves, mrec,
                lam, t, panel32, lds + (size_t)wv * bary_lds_doubles(n));
}

// levels < 0: a build whose sizes stay on the device (launched with
// kTopLdsWgMax bytes of LDS): the levels come from top_split, after the
// wide-level launches; when the wide levels outnumber those launches, the
// remaining levels are merged here in the global form (merge_level per level,
// the arithmetic of every form being merge_group's).
template <int NT>
__global__ __launch_bounds__(kTopLdsWaves * kBaryLanes) void merge_top_lds_kernel(
    int levels, const int *__restrict__ info, const int *__restrict__ b_left,
    const int2 *__restrict__ mrec, const double *__restrict__ lam, TreeArrays t,
    float *__restrict__ panel32) {
    extern __shared__ double lds[];
    const int *lvl = info + kInfoTopLevel;
    if (levels < 0) {
        int top, small;
        top_split(info, NT, top, small);
        if (top - small > kWideLaunches) {
            const int waves = blockDim.x / kBaryLanes, wv = threadIdx.x / kBaryLanes;
            double *mine = lds + (size_t)wv * bary_lds_doubles(NT);
            for (int d = top - kWideLaunches - 1; d >= 0; --d) {
                merge_level(NT, lvl[d], lvl[d + 1] - lvl[d], wv, waves, mrec, lam, t, panel32,
                            mine);
                __syncthreads();
            }
            return;
        }
        levels = small;
        if (levels == 0) return;
    }
    const int S = lvl[levels + 1], fin = lvl[levels];
    const int waves = blockDim.x / kBaryLanes;
    double *s_pos = lds;
    double *s_w = s_pos + (size_t)S * NT;
    double *bary = s_w + (size_t)S * NT;
    int *s_left = reinterpret_cast<int *>(bary + (size_t)waves * bary_lds_doubles(NT));
    int *s_k = s_left + S;
    double lr[NT];  // barycentric weights, loaded once
#pragma unroll
    for (int j = 0; j < NT; ++j) lr[j] = lam[j];
    for (int v = threadIdx.x; v < S; v += blockDim.x) {
        const int2 m = mrec[v];
        s_k[v] = m.x;
        s_left[v] = (v < fin && m.y >= 0) ? b_left[v] : -1;  // children of slot v: slots
    }
    lds_barrier();
    for (int e = threadIdx.x; e < S * NT; e += blockDim.x) {
        const int v = e / NT, j = e - v * NT;
        const size_t o = (size_t)s_k[v] * NT + j;
        s_pos[e] = t.nodes[o];
        s_w[e] = t.w[o];  // finished weights; merged nodes overwrite theirs
    }
    lds_barrier();
    constexpr int kScratch = kTopLdsWaves * (kBaryLanes * NT + kBaryLanes + kBaryLanes / 2);
    static_assert(quot_table_doubles(kQNarrow, NT) <= kScratch,
                  "the narrow levels' quotient tables fit the merge scratch");
    constexpr int pcap = quot_table_cap(kScratch, NT, kTopLdsWaves);
    // levels [0, dq) hold the first lvl[dq] <= kQNarrow slots: their tables
    // are built at once; each wider level builds its own, pcap nodes at a time
    int dq = 0;
    while (dq < levels && lvl[dq + 1] <= kQNarrow) ++dq;
    for (int d = levels - 1; d >= dq; --d)
        merge_level_quot<NT>(lvl[d], lvl[d + 1] - lvl[d], pcap, s_left, s_k, s_pos, s_w, lr, t,
                             panel32, bary);
    if (dq == 0) return;
    const QuotTables<NT> tb = quot_tables<NT>(bary, 0, lvl[dq], s_left, s_pos, lr);
    lds_barrier();
    for (int d = dq - 1; d >= 0; --d) {
        quot_chains<NT>(tb, lvl[d], lvl[d + 1] - lvl[d], s_left, s_k, s_w, t, panel32);
        lds_barrier();
    }
}

// LDS form of the subtree merges for the presets' orders: subtree f (its root
// and the contiguous slots of its levels below) staged in LDS with positions,
// weights, left-child slots and pre-order numbers, local node 0 being the
// root and slot v >= base1 the local node v - base1 + 1.  A subtree with more than `cap` nodes takes the
// global path in the same workgroup.
template <int NT>
__global__ __launch_bounds__(kMergeSubWaves * kBaryLanes) void merge_sub_lds_kernel(
    const int *__restrict__ lv_start, const int *__restrict__ lv_cnt,
    const int *__restrict__ sub_nodes, const int *__restrict__ b_left,
    const int2 *__restrict__ mrec, const double *__restrict__ lam, TreeArrays t,
    float *__restrict__ panel32, int cap, const int *__restrict__ info) {
    extern __shared__ double lds[];
    constexpr int G = kBaryLanes / (2 * NT);
    const int f = blockIdx.x;
    // a grid of kMaxFrontier workgroups when the frontier size stays on the
    // device (info set): those past it exit
    if (info && (f >= info[kInfoF] || dev_build_failed(info))) return;
    const int waves = blockDim.x / kBaryLanes, wv = threadIdx.x / kBaryLanes;
    int n_lv = 1;
    while (n_lv < kLv && lv_cnt[f * kLv + n_lv] > 0) ++n_lv;
    const int S = 1 + sub_nodes[f];
    double *s_pos = lds;
    double *s_w = s_pos + (size_t)cap * NT;
    double *bary = s_w + (size_t)cap * NT;
    double *mine = bary + (size_t)wv * bary_lds_doubles(NT);
    if (S > cap) {  // uniform per workgroup
        for (int l = n_lv - 2; l >= 0; --l) {
            merge_level(NT, lv_start[f * kLv + l], lv_cnt[f * kLv + l], wv, waves, mrec, lam, t,
                        panel32, mine);
            __syncthreads();
        }
        return;
    }
    int *s_left = reinterpret_cast<int *>(bary + (size_t)waves * bary_lds_doubles(NT));
    int *s_k = s_left + cap;
    const int root = lv_start[f * kLv], base1 = (n_lv > 1) ? lv_start[f * kLv + 1] : 0;
    double lr[NT];  // barycentric weights, loaded once
#pragma unroll
    for (int j = 0; j < NT; ++j) lr[j] = lam[j];
    for (int u = threadIdx.x; u < S; u += blockDim.x) {
        const int v = (u == 0) ? root : base1 + u - 1;
        const int2 m = mrec[v];
        s_k[u] = m.x;
        s_left[u] = (m.y >= 0) ? b_left[v] - base1 + 1 : -1;
    }
    lds_barrier();
    for (int e = threadIdx.x; e < S * NT; e += blockDim.x) {
        const int u = e / NT, j = e - u * NT;
        const size_t o = (size_t)s_k[u] * NT + j;
        s_pos[e] = t.nodes[o];
        s_w[e] = t.w[o];  // leaf weights; merged nodes overwrite theirs
    }
    lds_barrier();
    // Levels here are wide enough that one group's divisions and column
    // sums, wave-synchronous, beat quotient tables built between workgroup
    // barriers (36.0 against 31.6 us at 1e6, FAST).
    for (int l = n_lv - 2; l >= 0; --l) {
        const int lb = (l == 0) ? 0 : lv_start[f * kLv + l] - base1 + 1;
        const int cnt = lv_cnt[f * kLv + l];
        for (int c = wv * G; c < cnt; c += waves * G)
            merge_group_lds<NT>(lb, c, cnt, s_left, s_k, s_pos, s_w, lr, t, panel32, mine);
        lds_barrier();
    }
}

// ---- evaluation --------------------------------------------------------------

// Passes x through an asm move whose other inputs are loaded node data:
// the loads then complete before x's use (the leaf test of a visit), and
// the compiler cannot defer them behind that test into a second round trip.
// The asm has no side effects and touches no memory.
template <class T>
__device__ __forceinline__ int hold2(int x, T a, T b) {
    int y;
    asm("s_mov_b32 %0, %1" : "=s"(y) : "s"(x), "s"(a), "s"(b));
    return y;
}

struct EvalView {
    const double *sorted;
    const float *s_hi, *s_lo;
    const double *lo, *hi, *acc, *nodes, *w;
    const int4 *link;
    const float4 *rec32;
    const float *panel32;
    const int32_t *perm;
    int N, n;
    // set for a tree whose build kept its sizes on the device: the walks
    // read the node count from these info words
    const int *info = nullptr;
};

// The node count of a walk: v.N, or the build's info words for a tree whose
// sizes stayed on the device.  `bad` marks such a build that found its input
// not ascending (or its node slots exhausted): the outputs are then NaN.
__device__ __forceinline__ int walk_nodes(const EvalView &v, bool &bad) {
    if (!v.info) {
        bad = false;
        return v.N;
    }
    bad = dev_build_failed(v.info);
    return dev_node_count(v.info);
}

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
__device__ __forceinline__ bool accept_f64_of(double lo, double hi, double acc, double x,
                                              double eta2) {
    const double cl = fmax(fmax(lo - x, x - hi), 0.0);
    const double d_sq = __builtin_fma(cl, cl, eta2);
    return acc < d_sq;
}

__device__ __forceinline__ bool accept_f64(const EvalView &v, int node, double x, double eta2) {
    return accept_f64_of(v.lo[node], v.hi[node], v.acc[node], x, eta2);
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
// NT > 0 (the wave walk at order NT = n): a visit's link, interval,
// acceptance radius and panel are loaded at its start, from the node index
// alone, instead of three dependent loads; the arithmetic and its order are
// unchanged.  The compiler still issues the panel loads behind the leaf
// test, a second round trip for an internal node.
// kHold (NT > 0): asm moves that consume every loaded value keep the panel
// loads ahead of the leaf test, so a visit is one round trip.  That pays
// where the walk is latency-bound (fewer than about two wavefronts per SIMD)
// and costs 2-5 % where it is issue-bound.
template <bool kWave, int NT = 0, bool kHold = false>
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
    bool bad;
    const int N = walk_nodes(v, bad);
    int skip = valid ? 0 : N;
    int node = 0;
    constexpr int kP = (NT > 0) ? NT : 1;
    while (node < N) {
        if constexpr (kWave) node = __builtin_amdgcn_readfirstlane(node);
        int4 lk = v.link[node];  // right, esc, lo_idx, hi_idx
        double rlo = 0.0, rhi = 0.0, racc = 0.0, pts[kP], pws[kP];
        if constexpr (NT > 0) {
            rlo = v.lo[node];
            rhi = v.hi[node];
            racc = v.acc[node];
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                pts[j] = v.nodes[(size_t)node * NT + j];
                pws[j] = v.w[(size_t)node * NT + j];
            }
            if constexpr (kHold) {
                int leaf_bits = hold2(lk.x, rlo, rhi);
                leaf_bits = hold2(leaf_bits, racc, racc);
#pragma unroll
                for (int j = 0; j < NT; ++j) leaf_bits = hold2(leaf_bits, pts[j], pws[j]);
                lk.x = leaf_bits;
            }
        }
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
        const bool accept = (NT > 0) ? accept_f64_of(rlo, rhi, racc, x, eta2)
                                     : accept_f64(v, node, x, eta2);
        if (active && accept && NT > 0) {
            double afr0 = 0.0, afr1 = 0.0, afi0 = 0.0, afi1 = 0.0;
            // the terms' reciprocals four (and a last two) per v_rcp_f64
            // (rcp_many_f64)
            constexpr int kPair = (kP >= 2) ? (kP & ~1) : 2;  // (NT = 0 never runs this)
            double dd[kPair], den[kPair], inv[kPair];
#pragma unroll
            for (int j = 0; j < kPair; ++j) {
                dd[j] = x - pts[j];
                den[j] = __builtin_fma(dd[j], dd[j], eta2);
            }
            rcp_many_f64<kPair>(den, inv);
#pragma unroll
            for (int j = 0; j + 2 <= kP; j += 2) {
                const double q0 = pws[j] * inv[j];
                const double q1 = pws[j + 1] * inv[j + 1];
                afr0 = __builtin_fma(q0, dd[j], afr0);
                afr1 = __builtin_fma(q1, dd[j + 1], afr1);
                afi0 += q0;
                afi1 += q1;
            }
            double fr = afr0 + afr1, fi = afi0 + afi1;
            if (kP % 2) {
                const double dj = x - pts[kP - 1];
                const double q = pws[kP - 1] * rcp_f64(dj * dj + eta2);
                fr += q * dj;
                fi += q;
            }
            re += fr;
            im += eta * fi;
            skip = lk.y;
        } else if (active && accept) {
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
    out_re[o] = bad ? __builtin_nan("") : re * scale;
    out_im[o] = bad ? __builtin_nan("") : im * scale;
}

// NT > 0 as in eval_f64_kernel: record and f32 panel loaded with the link.
template <bool kWave, int NT = 0, bool kHold = false>
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
    bool bad;
    const int N = walk_nodes(v, bad);
    int skip = valid ? 0 : N;
    int node = 0;
    constexpr int kP = (NT > 0) ? NT : 1;
    while (node < N) {
        if constexpr (kWave) node = __builtin_amdgcn_readfirstlane(node);
        int4 lk = v.link[node];
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
        float pts[kP], pws[kP];
        if constexpr (NT > 0) {
            rec = v.rec32[node];
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                pts[j] = v.panel32[(size_t)node * 2 * NT + j];
                pws[j] = v.panel32[(size_t)node * 2 * NT + NT + j];
            }
            if constexpr (kHold) {
                int leaf_bits = hold2(lk.x, rec.x, rec.y);
                leaf_bits = hold2(leaf_bits, rec.z, rec.z);
#pragma unroll
                for (int j = 0; j < NT; ++j) leaf_bits = hold2(leaf_bits, pts[j], pws[j]);
                lk.x = leaf_bits;
            }
        }
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
        const bool accept = accept_fast((NT > 0) ? rec : v.rec32[node], x, eta2);
        if (active && accept && NT > 0) {
            // term j goes to accumulator j mod 4 (static indices)
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const float d = x - pts[j];
                const float q = pws[j] * __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2));
                far_re[j % 4] = __builtin_fmaf(q, d, far_re[j % 4]);
                far_im[j % 4] = far_im[j % 4] + q;
            }
            skip = lk.y;
        } else if (active && accept) {
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
    out_re[o] = bad ? __builtin_nan("") : re * scale;
    out_im[o] = bad ? __builtin_nan("") : im * scale;
}

// Sparse queries (a grid across the spectrum, fewer than kSparseMax points):
// one wavefront per query.  The wavefront walks its query's path in the
// reference's visit order; the lanes split a leaf's sources (lane l takes
// sources l, l + 64, ...) and a panel's terms (lane j takes term j), each
// into its own accumulators, and a fixed butterfly adds the 64 lanes at the
// end.  The per-lane walk leaves 63 of 64 lanes of a wavefront waiting on
// the one whose near field is largest, with a few hundred queries on a few
// CUs; this form spreads them over the chip.  Only the summation order
// differs from the per-lane walk (both within the presets' tolerances).
constexpr int kSparseMax = 16384;
constexpr int kSparseWaves = 4;  // queries per workgroup

__device__ __forceinline__ double wave_sum_fixed(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = ((threadIdx.x & o) ? w + v : v + w);  // the same operand order on both lanes
    }
    return v;
}

__global__ __launch_bounds__(kSparseWaves * 64) void eval_f64_sparse_kernel(
    EvalView v, const double *__restrict__ query, int nq, double eta, double scale,
    double *__restrict__ out_re, double *__restrict__ out_im) {
    const int q = blockIdx.x * kSparseWaves + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= nq) return;  // wavefront-uniform
    const double x = query[q];
    const double eta2 = eta * eta;
    const int n = v.n;
    double lr = 0.0, li = 0.0, fr = 0.0, fi = 0.0;
    bool bad;
    const int N = walk_nodes(v, bad);
    int node = 0;
    while (node < N) {
        node = __builtin_amdgcn_readfirstlane(node);
        const int4 lk = v.link[node];
        if (lk.x < 0) {
            for (int k = lk.z + lane; k < lk.w; k += 64) {
                const double d = x - v.sorted[k];
                const double inv = rcp_f64(__builtin_fma(d, d, eta2));
                lr = __builtin_fma(d, inv, lr);
                li += inv;
            }
            node = lk.y;
            continue;
        }
        if (accept_f64(v, node, x, eta2)) {
            for (int j = lane; j < n; j += 64) {
                const double d = x - v.nodes[(size_t)node * n + j];
                const double qv = v.w[(size_t)node * n + j] * rcp_f64(__builtin_fma(d, d, eta2));
                fr = __builtin_fma(qv, d, fr);
                fi += qv;
            }
            node = lk.y;
        } else {
            node = node + 1;
        }
    }
    const double re = wave_sum_fixed(lr + fr), im = wave_sum_fixed(li + fi);
    if (lane == 0) {
        out_re[q] = bad ? __builtin_nan("") : re * scale;
        out_im[q] = bad ? __builtin_nan("") : (eta * im) * scale;
    }
}

__global__ __launch_bounds__(kSparseWaves * 64) void eval_fast_sparse_kernel(
    EvalView v, const double *__restrict__ query, int nq, double eta, double scale,
    double *__restrict__ out_re, double *__restrict__ out_im) {
    const int q = blockIdx.x * kSparseWaves + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= nq) return;  // wavefront-uniform
    const double xd = query[q];
    const float x = (float)xd;
    const float x_lo = (float)(xd - (double)x);
    const float e32 = (float)eta;
    const float eta2 = e32 * e32;
    const float eta2n = (float)(eta * eta);
    const int n = v.n;
    double leaf_re = 0.0, leaf_inv = 0.0;
    float far_re = 0.f, far_im = 0.f;
    bool bad;
    const int N = walk_nodes(v, bad);
    int node = 0;
    while (node < N) {
        node = __builtin_amdgcn_readfirstlane(node);
        const int4 lk = v.link[node];
        if (lk.x < 0) {
            // f32 near field on the hi/lo split, folded into f64 per leaf
            float ar = 0.f, ai = 0.f;
            for (int k = lk.z + lane; k < lk.w; k += 64) {
                const float d = (x - v.s_hi[k]) + (x_lo - v.s_lo[k]);
                const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2n));
                ar = __builtin_fmaf(d, r, ar);
                ai += r;
            }
            leaf_re += (double)ar;
            leaf_inv += (double)ai;
            node = lk.y;
            continue;
        }
        if (accept_fast(v.rec32[node], x, eta2)) {
            const float *ts = v.panel32 + (size_t)node * 2 * n;
            for (int j = lane; j < n; j += 64) {
                const float d = x - ts[j];
                const float qv = ts[n + j] * __builtin_amdgcn_rcpf(__builtin_fmaf(d, d, eta2));
                far_re = __builtin_fmaf(qv, d, far_re);
                far_im = far_im + qv;
            }
            node = lk.y;
        } else {
            node = node + 1;
        }
    }
    const double re = wave_sum_fixed(leaf_re + (double)far_re);
    const double li = wave_sum_fixed(leaf_inv), fi = wave_sum_fixed((double)far_im);
    if (lane == 0) {
        out_re[q] = bad ? __builtin_nan("") : re * scale;
        out_im[q] = bad ? __builtin_nan("") : (eta * li + eta * fi) * scale;
    }
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

// ---- local expansions (M2L / L2P) at the tree's own points ------------------
//
// No counterpart in the reference: its far field is the per-query panel sum
// (M2P, chebcode.rs:607-711), and its notes leave a Chebyshev M2L to future
// work (docs/internals.md:580-599).  Targets are the topmost nodes no wider
// than kLocalWidth * eta, plus the leaves wider than that (which share one
// walk and sum the accepted panels at each query).  Each target T walks the
// tree once, with the reference's acceptance test (accept_f64_of) at T's gap
// to the panel, i.e. at the query of T nearest to it: a panel accepted for T
// is accepted by the reference for every query of T, so the far field is
// never coarser than the reference's.
// Accepted panels are evaluated at T's n Chebyshev points y_k (M2L: n^2
// terms per panel, shared by all of T's queries, instead of n per query);
// leaves the walk reaches are listed and summed exactly per query, as the
// reference sums every leaf it reaches.  Each query then interpolates T's
// far field with the barycentric formula on the Chebyshev points of the
// second kind (L2P, n terms) and adds the listed leaves' exact sums.  The far
// field sum_j w_j / (x - t_j - i eta) is analytic within eta of the real
// axis, so on an interval of width <= eta/4 its degree n-1 interpolant errs
// by about rho^-n, rho = 8 + sqrt(65) (5e-14 at n = 11).  A target whose
// lists overflow gets the per-query walk.
constexpr double kLocalWidth = 0.25;
constexpr int kLocalNear = 16;

// The reference traversal for one query (the per-lane walk of
// eval_f64_kernel<false>, same arithmetic and order).
__device__ __forceinline__ void walk_one_f64(const EvalView &v, double x, double eta,
                                             double eta2, double &re, double &im) {
    const int n = v.n;
    re = im = 0.0;
    int node = 0;
    while (node < v.N) {
        const int4 lk = v.link[node];
        if (lk.x < 0) {
            double lr, li;
            leaf_f64(v.sorted + lk.z, lk.w - lk.z, x, eta2, lr, li);
            re += lr;
            im += eta * li;
            node = lk.y;
            continue;
        }
        if (accept_f64(v, node, x, eta2)) {
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
            node = lk.y;
        } else {
            node = node + 1;
        }
    }
}

__global__ void local_zero_kernel(int *count) {
    if (threadIdx.x < 3 && blockIdx.x == 0) count[threadIdx.x] = 0;
}

// Target list: node T (code T) is a target when it is no wider than wmax and
// its parent is wider (or T is the root); a leaf wider than wmax under such
// parents is a fallback target (code ~T).  Every source lies in exactly one
// target.  The order of the list does not matter: targets are independent.
// Appends this thread's n (0..2) entries c0, c1 to a list with one atomic
// per workgroup (THREADS threads): ballots give each wavefront's offsets, a
// workgroup scan the wavefronts', and thread 0 reserves the whole block.
// Many wavefronts adding to one counter serialise at the L2.
template <int THREADS>
__device__ __forceinline__ void block_append(int n, int c0, int c1, int *__restrict__ list,
                                             int *__restrict__ counter, int *s_part) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t m1 = __ballot(n >= 1), m2 = __ballot(n >= 2);
    if (lane == 0) s_part[wid] = __popcll(m1) + __popcll(m2);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < THREADS / 64; ++w) {
            const int c = s_part[w];
            s_part[w] = tot;  // the wavefront's offset within the block
            tot += c;
        }
        s_part[THREADS / 64] = (tot > 0) ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const int at = s_part[THREADS / 64] + s_part[wid] + __popcll(m1 & below) + __popcll(m2 & below);
    if (n >= 1) list[at] = c0;
    if (n >= 2) list[at + 1] = c1;
    __syncthreads();  // s_part is reused by the next call
}

// Target lists: node T is a target (list, count[0]) when it is no wider than
// wmax and its parent is wider (or T is the root); a leaf wider than wmax
// under such parents is a wide target (wide_list, count[2]).  Every source
// lies in exactly one target.  List order does not matter: targets are
// independent.
constexpr int kPlanThreads = 1024;

__global__ __launch_bounds__(kPlanThreads) void local_plan_kernel(EvalView v, double wmax,
                                                                 int *__restrict__ list,
                                                                 int *__restrict__ wide_list,
                                                                 int *__restrict__ count) {
    __shared__ int s_part[kPlanThreads / 64 + 1];
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    int t[2] = {0, 0}, wd[2] = {0, 0}, nt = 0, nw = 0;
    if (node < v.N) {
        const int4 lk = v.link[node];
        const double w = v.hi[node] - v.lo[node];
        if (node == 0) {
            if (w <= wmax) t[nt++] = 0;
            else if (lk.x < 0) wd[nw++] = 0;
        }
        if (lk.x >= 0 && w > wmax) {
            const int kids[2] = {node + 1, lk.x};
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int k = kids[c];
                if (v.hi[k] - v.lo[k] <= wmax) t[nt++] = k;
                else if (v.link[k].x < 0) wd[nw++] = k;
            }
        }
    }
    block_append<kPlanThreads>(nt, t[0], t[1], list, count, s_part);
    block_append<kPlanThreads>(nw, wd[0], wd[1], wide_list, count + 2, s_part);
}

// Two kernels per evaluation, one wavefront per target (grid-stride over
// the targets, the wide leaves first).
//
// local_walk_kernel walks the tree for the target breadth-first: each
// level's candidate nodes are spread over the lanes, and ballots compact the
// outcomes (accepted panels, reached leaves, the children of rejected nodes:
// the next level) into lists, so a target costs one round of loads per tree
// level instead of one per visit.  The accepted panels and reached leaves go
// to the target's slots in global memory (kLocalPanels and kLocalNear of
// them); a target whose lists overflow goes to the fallback list instead.
// The walk is a chain of round trips with little arithmetic, so it is kept
// to few registers and little LDS, and many wavefronts share a SIMD.
//
// local_f64_kernel then evaluates each target from its lists.  Lane l
// evaluates the M2L terms u = l + 64 s (s < kS) of every accepted panel:
// target point k = u / NT, panel point j = u % NT; the per-term sums are
// added per k in j order.  A leaf wider than the target width instead sums
// the accepted panels at each of its queries (M2P; it has at most leaf_cap
// queries).
constexpr int kLocalPanels = 64;
constexpr int kLocalStage = 32;   // panels staged in LDS at a time for M2L

// Two targets per wavefront, one per 32-lane half, so that twice as many
// walks are in flight per SIMD: each half spreads its own front over its
// lanes, and its ballots are masked to its half.  Both halves step the
// levels together; a half whose walk has ended idles.
constexpr int kWalkFront = 256;  // front capacity per target

__global__ __launch_bounds__(64) void local_walk_kernel(EvalView v, const int *__restrict__ list,
                                                        const int *__restrict__ wide_list,
                                                        const int *__restrict__ count, double eta,
                                                        int *__restrict__ acc_n,
                                                        int *__restrict__ acc_buf,
                                                        int2 *__restrict__ near_buf) {
    __shared__ int s_front[2][2][kWalkFront];  // [half][buffer][entry]
    const int lane = threadIdx.x, h = lane >> 5, hl = lane & 31;
    const uint64_t hmask = h ? 0xffffffff00000000ull : 0x00000000ffffffffull;
    const uint64_t below = ((lane == 0) ? 0ull : (~0ull >> (64 - lane))) & hmask;
    const int n_wide = count[2];
    const int total = count[0] + n_wide;
    const double eta2 = eta * eta;
    for (int e0 = 2 * blockIdx.x; e0 < total; e0 += 2 * gridDim.x) {
        const int e = e0 + h;
        const bool has = e < total;
        const int T = has ? ((e < n_wide) ? wide_list[e] : list[e - n_wide]) : 0;
        const double tlo = v.lo[T], thi = v.hi[T];
        int *acc_out = acc_buf + (size_t)e * kLocalPanels;
        int2 *near_out = near_buf + (size_t)e * kLocalNear;
        bool ok = has;
        int n_near = 0, n_acc = 0, fs = has ? 1 : 0, cur = 0;
        if (hl == 0) s_front[h][0][0] = 0;
        __syncthreads();
        while (__any(fs > 0 && ok)) {
            const bool act = fs > 0 && ok;  // uniform within the half
            const int chunks = act ? (fs + 31) / 32 : 0;
            const int steps = max(chunks, __shfl_xor(chunks, 32, 64));
            int nfs = 0;
            for (int c = 0; c < steps; ++c) {
                const int idx = c * 32 + hl;
                const bool valid = c < chunks && idx < fs;
                const int node = valid ? s_front[h][cur][idx] : 0;
                const int4 lk = v.link[node];
                const double plo = v.lo[node], phi = v.hi[node], pacc = v.acc[node];
                const bool leaf = valid && lk.x < 0;
                // the reference's test at the target's point nearest the panel centre
                const bool acc =
                    valid && !leaf &&
                    accept_f64_of(plo, phi, pacc, fmin(fmax(0.5 * (plo + phi), tlo), thi), eta2);
                const bool desc = valid && !leaf && !acc;
                const uint64_t ml = __ballot(leaf) & hmask, ma = __ballot(acc) & hmask,
                               md = __ballot(desc) & hmask;
                if (leaf) {
                    const int at = n_near + __popcll(ml & below);
                    if (at < kLocalNear) near_out[at] = make_int2(lk.z, lk.w);
                }
                if (acc) {
                    const int at = n_acc + __popcll(ma & below);
                    if (at < kLocalPanels) acc_out[at] = node;
                }
                if (desc) {
                    const int at = nfs + 2 * __popcll(md & below);
                    if (at + 1 < kWalkFront) {
                        s_front[h][cur ^ 1][at] = node + 1;
                        s_front[h][cur ^ 1][at + 1] = lk.x;
                    }
                }
                n_near += __popcll(ml);
                n_acc += __popcll(ma);
                nfs += 2 * __popcll(md);
            }
            if (act && (n_near > kLocalNear || n_acc > kLocalPanels || nfs > kWalkFront)) ok = false;
            __syncthreads();
            if (act) {
                cur ^= 1;
                fs = nfs;
            }
        }
        // panels, then leaves (the near count in the upper half); -1: overflow
        if (has && hl == 0) acc_n[e] = ok ? (n_acc | (n_near << 16)) : -1;
    }
}

template <int NT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void local_f64_kernel(
    EvalView v, const int *__restrict__ list,
                                                       const int *__restrict__ wide_list,
                                                       const int *__restrict__ count, double eta,
                                                       double scale,
                                                       const int *__restrict__ acc_n,
                                                       const int *__restrict__ acc_buf,
                                                       const int2 *__restrict__ near_buf,
                                                       double *__restrict__ out_re,
                                                       double *__restrict__ out_im) {
    constexpr int kTerms = NT * NT;
    constexpr int kS = (kTerms + 63) / 64;
    // One LDS region, reused in turn: the staged panels (positions, then
    // weights), the M2L term sums, the reached leaves' sources.
    constexpr int kRegion = 2 * kLocalStage * NT;
    static_assert(2 * 64 * kS <= kRegion, "the term sums fit the staging region");
    constexpr int kNearStage = kRegion;
    __shared__ double s_region[kRegion];
    __shared__ double s_cr[NT], s_ci[NT], s_y[NT];
    __shared__ int2 s_near[kLocalNear];
    __shared__ int s_acc[kLocalPanels];
    double *s_pt = s_region, *s_pw = s_region + kLocalStage * NT;
    double *s_fr = s_region, *s_fi = s_region + 64 * kS, *s_ns = s_region;
    const int lane = threadIdx.x;
    const int n_wide = count[2];
    const int total = count[0] + n_wide;
    const double eta2 = eta * eta;
    for (int e = blockIdx.x; e < total; e += gridDim.x) {
        const int code = acc_n[e];
        const bool wide = e < n_wide;
        const int T = wide ? wide_list[e] : list[e - n_wide];
        if (code < 0) {
            // the walk's lists overflowed: the reference's per-query walk
            const int4 lkT = v.link[T];
            for (int i = lkT.z + lane; i < lkT.w; i += 64) {
                double re, im;
                walk_one_f64(v, v.sorted[i], eta, eta2, re, im);
                const int o = v.perm ? v.perm[i] : i;
                out_re[o] = re * scale;
                out_im[o] = im * scale;
            }
            continue;
        }
        const int n_acc = code & 0xffff, n_near = code >> 16;
        if (lane < n_acc) s_acc[lane] = acc_buf[(size_t)e * kLocalPanels + lane];
        if (lane < n_near) s_near[lane] = near_buf[(size_t)e * kLocalNear + lane];
        const int4 lkT = v.link[T];
        const int q0 = lkT.z, q1 = lkT.w;
        __syncthreads();
        if (wide && q1 - q0 <= 64) {
            // a leaf wider than the target width: the accepted panels at each
            // query (M2P), the walk shared; one query per lane, the panels
            // and then the reached leaves' sources staged in LDS as below
            const int i = q0 + lane;
            const bool q_ok = i < q1;
            const double x = q_ok ? v.sorted[i] : 0.0;
            double fr = 0.0, fi = 0.0;
            for (int a0 = 0; a0 < n_acc; a0 += kLocalStage) {
                const int cnt = min(kLocalStage, n_acc - a0);
                for (int e2 = lane; e2 < cnt * NT; e2 += 64) {
                    const int b = e2 / NT, j = e2 - b * NT;
                    const size_t o = (size_t)s_acc[a0 + b] * NT + j;
                    s_pt[e2] = v.nodes[o];
                    s_pw[e2] = v.w[o];
                }
                __syncthreads();
                if (q_ok) {
                    for (int b = 0; b < cnt; ++b) {
#pragma unroll
                        for (int j = 0; j < NT; ++j) {
                            const double d = x - s_pt[b * NT + j];
                            const double q = s_pw[b * NT + j] * rcp_f64(__builtin_fma(d, d, eta2));
                            fr = __builtin_fma(q, d, fr);
                            fi += q;
                        }
                    }
                }
                __syncthreads();  // the next chunk overwrites the stage
            }
            int n_ns = 0;
            for (int m = 0; m < n_near; ++m) n_ns += s_near[m].y - s_near[m].x;
            if (n_ns <= kNearStage) {
                int off = 0;
                for (int m = 0; m < n_near; ++m) {
                    const int2 lf = s_near[m];
                    for (int jj = lane; jj < lf.y - lf.x; jj += 64) s_ns[off + jj] = v.sorted[lf.x + jj];
                    off += lf.y - lf.x;
                }
                __syncthreads();
                if (q_ok)
                    for (int j = 0; j < n_ns; ++j) {
                        const double d = x - s_ns[j];
                        const double inv = rcp_f64(__builtin_fma(d, d, eta2));
                        fr = __builtin_fma(d, inv, fr);
                        fi += inv;
                    }
            } else if (q_ok) {
                for (int m = 0; m < n_near; ++m) {
                    const int2 lf = s_near[m];
                    for (int j = lf.x; j < lf.y; ++j) {
                        const double d = x - v.sorted[j];
                        const double inv = rcp_f64(__builtin_fma(d, d, eta2));
                        fr = __builtin_fma(d, inv, fr);
                        fi += inv;
                    }
                }
            }
            if (q_ok) {
                const int o = v.perm ? v.perm[i] : i;
                out_re[o] = fr * scale;
                out_im[o] = (eta * fi) * scale;
            }
        } else if (wide) {
            // the same with more than 64 queries (a leaf at the depth cap):
            // the panels read from global memory per query
            for (int c0 = q0; c0 < q1; c0 += 64) {
                const int i = c0 + lane;
                if (i >= q1) continue;
                const double x = v.sorted[i];
                double fr = 0.0, fi = 0.0;
                for (int a = 0; a < n_acc; ++a) {
                    const double *ts = v.nodes + (size_t)s_acc[a] * NT;
                    const double *ws = v.w + (size_t)s_acc[a] * NT;
#pragma unroll
                    for (int j = 0; j < NT; ++j) {
                        const double d = x - ts[j];
                        const double q = ws[j] * rcp_f64(__builtin_fma(d, d, eta2));
                        fr = __builtin_fma(q, d, fr);
                        fi += q;
                    }
                }
                for (int m = 0; m < n_near; ++m) {
                    const int2 lf = s_near[m];
                    for (int j = lf.x; j < lf.y; ++j) {
                        const double d = x - v.sorted[j];
                        const double inv = rcp_f64(__builtin_fma(d, d, eta2));
                        fr = __builtin_fma(d, inv, fr);
                        fi += inv;
                    }
                }
                const int o = v.perm ? v.perm[i] : i;
                out_re[o] = fr * scale;
                out_im[o] = (eta * fi) * scale;
            }
        } else {
            // M2L over the accepted panels
            double yk[kS], ar[kS], ai[kS];
            int jk[kS];
#pragma unroll
            for (int s = 0; s < kS; ++s) {
                const int u = lane + 64 * s;
                const int k = (u < kTerms) ? u / NT : 0;
                jk[s] = (u < kTerms) ? u % NT : 0;
                yk[s] = v.nodes[(size_t)T * NT + k];
                ar[s] = ai[s] = 0.0;
            }
            // kLocalStage panels at a time: their positions and weights are
            // staged in LDS with one round of loads, then each lane runs its
            // terms over them in list order
            for (int a0 = 0; a0 < n_acc; a0 += kLocalStage) {
                const int cnt = min(kLocalStage, n_acc - a0);
                for (int e = lane; e < cnt * NT; e += 64) {
                    const int b = e / NT, j = e - b * NT;
                    const size_t o = (size_t)s_acc[a0 + b] * NT + j;
                    s_pt[e] = v.nodes[o];
                    s_pw[e] = v.w[o];
                }
                __syncthreads();
#pragma unroll 4
                for (int b = 0; b < cnt; ++b) {
#pragma unroll
                    for (int s = 0; s < kS; ++s) {
                        if (lane + 64 * s < kTerms) {
                            const double d = yk[s] - s_pt[b * NT + jk[s]];
                            const double q = s_pw[b * NT + jk[s]] * rcp_f64(__builtin_fma(d, d, eta2));
                            ar[s] = __builtin_fma(q, d, ar[s]);
                            ai[s] += q;
                        }
                    }
                }
                __syncthreads();  // the next chunk overwrites the stage
            }
#pragma unroll
            for (int s = 0; s < kS; ++s) {
                s_fr[lane + 64 * s] = ar[s];
                s_fi[lane + 64 * s] = ai[s];
            }
            if (lane < NT) s_y[lane] = v.nodes[(size_t)T * NT + lane];
            __syncthreads();
            if (lane < NT) {
                double cr = 0.0, ci = 0.0;
#pragma unroll 1
                for (int j = 0; j < NT; ++j) {
                    cr += s_fr[lane * NT + j];
                    ci += s_fi[lane * NT + j];
                }
                s_cr[lane] = cr;
                s_ci[lane] = ci;
            }
            __syncthreads();  // the term sums are consumed
            // the reached leaves' sources, in leaf order, staged when they fit
            int n_ns = 0;
            for (int m = 0; m < n_near; ++m) n_ns += s_near[m].y - s_near[m].x;
            const bool near_staged = n_ns <= kNearStage;
            if (near_staged) {
                int off = 0;
                for (int m = 0; m < n_near; ++m) {
                    const int2 lf = s_near[m];
                    for (int jj = lane; jj < lf.y - lf.x; jj += 64) s_ns[off + jj] = v.sorted[lf.x + jj];
                    off += lf.y - lf.x;
                }
            }
            __syncthreads();
            for (int c0 = q0; c0 < q1; c0 += 64) {
                const int i = c0 + lane;
                if (i >= q1) continue;
                const double x = v.sorted[i];
                // L2P: barycentric weights (-1)^k, halved at both ends
                double fr = 0.0, fi = 0.0;
                int hit = -1;
                double num_r = 0.0, num_i = 0.0, den = 0.0;
#pragma unroll 2
                for (int k = 0; k < NT; ++k) {
                    const double dx = x - s_y[k];
                    if (dx == 0.0 && hit < 0) hit = k;
                    const double lam =
                        ((k & 1) ? -1.0 : 1.0) * ((k == 0 || k == NT - 1) ? 0.5 : 1.0);
                    const double r = lam * rcp_f64(dx);
                    num_r = __builtin_fma(r, s_cr[k], num_r);
                    num_i = __builtin_fma(r, s_ci[k], num_i);
                    den += r;
                }
                if (hit >= 0) {
                    fr = s_cr[hit];
                    fi = s_ci[hit];
                } else {
                    const double id = rcp_f64(den);
                    fr = num_r * id;
                    fi = num_i * id;
                }
                // the reached leaves, exactly (in leaf and source order)
                double nr = 0.0, ni = 0.0;
                if (near_staged) {
                    for (int j = 0; j < n_ns; ++j) {
                        const double d = x - s_ns[j];
                        const double inv = rcp_f64(__builtin_fma(d, d, eta2));
                        nr = __builtin_fma(d, inv, nr);
                        ni += inv;
                    }
                } else {
                    for (int m = 0; m < n_near; ++m) {
                        const int2 lf = s_near[m];
                        for (int j = lf.x; j < lf.y; ++j) {
                            const double d = x - v.sorted[j];
                            const double inv = rcp_f64(__builtin_fma(d, d, eta2));
                            nr = __builtin_fma(d, inv, nr);
                            ni += inv;
                        }
                    }
                }
                const int o = v.perm ? v.perm[i] : i;
                out_re[o] = (fr + nr) * scale;
                out_im[o] = (eta * (fi + ni)) * scale;
            }
        }
        __syncthreads();  // LDS is reused by the next target
    }
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
    v.info = t->info_dev;
    return v;
}

inline unsigned blocks_for(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void cheb_destroy(ChebTree *t) {
    if (!t) return;
    // Cached arrays of a transient tree: mark their last use (the work its
    // thread enqueued on the tree's stream); the next use waits for it.
    if (t->cached_src) (void)cache_release(CACHE_CHEB_SRC, t->stream);
    if (t->cached_nodes) (void)cache_release(CACHE_CHEB_NODES, t->stream);
    const bool own_src = t->arena && !t->cached_src, own_nodes = t->arena2 && !t->cached_nodes;
    if (own_src || own_nodes) {
        // Stream-ordered pool memory: wait until no stream can still read the
        // tree (the synchronisation hipFree would do), then return it.
        (void)hipDeviceSynchronize();
        if (own_src) (void)hipFreeAsync(t->arena, nullptr);
        if (own_nodes) (void)hipFreeAsync(t->arena2, nullptr);
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

// Marks the end of a cache slot's use when it goes out of scope.
struct CacheUse {
    int slot;
    hipStream_t s;
    bool on = false;
    ~CacheUse() {
        if (on) (void)cache_release(slot, s);
    }
};

}  // namespace

int cheb_build(const double *d_ev, size_t p, const ChebParams &prm, ChebTree **out,
               hipStream_t stream, int transient) {
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
    // The build waits once on the host for its topology (below), which a
    // stream under graph capture cannot do: refuse before enqueuing anything,
    // so the capture stays valid.
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        SHK_HIP(hipStreamIsCapturing(stream, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            set_error("chebcode build: a ChebCode build waits on the host for its topology and "
                      "cannot be captured in a graph; build the tree outside the capture "
                      "(shk_chebcode_build_dev) and capture its evaluations");
            return 6;
        }
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

    // Node slots: the top walk's BFS ids (at most kFrontier per level before
    // the frontier, which holds fewer than kMaxFrontier) and 2p for the
    // subtrees.
    const int cap_top = (int)std::min<size_t>(2 * p - 1, (size_t)(kMaxDepth + 3) * kFrontier);
    const size_t slots = (size_t)cap_top + 2 * p;

    // Scratch: info words, slot arrays, node list, subtree levels, node/
    // weight tables, sort temporaries.
    const int n_info = kInfoInts;
    Carver sc;
    (void)sc.take<int>(n_info);
    (void)sc.take<int>(6 * slots);
    (void)sc.take<int2>(slots);
    (void)sc.take<int>(3 * (size_t)kMaxFrontier * kLv + kMaxFrontier);
    (void)sc.take<double>(2 * (size_t)n);
    (void)sc.take<int32_t>(p);
    // hint buckets: kHintPerSource per source, so that a split search in the
    // bulk of a spectrum with outliers (which stretch the value range) meets
    // few sources per bucket
    const int R = (int)std::max<int64_t>(ip, std::min<int64_t>((int64_t)kHintPerSource * ip,
                                                                kHintMaxBuckets));
    (void)sc.take<int>((size_t)R + 1);
    (void)sc.take<int>(kGapInts);
    size_t sort_bytes = 0;
    SHK_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, d_ev, (double *)nullptr,
                                      (const int32_t *)nullptr, (int32_t *)nullptr, p, 0, 64,
                                      stream));
    (void)sc.take<char>(sort_bytes);
    void *scratch = nullptr;
    if (int st = cache_acquire(CACHE_CHEB_SCRATCH, sc.off, stream, &scratch)) return st;
    CacheUse scratch_use{CACHE_CHEB_SCRATCH, stream, true};
    sc.base = static_cast<char *>(scratch);
    sc.off = 0;
    int *d_info = sc.take<int>(n_info);
    int *b_buf = sc.take<int>(6 * slots);
    int2 *d_mrec = sc.take<int2>(slots);
    int *lv_buf = sc.take<int>(3 * (size_t)kMaxFrontier * kLv + kMaxFrontier);
    double *d_tab = sc.take<double>(2 * (size_t)n);
    int32_t *d_iota = sc.take<int32_t>(p);
    int *d_hint = sc.take<int>((size_t)R + 1);
    int *d_gap = sc.take<int>(kGapInts);
    void *d_sort_tmp = sc.take<char>(sort_bytes);
    int *b_lo = b_buf, *b_hi = b_buf + slots, *b_left = b_buf + 2 * slots,
        *b_size = b_buf + 3 * slots, *b_pre = b_buf + 4 * slots, *d_list = b_buf + 5 * slots;
    int *lv_start = lv_buf, *lv_cnt = lv_buf + (size_t)kMaxFrontier * kLv,
        *lv_off = lv_buf + 2 * (size_t)kMaxFrontier * kLv,
        *sub_nodes = lv_buf + 3 * (size_t)kMaxFrontier * kLv;

    // Arena part 1 (sizes known before the topology): sources.
    Carver ar;
    (void)ar.take<double>(p);
    (void)ar.take<int32_t>(p);
    (void)ar.take<float>(p);
    (void)ar.take<float>(p);
    const size_t src_bytes = ar.off;

    // 1. sortedness, then the sorted copy (+ permutation)
    std::vector<int> info(n_info, 0);
    int *mail = nullptr;
    if (int st = host_mailbox(&mail)) return st;
    thread_local int mail_seq = 0;
    // The topology needs the sorted values; its node count sizes the rest of
    // the arena, so the arena is allocated after a first pass with the
    // sources in scratch-independent storage: allocate the source part now.
    void *src_arena = nullptr;
    t->stream = stream;
    if (transient) {
        if (int st = cache_acquire(CACHE_CHEB_SRC, src_bytes, stream, &src_arena)) return st;
        t->cached_src = true;
    } else {
        SHK_HIP(hipMallocAsync(&src_arena, src_bytes, stream));
    }
    ar.base = static_cast<char *>(src_arena);
    ar.off = 0;
    t->sorted = ar.take<double>(p);
    int32_t *perm_buf = ar.take<int32_t>(p);
    t->s_hi = ar.take<float>(p);
    t->s_lo = ar.take<float>(p);
    t->arena = src_arena;  // owned from here on (freed by cheb_destroy)

    // Chebyshev nodes of the second kind s_j = cos(j*pi/(n-1)) and
    // barycentric weights (-1)^j halved at both ends (chebcode.rs:242-261).
    ChebTab tab{};
    for (int j = 0; j < n; ++j) {
        tab.v[j] = std::cos((double)j * kPi / (double)(n - 1));
        const double sign = (j % 2 == 0) ? 1.0 : -1.0;
        tab.v[n + j] = (j == 0 || j == n - 1) ? 0.5 * sign : sign;
    }
    // 1-2. Speculate that the input is ascending (eigenvalues usually are):
    // copy it, check it and build the topology on the copy, with one host
    // synchronisation for the check and the node count together.  An input
    // that is not ascending is then sorted (+ permutation) and the topology
    // rebuilt; on unsorted values the first topology still stays in bounds
    // (every split lies strictly inside its node), it is only wasted work.
    auto topology = [&](bool first) -> int {
        hipLaunchKernelGGL(hint_clear_kernel, dim3(blocks_for(hint_clear_threads(R + 1), 256)),
                           dim3(256), 0, stream, d_hint, R + 1, d_gap, d_info);
        SHK_HIP(hipGetLastError());
        if (first)  // copy, check, hint and split of the input, and the table
            hipLaunchKernelGGL(ingest_kernel, dim3(blocks_for(std::max<size_t>(p, 2 * n), 256)),
                               dim3(256), 0, stream, d_ev, ip, R, t->sorted, d_hint, d_gap,
                               t->s_hi, t->s_lo, d_info, tab, 2 * n, d_tab);
        else  // the sorted copy
            hipLaunchKernelGGL(hint_scatter_kernel, dim3(blocks_for(p, 256)), dim3(256), 0,
                               stream, t->sorted, ip, R, d_hint, d_gap, t->s_hi, t->s_lo);
        SHK_HIP(hipGetLastError());
        const TopoCtx c{t->sorted, ip,  leaf_cap, d_hint, R,      Hint{}, d_gap,
                        0,         b_lo, b_hi,     b_left, b_size, b_pre};
        hipLaunchKernelGGL(topo_top_kernel, dim3(1), dim3(kTopoThreads), 0, stream, c, cap_top,
                           d_info);
        SHK_HIP(hipGetLastError());
        hipLaunchKernelGGL(topo_sub_kernel, dim3(kMaxFrontier), dim3(kSubThreads), 0, stream, c,
                           d_info, lv_start, lv_cnt, sub_nodes);
        SHK_HIP(hipGetLastError());
        const int seq = ++mail_seq;
        hipLaunchKernelGGL(topo_mid_kernel, dim3(1), dim3(kTopoThreads), 0, stream, c, d_info,
                           sub_nodes, lv_off, d_list, mail, seq);
        SHK_HIP(hipGetLastError());
        hipLaunchKernelGGL(topo_fin_kernel, dim3(kMaxFrontier), dim3(kSubThreads), 0, stream, c,
                           d_info, lv_start, lv_cnt, lv_off, d_list);
        SHK_HIP(hipGetLastError());
        // the host continues as soon as C has published the info words; D
        // and everything after it follow in stream order
        if (int st = mailbox_wait(mail, seq, stream)) return st;
        std::copy(mail + kMailData, mail + kMailData + n_info, info.begin());
        return 0;
    };
    if (int st = topology(true)) return st;
    if (info[kInfoUnsorted] != 0) {
        t->perm = perm_buf;
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(p, 256)), dim3(256), 0, stream, d_iota, ip);
        SHK_HIP(hipGetLastError());
        SHK_HIP(rocprim::radix_sort_pairs(d_sort_tmp, sort_bytes, d_ev, t->sorted, d_iota, t->perm,
                                          p, 0, 64, stream));
        if (int st = topology(false)) return st;
    }
    if (info[kInfoOverflow]) {
        set_error("chebcode build: node capacity exceeded");
        return 6;
    }
    const int N = info[kInfoN];
    t->N = (size_t)N;

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
    if (transient) {
        if (int st = cache_acquire(CACHE_CHEB_NODES, nr.off, stream, &node_arena)) return st;
        t->cached_nodes = true;
    } else {
        SHK_HIP(hipMallocAsync(&node_arena, nr.off, stream));
    }
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

    // 3. node records and mirrors, leaf weights: one wavefront per node
    TreeArrays ta{t->lo, t->hi, t->acc, t->nodes, t->w, t->lo_idx, t->hi_idx, t->right, t->esc};
    const size_t lds = bary_lds_bytes(n);
    // Grouped at the presets' orders when leaves fit a lane group's rows;
    // wide leaves (the grid path's leaf_cap) keep one wavefront per node.
    if ((n == 8 || n == 11 || n == 12) && prm.leaf_cap <= kGroupRowsMax) {
        const dim3 grid((unsigned)((N + kNodeGroup - 1) / kNodeGroup));
        const size_t bytes = sizeof(double) * (size_t)node_group_lds_doubles(n);
        auto kern = (n == 8) ? node_group_kernel<8>
                             : (n == 11) ? node_group_kernel<11> : node_group_kernel<12>;
        hipLaunchKernelGGL(kern, grid, dim3(kBaryLanes), bytes, stream, t->sorted, N,
                           t->theta_sq, d_list, b_lo, b_hi, b_left, b_size, b_pre, d_tab, ta,
                           d_mrec, t->rec32, t->link, t->panel32, nullptr);
    } else {
        hipLaunchKernelGGL(node_kernel, dim3(N), dim3(kBaryLanes), lds + sizeof(double) * n,
                           stream, t->sorted, n, t->theta_sq, d_list, b_lo, b_hi, b_left, b_size,
                           b_pre, d_tab, ta, d_mrec, t->rec32, t->link, t->panel32, nullptr);
    }
    SHK_HIP(hipGetLastError());
    // 4. parents, deepest first.  The subtree workgroups keep their LDS
    // within 80 KiB (two per CU); the one top-level workgroup may use the
    // whole 160 KiB of a CU.
    const int lds_waves = (int)std::max<size_t>(1, 65536 / lds);
    const int F = info[kInfoF], D = info[kInfoD];
    if (F > 0 && (n == 8 || n == 11 || n == 12)) {
        // LDS form: subtrees of up to `cap` nodes in an 80 KiB workgroup
        const size_t fixed = (size_t)kMergeSubWaves * lds;
        const int cap = (int)((kSubLdsMax - fixed) / (2 * sizeof(int) + 2 * sizeof(double) * n));
        const size_t bytes = fixed + (size_t)cap * (2 * sizeof(int) + 2 * sizeof(double) * n);
        const dim3 blk(kMergeSubWaves * kBaryLanes);
        if (n == 8)
            hipLaunchKernelGGL(merge_sub_lds_kernel<8>, dim3(F), blk, bytes, stream, lv_start,
                               lv_cnt, sub_nodes, b_left, d_mrec, d_tab + n, ta, t->panel32, cap,
                               nullptr);
        else if (n == 11)
            hipLaunchKernelGGL(merge_sub_lds_kernel<11>, dim3(F), blk, bytes, stream, lv_start,
                               lv_cnt, sub_nodes, b_left, d_mrec, d_tab + n, ta, t->panel32, cap,
                               nullptr);
        else
            hipLaunchKernelGGL(merge_sub_lds_kernel<12>, dim3(F), blk, bytes, stream, lv_start,
                               lv_cnt, sub_nodes, b_left, d_mrec, d_tab + n, ta, t->panel32, cap,
                               nullptr);
        SHK_HIP(hipGetLastError());
    } else if (F > 0) {
        const int waves = std::min(lds_waves, kMergeSubWaves);
        hipLaunchKernelGGL(merge_sub_kernel, dim3(F), dim3(waves * kBaryLanes), waves * lds, stream,
                           n, lv_start, lv_cnt, d_mrec, d_tab + n, ta, t->panel32);
        SHK_HIP(hipGetLastError());
    }
    // Top levels with internal nodes: [0, D) below a frontier, else [0, D-1)
    // (the last level of a tree that ended in the top walk holds leaves).
    const int top = (F > 0) ? D : std::max(D - 1, 0);
    const int *lvl = info.data() + kInfoTopLevel;
    const int top_waves = std::min(lds_waves, kMergeMaxThreads / kBaryLanes);
    const int group = merge_group_size(n);
    auto groups = [&](int d) { return (lvl[d + 1] - lvl[d] + group - 1) / group; };
    int small = 0;  // leading levels merged together in one workgroup
    // LDS form (the presets' orders): as many top levels as fit, with their
    // finished children
    const bool top_lds = (n == 8 || n == 11 || n == 12);
    if (top_lds)
        while (small < top && top_lds_bytes(lvl[small + 2], n, kTopLdsWaves) <= kTopLdsWgMax &&
               groups(small) <= kMergeTopRounds * kTopLdsWaves)
            ++small;
    else
        while (small < top && groups(small) <= kMergeTopRounds * top_waves) ++small;
    for (int d = top - 1; d >= small; --d) {
        hipLaunchKernelGGL(merge_range_kernel, dim3(groups(d)), dim3(kBaryLanes), lds, stream, n,
                           lvl[d], lvl[d + 1] - lvl[d], d_mrec, d_tab + n, ta, t->panel32);
        SHK_HIP(hipGetLastError());
    }
    if (small > 0 && top_lds) {
        const size_t bytes = top_lds_bytes(lvl[small + 1], n, kTopLdsWaves);
        const dim3 blk(kTopLdsWaves * kBaryLanes);
        if (n == 8)
            hipLaunchKernelGGL(merge_top_lds_kernel<8>, dim3(1), blk, bytes, stream, small, d_info,
                               b_left, d_mrec, d_tab + n, ta, t->panel32);
        else if (n == 11)
            hipLaunchKernelGGL(merge_top_lds_kernel<11>, dim3(1), blk, bytes, stream, small,
                               d_info, b_left, d_mrec, d_tab + n, ta, t->panel32);
        else
            hipLaunchKernelGGL(merge_top_lds_kernel<12>, dim3(1), blk, bytes, stream, small,
                               d_info, b_left, d_mrec, d_tab + n, ta, t->panel32);
        SHK_HIP(hipGetLastError());
    } else if (small > 0) {
        hipLaunchKernelGGL(merge_top_kernel, dim3(1), dim3(top_waves * kBaryLanes),
                           top_waves * lds, stream, n, small, d_info, d_mrec, d_tab + n, ta,
                           t->panel32);
        SHK_HIP(hipGetLastError());
    }
    // Evaluations on any stream wait for this event (cheb_eval*).  A
    // transient tree is evaluated on its build stream only and needs none.
    if (!transient) {
        SHK_HIP(hipEventCreateWithFlags(&t->ready, hipEventDisableTiming));
        SHK_HIP(hipEventRecord(t->ready, stream));
    }
    *out = t.release();
    return 0;
}

// ---- one-shot transform in a caller's workspace -----------------------------
//
// cheb_build waits once on the host: the node count sizes the node arena and
// the top merges' launches, and the sortedness check decides the sort.  The
// workspace form takes ascending input and keeps every size on the device:
//   - the node arena is carved at its bound (every node slot, cap_top + 2p),
//     so no count is needed to place it;
//   - the node kernels run a fixed grid that strides over the node count
//     read from the info words, the subtree merges a grid of kMaxFrontier
//     workgroups past the frontier size of which exit, and the top merges
//     one workgroup that reads its levels there;
//   - the walks read the node count there too.
// Nothing waits on the host and nothing is allocated, so a captured graph can
// hold the whole transform.  The kernels and their arithmetic are those of
// cheb_build and cheb_eval (the top merges take merge_top_kernel's form for
// every top level), so on ascending input the results equal the transient
// build's bit for bit.  Input that is not ascending (ingest_kernel's check)
// gets an empty topology and NaN outputs; info word kInfoUnsorted reports it.

namespace {

struct WsParts {
    int *info = nullptr, *b_buf = nullptr, *lv_buf = nullptr, *hint = nullptr, *gap = nullptr;
    int2 *mrec = nullptr;
    double *tab = nullptr, *sorted = nullptr;
    float *s_hi = nullptr, *s_lo = nullptr;
    double *lo = nullptr, *hi = nullptr, *acc = nullptr, *nodes = nullptr, *w = nullptr;
    int32_t *lo_idx = nullptr, *hi_idx = nullptr, *right = nullptr, *esc = nullptr;
    float4 *rec32 = nullptr;
    int4 *link = nullptr;
    float *panel32 = nullptr;
    size_t slots = 0;
    int cap_top = 0, R = 0;
};

// Carves the parts from base (sizes only when base is null); returns the
// bytes used.  The info words come first (offset 0).
size_t carve_ws(char *base, size_t p, int n, WsParts &w) {
    Carver c;
    c.base = base;
    w.cap_top = (int)std::min<size_t>(2 * p - 1, (size_t)(kMaxDepth + 3) * kFrontier);
    w.slots = (size_t)w.cap_top + 2 * p;
    const size_t S = w.slots;
    const int ip = (int)p;
    w.R = (int)std::max<int64_t>(ip, std::min<int64_t>((int64_t)kHintPerSource * ip,
                                                       kHintMaxBuckets));
    w.info = c.take<int>(kInfoInts);
    w.b_buf = c.take<int>(6 * S);
    w.mrec = c.take<int2>(S);
    w.lv_buf = c.take<int>(3 * (size_t)kMaxFrontier * kLv + kMaxFrontier);
    w.tab = c.take<double>(2 * (size_t)n);
    w.hint = c.take<int>((size_t)w.R + 1);
    w.gap = c.take<int>(kGapInts);
    w.sorted = c.take<double>(p);
    w.s_hi = c.take<float>(p);
    w.s_lo = c.take<float>(p);
    w.lo = c.take<double>(S);
    w.hi = c.take<double>(S);
    w.acc = c.take<double>(S);
    w.lo_idx = c.take<int32_t>(S);
    w.hi_idx = c.take<int32_t>(S);
    w.right = c.take<int32_t>(S);
    w.esc = c.take<int32_t>(S);
    w.nodes = c.take<double>(S * n);
    w.w = c.take<double>(S * n);
    w.rec32 = c.take<float4>(S);
    w.link = c.take<int4>(S);
    w.panel32 = c.take<float>(S * 2 * n);
    return (c.off + 255) / 256 * 256;
}

constexpr int kWsNodeGrid = 16384;  // workgroups of the node kernels (stride over the nodes)

}  // namespace

size_t cheb_ws_bytes(size_t p, int n) {
    if (p == 0 || n < 2 || n > kMaxN) return 0;
    WsParts w;
    return carve_ws(nullptr, p, n, w);
}

int cheb_oneshot_ws(const double *d_ev, size_t p, const ChebParams &prm, const double *d_query,
                    size_t nq, double eta, double scale, int queries_are_sources, void *ws,
                    size_t ws_bytes, double *d_out_re, double *d_out_im, hipStream_t stream) {
    if (p == 0 || p > (size_t)INT32_MAX / 2) {
        set_error("chebcode build: p must be in [1, 2^30)");
        return 6;
    }
    const int n = prm.n;
    if (!(n == 8 || n == 11 || n == 12) || !(prm.theta > 0.0) || prm.leaf_cap < 1) {
        set_error("chebcode workspace build: the presets' orders (8, 11, 12), theta > 0 and "
                  "leaf_cap >= 1");
        return 6;
    }
    if (reinterpret_cast<uintptr_t>(ws) % 256 != 0) {
        set_error("chebcode workspace: the workspace must be 256-byte aligned");
        return 6;
    }
    const size_t need = cheb_ws_bytes(p, n);
    if (!ws || ws_bytes < need) {
        set_error("chebcode workspace: " + std::to_string(ws_bytes) + " bytes given, " +
                  std::to_string(need) + " needed (shk_chebcode_workspace_bytes)");
        return 6;
    }
    WsParts w;
    (void)carve_ws(static_cast<char *>(ws), p, n, w);
    const int ip = (int)p, R = w.R;
    const int leaf_cap = (int)std::min<int64_t>(prm.leaf_cap, INT32_MAX);
    const size_t S = w.slots;
    int *b_lo = w.b_buf, *b_hi = w.b_buf + S, *b_left = w.b_buf + 2 * S, *b_size = w.b_buf + 3 * S,
        *b_pre = w.b_buf + 4 * S, *d_list = w.b_buf + 5 * S;
    int *lv_start = w.lv_buf, *lv_cnt = w.lv_buf + (size_t)kMaxFrontier * kLv,
        *lv_off = w.lv_buf + 2 * (size_t)kMaxFrontier * kLv,
        *sub_nodes = w.lv_buf + 3 * (size_t)kMaxFrontier * kLv;
    ChebTab tab{};
    for (int j = 0; j < n; ++j) {
        tab.v[j] = std::cos((double)j * kPi / (double)(n - 1));
        const double sign = (j % 2 == 0) ? 1.0 : -1.0;
        tab.v[n + j] = (j == 0 || j == n - 1) ? 0.5 * sign : sign;
    }
    // 1. copy, ascending check, hint and split of the input, and the table
    hipLaunchKernelGGL(hint_clear_kernel, dim3(blocks_for(hint_clear_threads(R + 1), 256)),
                       dim3(256), 0, stream, w.hint, R + 1, w.gap, w.info);
    hipLaunchKernelGGL(ingest_kernel, dim3(blocks_for(std::max<size_t>(p, 2 * n), 256)), dim3(256),
                       0, stream, d_ev, ip, R, w.sorted, w.hint, w.gap, w.s_hi, w.s_lo, w.info,
                       tab, 2 * n, w.tab);
    // 2. topology (no mailbox)
    const TopoCtx c{w.sorted, ip,  leaf_cap, w.hint, R,      Hint{}, w.gap,
                    0,        b_lo, b_hi,    b_left, b_size, b_pre};
    hipLaunchKernelGGL(topo_top_kernel, dim3(1), dim3(kTopoThreads), 0, stream, c, w.cap_top,
                       w.info);
    hipLaunchKernelGGL(topo_sub_kernel, dim3(kMaxFrontier), dim3(kSubThreads), 0, stream, c,
                       w.info, lv_start, lv_cnt, sub_nodes);
    hipLaunchKernelGGL(topo_mid_kernel, dim3(1), dim3(kTopoThreads), 0, stream, c, w.info,
                       sub_nodes, lv_off, d_list, nullptr, 0);
    hipLaunchKernelGGL(topo_fin_kernel, dim3(kMaxFrontier), dim3(kSubThreads), 0, stream, c,
                       w.info, lv_start, lv_cnt, lv_off, d_list);
    SHK_HIP(hipGetLastError());
    // 3. node records and leaf weights, striding over the node count
    TreeArrays ta{w.lo, w.hi, w.acc, w.nodes, w.w, w.lo_idx, w.hi_idx, w.right, w.esc};
    const double theta_sq = prm.theta * prm.theta;
    if (prm.leaf_cap <= kGroupRowsMax) {
        const dim3 grid((unsigned)std::min<size_t>((S + kNodeGroup - 1) / kNodeGroup, kWsNodeGrid));
        const size_t bytes = sizeof(double) * (size_t)node_group_lds_doubles(n);
        auto kern = (n == 8) ? node_group_kernel<8>
                             : (n == 11) ? node_group_kernel<11> : node_group_kernel<12>;
        hipLaunchKernelGGL(kern, grid, dim3(kBaryLanes), bytes, stream, w.sorted, 0, theta_sq,
                           d_list, b_lo, b_hi, b_left, b_size, b_pre, w.tab, ta, w.mrec, w.rec32,
                           w.link, w.panel32, w.info);
    } else {
        const dim3 grid((unsigned)std::min<size_t>(S, kWsNodeGrid));
        hipLaunchKernelGGL(node_kernel, grid, dim3(kBaryLanes),
                           bary_lds_bytes(n) + sizeof(double) * n, stream, w.sorted, n, theta_sq,
                           d_list, b_lo, b_hi, b_left, b_size, b_pre, w.tab, ta, w.mrec, w.rec32,
                           w.link, w.panel32, w.info);
    }
    SHK_HIP(hipGetLastError());
    // 4. parents: the subtrees (workgroups past the frontier exit), then
    // every top level in one workgroup
    const size_t lds = bary_lds_bytes(n);
    {
        const size_t fixed = (size_t)kMergeSubWaves * lds;
        const int cap = (int)((kSubLdsMax - fixed) / (2 * sizeof(int) + 2 * sizeof(double) * n));
        const size_t bytes = fixed + (size_t)cap * (2 * sizeof(int) + 2 * sizeof(double) * n);
        const dim3 blk(kMergeSubWaves * kBaryLanes);
        auto kern = (n == 8) ? merge_sub_lds_kernel<8>
                             : (n == 11) ? merge_sub_lds_kernel<11> : merge_sub_lds_kernel<12>;
        hipLaunchKernelGGL(kern, dim3(kMaxFrontier), blk, bytes, stream, lv_start, lv_cnt,
                           sub_nodes, b_left, w.mrec, w.tab + n, ta, w.panel32, cap, w.info);
    }
    {
        // a top level holds fewer than 2 kFrontier slots
        const int groups = (2 * kFrontier + merge_group_size(n) - 1) / merge_group_size(n);
        const dim3 grid((unsigned)((groups + kWideWaves - 1) / kWideWaves));
        for (int k = 0; k < kWideLaunches; ++k)
            hipLaunchKernelGGL(merge_wide_kernel, grid, dim3(kWideWaves * kBaryLanes),
                               kWideWaves * lds, stream, n, k, w.info, w.mrec, w.tab + n, ta,
                               w.panel32);
        const dim3 blk(kTopLdsWaves * kBaryLanes);
        auto kern = (n == 8) ? merge_top_lds_kernel<8>
                             : (n == 11) ? merge_top_lds_kernel<11> : merge_top_lds_kernel<12>;
        hipLaunchKernelGGL(kern, dim3(1), blk, kTopLdsWgMax, stream, -1, w.info, b_left, w.mrec,
                           w.tab + n, ta, w.panel32);
    }
    SHK_HIP(hipGetLastError());
    // 5. the walks, with the node count from the info words
    ChebTree t;
    t.p = p;
    t.n = n;
    t.theta = prm.theta;
    t.theta_sq = theta_sq;
    t.leaf_cap = prm.leaf_cap;
    t.f32_far = prm.f32_far;
    t.sorted = w.sorted;
    t.s_hi = w.s_hi;
    t.s_lo = w.s_lo;
    t.lo = w.lo;
    t.hi = w.hi;
    t.acc = w.acc;
    t.lo_idx = w.lo_idx;
    t.hi_idx = w.hi_idx;
    t.right = w.right;
    t.esc = w.esc;
    t.nodes = w.nodes;
    t.w = w.w;
    t.rec32 = w.rec32;
    t.link = w.link;
    t.panel32 = w.panel32;
    t.stream = stream;
    t.info_dev = w.info;
    return cheb_eval(&t, d_query, nq, eta, scale, queries_are_sources, d_out_re, d_out_im,
                     stream);
}

int cheb_ws_status(const void *ws, hipStream_t stream, int *ascending) {
    int words[kInfoUnsorted + 1] = {};
    SHK_HIP(hipMemcpyAsync(words, ws, sizeof(words), hipMemcpyDeviceToHost, stream));
    SHK_HIP(hipStreamSynchronize(stream));
    *ascending = (words[kInfoUnsorted] == 0 && words[kInfoOverflow] == 0) ? 1 : 0;
    return 0;
}

size_t cheb_node_count(const ChebTree *t) { return t ? t->N : 0; }
size_t cheb_p(const ChebTree *t) { return t ? t->p : 0; }

// Orders `stream` after the tree's build.  A graph being captured cannot
// wait on work outside it, so under capture the build must already have
// finished, which the host checks on the build's event.
int wait_tree(const ChebTree *t, hipStream_t stream) {
    if (!t->ready) return 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SHK_HIP(hipStreamIsCapturing(stream, &cs));
    if (cs == hipStreamCaptureStatusNone) {
        SHK_HIP(hipStreamWaitEvent(stream, t->ready, 0));
        return 0;
    }
    const hipError_t q = hipEventQuery(t->ready);
    if (q == hipSuccess) return 0;
    if (q == hipErrorNotReady) {
        set_error("chebcode: an evaluation is being captured before its tree's build has "
                  "finished; synchronise the build's stream first");
        return 6;
    }
    SHK_HIP(q);
    return 0;
}

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
    if (int st = wait_tree(t, stream)) return st;
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

// Wave walks over fewer queries than this hold each visit's loads together.
constexpr int kEvalHoldMax = 1 << 19;

int cheb_eval(const ChebTree *t, const double *d_query, size_t nq, double eta, double scale,
              int queries_are_sources, double *d_out_re, double *d_out_im, hipStream_t stream) {
    if (nq == 0) return 0;
    if (nq > (size_t)INT32_MAX) {
        set_error("chebcode eval: too many query points");
        return 6;
    }
    if (int st = wait_tree(t, stream)) return st;
    EvalView v = view_of(t);
    // Dense queries (the tree's own sources, or at least as many points as
    // sources) share near fields within a wavefront: walk together.
    const bool wave = queries_are_sources || nq >= t->p;
    const dim3 grid(blocks_for(nq, 256)), block(256);
    ProfileMark m = profile_begin(stream);
    const bool hold = nq < (size_t)kEvalHoldMax;
    if (t->f32_far && wave && t->n == 8 && hold)
        hipLaunchKernelGGL((eval_fast_kernel<true, 8, true>), grid, block, 0, stream, v, d_query,
                           (int)nq, queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (t->f32_far && wave && t->n == 8)
        hipLaunchKernelGGL((eval_fast_kernel<true, 8>), grid, block, 0, stream, v, d_query,
                           (int)nq, queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (t->f32_far && wave)
        hipLaunchKernelGGL(eval_fast_kernel<true>, grid, block, 0, stream, v, d_query, (int)nq,
                           queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (t->f32_far && nq <= (size_t)kSparseMax)
        hipLaunchKernelGGL(eval_fast_sparse_kernel,
                           dim3(blocks_for(nq, kSparseWaves)), dim3(kSparseWaves * 64), 0,
                           stream, v, d_query, (int)nq, eta, scale, d_out_re, d_out_im);
    else if (t->f32_far)
        hipLaunchKernelGGL(eval_fast_kernel<false>, grid, block, 0, stream, v, d_query, (int)nq,
                           queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (wave && t->n == 11 && hold)
        hipLaunchKernelGGL((eval_f64_kernel<true, 11, true>), grid, block, 0, stream, v, d_query,
                           (int)nq, queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (wave && t->n == 11)
        hipLaunchKernelGGL((eval_f64_kernel<true, 11>), grid, block, 0, stream, v, d_query,
                           (int)nq, queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (wave && t->n == 12 && hold)
        hipLaunchKernelGGL((eval_f64_kernel<true, 12, true>), grid, block, 0, stream, v, d_query,
                           (int)nq, queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (wave && t->n == 12)
        hipLaunchKernelGGL((eval_f64_kernel<true, 12>), grid, block, 0, stream, v, d_query,
                           (int)nq, queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (wave)
        hipLaunchKernelGGL(eval_f64_kernel<true>, grid, block, 0, stream, v, d_query, (int)nq,
                           queries_are_sources, eta, scale, d_out_re, d_out_im);
    else if (nq <= (size_t)kSparseMax)
        hipLaunchKernelGGL(eval_f64_sparse_kernel, dim3(blocks_for(nq, kSparseWaves)),
                           dim3(kSparseWaves * 64), 0, stream, v, d_query, (int)nq, eta, scale,
                           d_out_re, d_out_im);
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
    if (int st = wait_tree(t, stream)) return st;
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

int cheb_eval_local(const ChebTree *t, double eta, double scale, double *d_out_re,
                    double *d_out_im, hipStream_t stream) {
    if (t->f32_far) {
        set_error("local-expansion evaluation needs a tree with an f64 far field");
        return 6;
    }
    if (t->n != 11 && t->n != 12) {
        set_error("local-expansion evaluation is built for orders 11 and 12");
        return 6;
    }
    if (int st = wait_tree(t, stream)) return st;
    EvalView v = view_of(t);
    const int N = (int)t->N;
    // counts (targets, unused, wide leaves), the two target lists, and per
    // target (at most N) its list counts, panels and reached leaves
    int *buf = nullptr;
    CachedBuf cache(CACHE_CHEB_LOCAL, stream);
    const size_t per_target = 1 + kLocalPanels + 2 * kLocalNear;
    if (int e = cache.get(4 + (2 + per_target) * (size_t)N, &buf)) return e;
    int *count = buf, *list = buf + 4, *wide_list = list + N;
    int *acc_n = wide_list + N, *acc_buf = acc_n + N;
    int2 *near_buf = reinterpret_cast<int2 *>(acc_buf + (size_t)N * kLocalPanels);
    hipLaunchKernelGGL(local_zero_kernel, dim3(1), dim3(64), 0, stream, count);
    hipLaunchKernelGGL(local_plan_kernel, dim3(blocks_for((size_t)N, kPlanThreads)),
                       dim3(kPlanThreads), 0, stream, v, kLocalWidth * eta, list, wide_list, count);
    SHK_HIP(hipGetLastError());
    const unsigned grid = (unsigned)std::min(N, 32768);
    ProfileMark m = profile_begin(stream);
    hipLaunchKernelGGL(local_walk_kernel, dim3((grid + 1) / 2), dim3(64), 0, stream, v, list,
                       wide_list, count, eta, acc_n, acc_buf, near_buf);
    SHK_HIP(hipGetLastError());
    if (t->n == 11)
        hipLaunchKernelGGL(local_f64_kernel<11>, dim3(grid), dim3(64), 0, stream, v, list,
                           wide_list, count, eta, scale, acc_n, acc_buf, near_buf, d_out_re,
                           d_out_im);
    else
        hipLaunchKernelGGL(local_f64_kernel<12>, dim3(grid), dim3(64), 0, stream, v, list,
                           wide_list, count, eta, scale, acc_n, acc_buf, near_buf, d_out_re,
                           d_out_im);
    profile_end(m, stream);
    SHK_HIP(hipGetLastError());
    return 0;
}

int cheb_eval_x2(const ChebTree *t, double eta0, double eta1, double *d_re0, double *d_im0,
                 double *d_re1, double *d_im1, hipStream_t stream) {
    if (int st = wait_tree(t, stream)) return st;
    EvalView v = view_of(t);
    const size_t nq = t->p;
    hipLaunchKernelGGL(eval_x2_kernel, dim3(blocks_for(nq, 256)), dim3(256), 0, stream, v, (int)nq,
                       eta0, eta1, t->theta_sq, d_re0, d_im0, d_re1, d_im1);
    SHK_HIP(hipGetLastError());
    return 0;
}

}  // namespace shk
