// Internal declarations shared by the HIP translation units of
// libshrinkers_hip.so.  Not part of the C ABI (see include/).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace shk {

// ---- error plumbing --------------------------------------------------------

void set_error(const std::string &msg);
const char *last_error();

// Checks once that a gfx950 device is visible (SHK_ERR_NO_DEVICE otherwise).
int device_ready();

// ---- per-thread device buffer cache ---------------------------------------
// One buffer per slot, per calling thread and device, grown on demand (and
// never shrunk until cache_free_all).  Each slot records an event at its
// last use; a use on another stream first waits for it, so reuse never
// races with work still in flight.  Growing waits for the last use and
// frees the old buffer.
enum CacheSlot : int {
    CACHE_CHEB_SCRATCH = 0,  // tree build scratch
    CACHE_CHEB_SRC = 1,      // transient tree: sorted sources
    CACHE_CHEB_NODES = 2,    // transient tree: node records and panels
    CACHE_EXACT_PART = 3,    // exact / symmetric kernels: FP64 partial sums
    CACHE_ADJ_PART = 4,      // adjoint sweeps: FP64 partial sums
    CACHE_EK_SCR = 5,        // El Karoui fit kernels: per-point scratch
    CACHE_EK_STEP = 6,       // fused fit step: sample transform and its gradient
    CACHE_ADJ_TMP = 7,       // all-points adjoint by two sweeps: the query gradient
    CACHE_CHEB_LOCAL = 8,    // ChebCode local-expansion evaluation: target list
    CACHE_HOST_0 = 9,        // host-array calls: the k-th device copy of a call
    kCacheHostSlots = 6,     //   takes slot CACHE_HOST_0 + k
    kCacheSlots = CACHE_HOST_0 + kCacheHostSlots
};
// Returns the slot's buffer (>= bytes) in *out, ordered on stream s.
int cache_acquire(int slot, size_t bytes, hipStream_t s, void **out);
// Marks the end of the slot's use by the work enqueued on s so far.
int cache_release(int slot, hipStream_t s);
// Frees the calling thread's cached buffers on the current device (waits
// for their last uses).
int cache_free_all();

// A cache slot held for the rest of the enclosing scope: get() acquires it on
// stream s, and leaving the scope marks the end of its use there.
struct CachedBuf {
    int slot;
    hipStream_t s;
    void *p = nullptr;
    CachedBuf(int sl, hipStream_t st) : slot(sl), s(st) {}
    CachedBuf(const CachedBuf &) = delete;
    CachedBuf &operator=(const CachedBuf &) = delete;
    template <class T>
    int get(size_t count, T **out) {
        const int st = cache_acquire(slot, sizeof(T) * count, s, &p);
        *out = static_cast<T *>(p);
        return st;
    }
    ~CachedBuf() {
        if (p) (void)cache_release(slot, s);
    }
};

// Evaluates a HIP call; on failure records the message and returns
// SHK_ERR_HIP (7) from the enclosing function.
#define SHK_HIP(expr)                                                           \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::shk::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return 7;                                                           \
        }                                                                       \
    } while (0)

// Propagates a non-zero status.
#define SHK_TRY(expr)                  \
    do {                               \
        int s_ = (expr);               \
        if (s_ != 0) return s_;        \
    } while (0)

// ---- measurement -----------------------------------------------------------

// Brackets the dominant kernel of an evaluation with events on `stream` when
// profiling is enabled.  begin() returns an opaque token consumed by end().
struct ProfileMark {
    hipEvent_t a = nullptr, b = nullptr;
};
ProfileMark profile_begin(hipStream_t stream);
void profile_end(const ProfileMark &m, hipStream_t stream);

// ---- exact family (shk_exact.hip) -------------------------------------------

// Range of eta in which the FP32-compensated kernels are used.  They form
// eta^2 and 1/(d^2 + eta^2) in f32 and add runs of up to 32 such terms, so
// the diagonal term 1/eta^2 and its runs must stay inside f32's normal range
// (eta^2 >= 1e-36 > 1.2e-38, 32/eta^2 <= 3.2e37 < 3.4e38).  Outside it the
// entry points run the FP64 kernels instead (more accurate, never worse).
// The adjoint's FP32 terms are G/D^2, so its floor is higher: eta >=
// kSymAdjMinEta (below).
constexpr double kF32cMinEta = 1.0e-18;
constexpr double kF32cMaxEta = 1.0e18;
// Largest |eigenvalue| the FP32-compensated kernels accept: pair differences
// are formed in FP64 relative to a tile centre and rounded to f32, and d^2
// must not overflow f32.  Host-array entry points check it and run FP64
// beyond it; device entry points document it as a precondition.
constexpr double kF32cMaxAbs = 1.0e18;
inline bool f32c_eta_ok(double eta) { return eta >= kF32cMinEta && eta <= kF32cMaxEta; }

// Which accumulation the exact kernels perform.
enum ExactKind : int {
    EXACT_PLAIN = 0,   // re = sum d*inv, im = eta * sum inv
    EXACT_CUTOFF = 1,  // both parts skip |d| > cut
    EXACT_HYBRID = 2,  // re from every term, im skips |d| > cut
    EXACT_DERIV = 3,   // values + d/dx (autovec.rs:38-59)
    EXACT_RAY = 4,     // eta_i = scale * x_i per query (mod.rs:304-319)
};

// Raw or scaled exact sums at nq query points against p sources, FP64.
//   eta      : eta (EXACT_RAY: the ray scale h)
//   cut      : absolute cutoff distance (ratio*eta) for CUTOFF/HYBRID
//   scale    : multiplier applied to every output (1 or 1/p)
//   out[0..n_out) : re, im (+ deriv re, deriv im for EXACT_DERIV)
//   workspace: device scratch of exact_workspace_doubles(nq, p, kind)
//              doubles, or NULL to allocate it stream-ordered per call
int exact_f64(const double *q, size_t nq, const double *src, size_t p, double eta,
              ExactKind kind, double cut, double scale, double *const *out,
              hipStream_t stream, double *workspace = nullptr);

// FP32-compensated exact sums (no cutoff).  out_re/out_im are double, or
// float when out_f32 != 0.  workspace as for exact_f64 (kind EXACT_PLAIN).
int exact_f32c(const double *q, size_t nq, const double *src, size_t p, double eta,
               double scale, void *out_re, void *out_im, int out_f32, hipStream_t stream,
               double *workspace = nullptr);

// Scratch doubles the exact kernels need for nq queries against p sources.
size_t exact_workspace_doubles(size_t nq, size_t p, ExactKind kind);
// Threads of a grid-point workgroup (below); each sums every kGridThreads-th
// source.
constexpr int kGridThreads = 512;
// A few query points against many sources (at most 4 096 points, at least 8
// sources per point): the forward and adjoint sums at such points run one
// workgroup per point instead of the split plan, with no partial sums.  The
// kernels index sources in 32 bits and step up to 16 kGridThreads past the
// last one, so p stays that far below 2^31.
inline bool grid_queries(size_t nq, size_t p) {
    return nq <= 4096 && 8 * nq <= p && p <= (size_t)INT32_MAX - 16 * (size_t)kGridThreads;
}
// In-order sum of n_split FP64 partials laid out [re/im][split][query], times
// scale, with eta applied to the imaginary part (kind EXACT_PLAIN).
int exact_reduce_plain(const double *part, int n_split, size_t nq, double eta, double scale,
                       const double *q, double *out_re, double *out_im, hipStream_t stream);

// Symmetric all-points kernels (each unordered pair once, credited to both
// ends); partials of 16 * nt * p bytes, nt = ceil(p/512).  The FP32-
// compensated one is used by the all-points entry for kSymMinP <= p <=
// kSymMaxP.
// With the one-dimensional grid of tile pairs the symmetric kernel is ahead
// from about 1e4 (measured on MI355X, all-points entry: 4.5 % at 1e4, 7 % at
// 1.4e4, 10 % at 2.2e4, 7 % at 2.4e4; equal at 3e4 where it already ran).
#ifndef SHK_SYM_MIN_P
#define SHK_SYM_MIN_P 8192
#endif
constexpr size_t kSymMinP = SHK_SYM_MIN_P;
// Upper end of the symmetric kernels' range.  Their partials are bounded by
// running the triangle of tile pairs in bands (SymBand): one band when the
// whole triangle's partials fit in kSymScratchBudget bytes, otherwise bands of
// R rows whose partials fit, folded in slot order after each band.
constexpr size_t kSymMaxP = (size_t)1 << 22;
constexpr size_t kSymScratchBudget = (size_t)512 << 20;
// The current budget: kSymScratchBudget unless shk_set_symmetric_scratch_limit
// set another (process-wide).
size_t sym_scratch_budget();
// Tiles of the symmetric kernels (forward and adjoint).
constexpr int kSymTile = 512;

// One band of rows [r0, r1) of a symmetric kernel's triangle of tile pairs
// (a, b), a <= b, and where its partials go.  Tile t's slot m holds the
// partial of the pair {t, m} at t's points.  Band tiles (r0 <= t < r1) get
// their slots m in [r0, nt) in A (the j-sides of band rows m < t and the
// i-sides of row t); tiles beyond the band (t >= r1) get the band rows'
// slots m in [r0, r1) in B.  Slots m < r0 were folded by earlier bands.
// With r0 = 0, r1 = nt, A is the single-launch layout part[c][m][i].
struct SymBand {
    double *A;  // [2][nt - r0][na]: points [r0 T, r0 T + na)
    double *B;  // [2][r1 - r0][nb]: points [r1 T, r1 T + nb)
    int r0, r1, nt, na, nb;
    // Component c's partial for tile t, slot m: pointer at the tile's first point.
    __host__ __device__ double *slot(int c, int t, int m) const {
        if (t < r1)
            return A + (size_t)(c * (nt - r0) + (m - r0)) * na + (size_t)(t - r0) * kSymTile;
        return B + (size_t)(c * (r1 - r0) + (m - r0)) * nb + (size_t)(t - r1) * kSymTile;
    }
};

// Rows per band for p eigenvalues with `avail` doubles of scratch: nt (one
// band) when the whole triangle's partials fit (2 nt p doubles), else the
// largest R whose band partials and running sums fit, 0 if none does.
int sym_band_rows(size_t p, size_t avail);
// Scratch doubles a run with R rows per band uses.
size_t sym_scratch_doubles(size_t p, int R);
// Launches one band's kernel: grid = the band's tile pairs, id0 = the index
// of its first pair in the row-major triangle.
using SymLaunch = int (*)(const SymBand &bd, int grid, int id0, void *ctx);
// Runs the triangle band by band on `scratch` (avail doubles): each band's
// launch, then an in-order fold of its partials -- final outputs for the band
// tiles, out0 = s0 * scale and out1 = (mul * s1) * scale, running sums for
// the tiles beyond.  The additions per point happen in slot order 0..nt-1
// for every band count, so results are the same bits for every budget.
int sym_run_bands(size_t p, double *scratch, size_t avail, double mul, double scale,
                  double *out0, double *out1, SymLaunch launch, void *ctx, hipStream_t stream);

// f64 != 0: the FP64 kernel (pair reciprocals as in rcp_pairs_f64), used for
// kSym64MinP <= p <= kSymMaxP.
int exact_sym(const double *ev, size_t p, double eta, double scale, int f64, double *out_re,
              double *out_im, hipStream_t stream);
// The same on caller-provided scratch of scratch_doubles doubles (at least
// sym_scratch_doubles(p, 1)); the band height follows from its size.
int exact_sym_on(const double *ev, size_t p, double eta, double scale, int f64, double *out_re,
                 double *out_im, double *scratch, size_t scratch_doubles, hipStream_t stream);
// FP64 crossover (measured on MI355X): split ahead by 19 % at 1.64e4,
// symmetric ahead by 9 % at 2e4, 52 % at 5e4 and 57 % at 1e5.
#ifndef SHK_SYM64_MIN_P
#define SHK_SYM64_MIN_P 18432
#endif
constexpr size_t kSym64MinP = SHK_SYM64_MIN_P;

// One spectrum of a segmented all-points launch (device table entry).
struct SegSpectrum {
    int64_t off;    // index of its first eigenvalue (and first output)
    int32_t p;      // eigenvalues
    int32_t per;    // chunk length: the single-spectrum split length for p
    int32_t tile0;  // its first query tile in the launch
    int32_t pad;
    double eta;
    double scale;   // 1/p or 1
};

// Fills table[0, *n_table) for the non-empty spectra in `order` (offsets[k],
// offsets[k+1] bound spectrum k) and counts their query tiles of 64*r
// queries (r = 1, 2 or 4 queries per lane).
int exact_many_plan(const int64_t *offsets, const size_t *order, size_t n, const double *etas,
                    int scaled, SegSpectrum *table, int *n_table, int *n_tiles, int r = 2);
// All-points sums (kind EXACT_PLAIN) of every spectrum in the device table in
// one launch, written at each spectrum's offset in out_re/out_im.  With r = 2
// (the single-spectrum tiles) each result is bit-equal to exact_f32c /
// exact_f64 on that spectrum alone; FP32-compensated r = 1 or 4 use other
// tiles (their own centres and pair forms).  FP64 takes r = 2 only.
// d_stamps, if not NULL, receives 2 * n_tiles device wall-clock ticks: each
// workgroup's start and end.  The table must match r.
int exact_many(const double *ev, const SegSpectrum *d_table, int n_table, int n_tiles,
               int f32c, double *out_re, double *out_im, long long *d_stamps,
               hipStream_t stream, int r = 2);

// Segmented symmetric launch (prepared batches): the FP32-compensated
// symmetric kernel over the tile pairs of many spectra, in groups of
// consecutive table entries whose partials fit the scratch.  Per spectrum:
struct SymSegSpectrum {
    int64_t off;    // index of its first eigenvalue (and first output)
    int64_t part;   // its first partial (a double) in the group's scratch
    int32_t p, nt;  // eigenvalues, tiles of kSymTile
    int32_t pair0;  // its first workgroup in the group's launch
    int32_t fold0;  // its first workgroup in the group's fold
    double eta;
    double scale;   // 1/p or 1
};
// Per group: table entries [first, first + count), `pairs` workgroups (tile
// pairs) from global pair index pair_base on, `folds` fold workgroups, and
// the scratch doubles its partials take.
struct SymSegGroup {
    int first, count, pairs, folds;
    int64_t pair_base;
    size_t part_doubles;
};
// Fills the table (spectra in `order`, every one non-empty) and cuts it into
// groups whose partials (2 * nt * p doubles per spectrum) fit budget_doubles.
// Fails when one spectrum's partials alone exceed the budget.  *n_pairs
// receives the total tile pairs (the stamp count is twice that).
int sym_many_plan(const int64_t *offsets, const size_t *order, size_t n, const double *etas,
                  int scaled, size_t budget_doubles, SymSegSpectrum *table,
                  std::vector<SymSegGroup> &groups, int64_t *n_pairs);
// Enqueues each group's kernel and in-order fold on `stream`: every result
// equals the single-spectrum symmetric FP32-compensated transform (exact_sym)
// of that spectrum bit for bit.  d_table holds the whole table; scratch holds
// the largest group's partials; d_stamps, if not NULL, receives 2 * n_pairs
// device wall-clock ticks (each workgroup's start and end, by global pair).
int sym_many(const double *ev, const SymSegSpectrum *d_table,
             const std::vector<SymSegGroup> &groups, double *scratch, double *out_re,
             double *out_im, long long *d_stamps, hipStream_t stream);

// float -> double widening on device (for the f32 entry points).
int widen_f32(const float *in, double *out, size_t n, hipStream_t stream);

// Per-thread, per-device pinned host mailbox (coherent host memory that
// kernels write directly): mail[0] is a sequence number a kernel stores last,
// with system-scope release, after the payload in mail[kMailData, ...).
// Lives until the process ends.
constexpr int kMailData = 16;
constexpr int kMailInts = 256;
int host_mailbox(int **mail);
// Waits until mail[0] == seq: spins on the host for up to spin_us, then
// synchronises `stream` and checks again (a launch that failed or faulted
// reports its error there).
int mailbox_wait(volatile int *mail, int seq, hipStream_t stream, int spin_us = 2000);

// ---- reverse mode (shk_adjoint.hip) -----------------------------------------

// Gradients of a real loss of the raw at-points sums, given the upstream
// gradients gre/gim (w.r.t. Re S_m, Im S_m).  Outputs: grad_ev (p),
// grad_q (nq), grad_eta (one device double); any may be NULL.  f32c != 0
// selects the FP32-compensated pair arithmetic.
// accumulate_eta != 0 adds the eta gradient to *grad_eta instead of storing it.
int exact_adjoint(const double *q, size_t nq, const double *ev, size_t p, double eta,
                  const double *gre, const double *gim, int f32c, double *grad_ev,
                  double *grad_q, double *grad_eta, hipStream_t stream, int accumulate_eta = 0);

// Gradient of a real loss of the raw all-points sums S_i = sum_j
// 1/((lam_i - lam_j) - i eta) with respect to every eigenvalue (query and
// source roles together) into grad (p), and with respect to eta into
// *grad_eta (may be NULL; accumulate_eta != 0 adds to it).  Symmetric
// kernels (each unordered pair once; shk_symadj.hip), p <= kSymMaxP; f64 != 0
// selects FP64, else FP32-compensated.
int exact_adjoint_sym(const double *ev, size_t p, double eta, const double *gre,
                      const double *gim, int f64, double *grad, double *grad_eta,
                      hipStream_t stream, int accumulate_eta = 0);

// The FP32-compensated adjoint kernels need eta >= kSymAdjMinEta, so that
// r^2 = 1/D^2 stays within f32 range.
constexpr double kSymAdjMinEta = 1.0e-9;
// eta range of the FP32-compensated adjoint kernels (terms G/D^2 in f32).
inline bool adj_f32c_eta_ok(double eta) { return eta >= kSymAdjMinEta && eta <= kF32cMaxEta; }

// *out = sum(v[0, n)) in a fixed order (one workgroup); accumulate != 0 adds
// to *out.
int adjoint_ordered_sum(const double *v, size_t n, double *out, int accumulate,
                        hipStream_t stream);

// out[i] += in[i] (one kernel); n = 0 is a no-op.
int add_into(double *out, const double *in, size_t n, hipStream_t stream);

// Writes n zeros with a kernel (graph-capturable); n = 0 is a no-op.
int zero_fill(double *p, size_t n, hipStream_t stream);

// El Karoui fit: loss = mean_m |g_m - sum_k w_k / (t_k a_m - z_m)|^2 with
// a_m = 1 - c - c z_m g_m, z_m = x_m + i eta, and its gradients with respect
// to (Re g, Im g), the weights and eta (through z only).  Device pointers;
// every output may be NULL.
int ek_fit(const double *x, const double *g_re, const double *g_im, size_t m, double eta,
           double c, const double *t, const double *w, size_t k, double *loss,
           double *grad_g_re, double *grad_g_im, double *grad_w, double *grad_eta,
           hipStream_t stream);

// One El Karoui fit step on device: the raw exact sums S at z_m = x_m + i eta
// (FP64 or FP32-compensated), g = -conj(S)/p, the fit loss and the gradients
// with respect to the eigenvalues (p), eta (one double: the direct term plus
// the term through S) and the weights (k).  Gradient outputs may be NULL.
int ek_fit_step(const double *ev, size_t p, const double *x, size_t m, double eta, double c,
                const double *t, const double *w, size_t k, int f32c, double *loss,
                double *grad_ev, double *grad_eta, double *grad_w, hipStream_t stream);

// ---- ChebCode (shk_chebcode.hip) --------------------------------------------

struct ChebTree;  // device-resident tree

struct ChebParams {
    double theta;
    int n;
    int64_t leaf_cap;
    int f32_far;
};

// Builds the tree on the GPU from device eigenvalues (any order).  A
// transient tree (one evaluated on its build stream and destroyed by the same
// thread before its next build) keeps its arrays in that thread's buffer
// cache instead of allocating them; its destruction then needs no
// synchronisation, and its evaluations no event.
int cheb_build(const double *d_ev, size_t p, const ChebParams &prm, ChebTree **out,
               hipStream_t stream, int transient = 0);
// Build + evaluation in a caller's workspace with no host wait and no
// allocation (every size read on the device), for ascending input of the
// presets' orders: the transient build's kernels and results.  The sums
// go to nq device query points, or to the sources themselves when
// queries_are_sources (d_query unused), times scale.  Input that is not
// ascending gets NaN outputs, which cheb_ws_status reports.
size_t cheb_ws_bytes(size_t p, int n);
int cheb_oneshot_ws(const double *d_ev, size_t p, const ChebParams &prm, const double *d_query,
                    size_t nq, double eta, double scale, int queries_are_sources, void *ws,
                    size_t ws_bytes, double *d_out_re, double *d_out_im, hipStream_t stream);
// Whether the last workspace build found its input ascending.  Synchronises.
int cheb_ws_status(const void *ws, hipStream_t stream, int *ascending);
void cheb_destroy(ChebTree *t);
size_t cheb_node_count(const ChebTree *t);
size_t cheb_p(const ChebTree *t);
int cheb_params(const ChebTree *t, ChebParams *prm);
// Copies the pre-order tree arrays to host memory (NULL pointers skipped).
int cheb_export(const ChebTree *t, double *lo, double *hi, int64_t *lo_idx, int64_t *hi_idx,
                int32_t *right, double *nodes, double *w, hipStream_t stream);
// Raw sums at nq device query points; order_by_perm != 0 means the queries
// are the tree's own sorted eigenvalues and the outputs are scattered back
// to the build input order.  scale multiplies the outputs.
int cheb_eval(const ChebTree *t, const double *d_query, size_t nq, double eta, double scale,
              int queries_are_sources, double *d_out_re, double *d_out_im,
              hipStream_t stream);
// Traversal work at nq device query points (NULL: the tree's own sources):
// near-field pair evaluations and far-field panel terms.  Synchronises.
int cheb_count(const ChebTree *t, const double *d_query, size_t nq, double eta,
               int64_t *near_pairs, int64_t *far_terms, hipStream_t stream);
// Local-expansion evaluation at the tree's own points (input order, times
// scale): M2L into the Chebyshev points of target nodes no wider than eta/4,
// L2P per query, exact sums over the leaves the target walks reach; f64 far
// field, n = 11 or 12.
int cheb_eval_local(const ChebTree *t, double eta, double scale, double *d_out_re,
                    double *d_out_im, hipStream_t stream);
// Two-eta traversal for evaluate_many: acceptance at min(eta0, eta1).
int cheb_eval_x2(const ChebTree *t, double eta0, double eta1, double *d_re0, double *d_im0,
                 double *d_re1, double *d_im1, hipStream_t stream);

}  // namespace shk
