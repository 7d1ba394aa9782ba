/*
 * shrinkers_stieltjes.h -- C ABI of the MI355X Stieltjes-transform library.
 *
 * Drop-in boundary for the `src/stieltjes` method API of alexisrosuel/shrinkers
 * (Rust).  Every entry point takes plain pointers and sizes; none takes a
 * torch or HIP type (streams are passed as `void*` and may be NULL).
 *
 * Conventions, as in the reference (src/stieltjes/mod.rs:22-34,
 * src/stieltjes/term.rs:3-7):
 *   S_i = sum_j 1/((x_i - lam_j) - i*eta), diagonal term included,
 *   Re S = sum d/(d^2+eta^2), Im S = eta * sum 1/(d^2+eta^2) > 0.
 *   "raw" results are these sums; "scaled" results are multiplied by 1/p.
 *   Outputs are caller-owned structure-of-arrays (re[], im[]).
 *
 * Error convention: every function returns an shk_status and never aborts.
 * The reference's Rust API returns empty output for p = 0 and panics on NaN;
 * its Python boundary raises ValueError (src/python.rs:138-318).  Here p = 0
 * returns SHK_OK with nothing written (like the Rust API), except where a
 * function says otherwise; the text of the last error of the calling thread
 * is available from shk_last_error().
 *
 * Host-pointer entry points copy inputs to the GPU, run on an internal
 * stream of the current device and return after the results are back in
 * host memory.  Device-pointer entry points (suffix _dev) enqueue work on the
 * given stream and return without synchronising; the caller keeps the
 * buffers alive until the stream has passed the work.  The exception is a
 * ChebCode build (shk_chebcode_build_dev, and the ChebCode methods of
 * shk_all_points_dev, shk_at_points_dev and shk_all_points_many_dev): the
 * tree's node count sizes the rest of its build, so the call waits once for
 * its topology kernel, and with it for the work queued before it on the
 * stream.  The ChebCode workspace entries (shk_chebcode_all_points_ws_dev,
 * shk_chebcode_at_points_ws_dev) read every size on the device and do not
 * wait.
 *
 * HIP graph capture: entry points whose work is captured whole are
 * shk_all_points_noalloc_dev, the exact at-points transforms at a grid of
 * points (shk_at_points_dev with an exact method and no cutoff, and
 * shk_exact_at_points_dev, for at most 4 096 points and at least 8
 * eigenvalues per point: one workgroup per point, no scratch), the
 * evaluations of a built tree
 * (shk_chebcode_evaluate_dev, shk_chebcode_evaluate_many_dev; SHK_ERR_BAD_ARG
 * if the tree's build has not finished) and the prepared batch
 * (shk_batch_run, which captures its own graph), and the ChebCode one-shot
 * transforms in a caller's workspace (shk_chebcode_all_points_ws_dev,
 * shk_chebcode_at_points_ws_dev).  On a stream under capture,
 * a ChebCode build (including the ChebCode methods of shk_all_points_dev and
 * shk_at_points_dev), shk_all_points_many_dev, and every other _dev entry
 * point that needs scratch memory (the cached buffers below cannot be shared
 * with a graph) return SHK_ERR_BAD_ARG before enqueuing any work, so the
 * capture stays valid.
 *
 * Thread safety: calls on distinct streams may run concurrently; scratch
 * memory is stream-ordered per call (pool allocations, or the calling
 * thread's cached buffers, whose reuse on another stream waits for the
 * previous use).
 *
 * Memory: scratch buffers and ChebCode trees built through the handle API
 * come from the device's default stream-ordered memory pool.  On first use
 * the library raises that pool's release threshold so freed blocks stay in
 * the pool for reuse (the pool then holds as much as the largest calls
 * needed until the process ends).  The ChebCode build scratch, and the
 * trees of one-shot ChebCode transforms (all-points and at-points calls),
 * live in buffers cached per calling thread and device, which grow to the
 * largest size used; shk_release_cached_memory() frees them.
 */
#ifndef SHRINKERS_STIELTJES_H
#define SHRINKERS_STIELTJES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum shk_status {
    SHK_OK = 0,
    SHK_ERR_EMPTY = 1,            /* empty input where one is required        */
    SHK_ERR_NONFINITE = 2,        /* NaN or infinity in the input             */
    SHK_ERR_BAD_ETA = 3,          /* eta not a positive finite number         */
    SHK_ERR_BAD_METHOD = 4,       /* unknown method id or name                */
    SHK_ERR_UNSUPPORTED = 5,      /* method known to the reference but not
                                     built for the GPU (FFT/NUFFT/HODLR/...) */
    SHK_ERR_BAD_ARG = 6,          /* NULL pointer, bad size or bad option     */
    SHK_ERR_HIP = 7,              /* a HIP runtime call failed                */
    SHK_ERR_NO_DEVICE = 8         /* no usable gfx950 device                  */
} shk_status;

/* Method vocabulary, in the order of StieltjesMethod
 * (src/config.rs:19-89); names as in StieltjesMethod::name (:92-118) and the
 * Python aliases of parse_method (src/python.rs:321-350). */
typedef enum shk_method {
    SHK_NAIVE = 0,
    SHK_AUTOVEC = 1,
    SHK_BLOCKED = 2,
    SHK_BLOCKED_AUTOVEC = 3,
    SHK_BLOCKED_TILED = 4,
    SHK_BLOCKED_WINDOWED = 5,
    SHK_BLOCKED_HYBRID = 6,
    SHK_ADAPTIVE = 7,
    SHK_FFT5 = 8,
    SHK_FFT3 = 9,
    SHK_FFT2 = 10,
    SHK_NUFFT = 11,
    SHK_TREECODE = 12,
    SHK_CHEBCODE = 13,
    SHK_CHEBCODE_FAST = 14,
    SHK_CHEBCODE_XTREME = 15,
    SHK_CHEBCODE_BALANCED = 16,
    SHK_EWALD = 17,
    SHK_DST = 18,
    SHK_AUTO = 19,
    SHK_HODLR = 20,
    SHK_SPEED_AUTO = 21,
    SHK_ACCURACY_AUTO = 22
} shk_method;

/* Arithmetic of the exact family.
 *   SHK_PREC_F64  : FP64 everywhere (the correctness anchor).
 *   SHK_PREC_F32C : FP32 pair arithmetic on eigenvalues split into f32
 *                   hi/lo parts relative to a per-tile centre, FP32 partial
 *                   sums over short runs folded into FP64 accumulators, FP64
 *                   cross-block reduction.  Normwise relative error ~1e-7.
 *                   Range: 1e-18 <= eta <= 1e18 (the diagonal term 1/eta^2
 *                   and its f32 runs stay in f32's normal range) and
 *                   |eigenvalue| <= 1e18.  Forward calls with eta outside
 *                   that range run the FP64 kernels instead; host-array calls
 *                   also do so for larger eigenvalues, while device-pointer
 *                   calls take the magnitude bound as a precondition.  The
 *                   reverse-mode entries use FP32-compensated arithmetic for
 *                   1e-9 <= eta <= 1e18 only (their terms are G/D^2), and
 *                   FP64 otherwise.
 *   SHK_PREC_F32  : the reference's f32 arm (input and output float,
 *                   src/stieltjes/mod.rs:805-851); runs the F32C kernel
 *                   (FP64 outside the range above). */
typedef enum shk_precision {
    SHK_PREC_F64 = 0,
    SHK_PREC_F32C = 1,
    SHK_PREC_F32 = 2
} shk_precision;

/* ---- library ------------------------------------------------------------ */

/* Text of the last error raised on the calling thread ("" if none). */
const char *shk_last_error(void);

/* Library version string. */
const char *shk_version(void);

/* Frees the buffers the library caches for the calling thread on the
 * current device (ChebCode build scratch, one-shot ChebCode trees), after
 * their last uses have finished.  No counterpart in the reference; for
 * long-running callers that want the memory back. */
shk_status shk_release_cached_memory(void);

/* Resolve a Python method name or alias (src/python.rs:321-350) to its id.
 * Returns SHK_ERR_BAD_METHOD for an unknown name. */
shk_status shk_method_from_name(const char *name, int *method_out);

/* Canonical name of a method id (src/config.rs:92-118), NULL if invalid. */
const char *shk_method_name(int method);

/* The reference's default eta = 0.1/sqrt(p) (src/stieltjes/mod.rs:105-108). */
double shk_default_eta(size_t p);

/* Eigenvalue counts [*min_p, *max_p] for which the all-points entries
 * (shk_compute_all_stieltjes, shk_all_points_dev) take the symmetric kernel,
 * which evaluates each unordered pair once, for precision SHK_PREC_F64 or
 * SHK_PREC_F32C (with the exact family, no cutoff).  No counterpart in the
 * reference (its sequential sweep is always symmetric and its parallel path
 * never, cacheblock.rs:841-954, 1454-1513); for callers and benchmarks that
 * report which kernel ran. */
shk_status shk_symmetric_range(int precision, size_t *min_p, size_t *max_p);

/* Caps the device scratch that the symmetric kernels' partial sums may take
 * (forward and adjoint; default 512 MiB, bytes = 0 restores it).  When a
 * transform's partials (16 * ceil(p/512) * p bytes) exceed the cap, the
 * triangle of tile pairs runs in bands of rows whose partials fit, each band
 * followed by an in-order fold; results are the same bits for every cap.
 * Process-wide; calls in flight keep the cap they started with.  No
 * counterpart in the reference (its kernels keep no partial sums). */
shk_status shk_set_symmetric_scratch_limit(size_t bytes);

/* Resolve the auto presets (src/config.rs:543-555) for this library, by
 * size as the reference's measured table does (src/config/pareto_autogen.rs):
 * Auto/SpeedAuto -> BLOCKED for p <= 20 000 (the exact transform is the
 * faster one there on MI355X), CHEBCODE_FAST above; AccuracyAuto -> BLOCKED.
 * The at-points entries resolve Auto/SpeedAuto by the pair count instead
 * (BLOCKED up to 2e8 query-eigenvalue pairs), and shk_all_points_many_dev by
 * its largest spectrum.  A concrete method is returned unchanged. */
int shk_resolve_auto_method(int method, int parallel, size_t p);

/* ---- exact and treecode all-points transform ---------------------------- */

/* compute_all_stieltjes (src/stieltjes/mod.rs:601-790): transform at every
 * eigenvalue, SCALED by 1/p.  cutoff_ratio <= 0 disables the far-field
 * cutoff (CutoffConfig::Disabled).  `parallel` is accepted for interface
 * parity; the GPU path is the same for both values.  precision is an
 * shk_precision (SHK_PREC_F32 is rejected here: use the _f32 entry). */
shk_status shk_compute_all_stieltjes(const double *eigenvalues, size_t p, double eta,
                                     int method, double cutoff_ratio, int parallel,
                                     int precision, double *out_re, double *out_im);

/* compute_all_stieltjes_f32 (src/stieltjes/mod.rs:805-851): float input,
 * float output, scaled by 1/p. */
shk_status shk_compute_all_stieltjes_f32(const float *eigenvalues, size_t p, float eta,
                                         int method, double cutoff_ratio, float *out_re,
                                         float *out_im);

/* compute_stieltjes_at_points (src/stieltjes/mod.rs:394-494): RAW sums at
 * arbitrary query points. */
shk_status shk_compute_stieltjes_at_points(const double *query_points, size_t nq,
                                           const double *eigenvalues, size_t p, double eta,
                                           int method, double cutoff_ratio, int parallel,
                                           double *out_re, double *out_im);

/* compute_all_stieltjes_with_deriv (src/stieltjes/autovec.rs:63-76): RAW
 * values and their derivative w.r.t. the real query point,
 * S'(x) = -sum 1/(x - lam_j - i*eta)^2. */
shk_status shk_compute_all_stieltjes_with_deriv(const double *eigenvalues, size_t p, double eta,
                                                double *val_re, double *val_im,
                                                double *der_re, double *der_im);

/* compute_stieltjes_scaled_ray (src/stieltjes/mod.rs:304-319): RAW sums at
 * z_i = lam_i (1 + i*scale), i.e. eta_i = scale*lam_i per query. */
shk_status shk_compute_stieltjes_scaled_ray(const double *eigenvalues, size_t p, double scale,
                                            double *out_re, double *out_im);

/* ---- device-pointer variants --------------------------------------------- */

/* All-points transform on device buffers.  scaled != 0 multiplies by 1/p.
 * method must be a concrete exact-family or ChebCode id (autos resolved). */
shk_status shk_all_points_dev(const double *d_eigenvalues, size_t p, double eta, int method,
                              double cutoff_ratio, int precision, int scaled,
                              double *d_out_re, double *d_out_im, void *stream);

/* shk_all_points_dev for the exact family (no cutoff) without any
 * allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree or
 * allocating library routine: the symmetric kernels (each unordered pair
 * once) run band by band on a 64 MiB device array of the library, one per
 * device, held from library load until the process ends.  p <= 262 144
 * (SHK_ERR_BAD_ARG above).  precision SHK_PREC_F64 or SHK_PREC_F32C (FP64
 * for eta outside [1e-18, 1e18]); scaled != 0 multiplies by 1/p.  For p in
 * shk_symmetric_range the results equal shk_all_points_dev's bit for bit.
 * Calls share their device's array: calls that overlap on one device (other
 * streams, other threads) race and must be serialised by the caller. */
shk_status shk_all_points_noalloc_dev(const double *d_eigenvalues, size_t p, double eta,
                                      int precision, int scaled, double *d_out_re,
                                      double *d_out_im, void *stream);

/* Many independent spectra in one call (no counterpart in the reference,
 * which batches only across eta on one spectrum, chebcode.rs:1312).
 * d_eigenvalues holds the spectra back to back; offsets (host, n_spectra+1
 * entries, offsets[0] = 0) delimit them; etas (host) gives one eta each.
 * Each spectrum's result is its all-points transform for the method and
 * precision (scaled by 1/p_s when scaled != 0), written at the same offsets
 * of d_out_re / d_out_im.  For the exact family it is bit for bit the
 * split-tile single-spectrum kernel's: shk_all_points_dev's, except where
 * shk_all_points_dev takes the symmetric kernel (shk_symmetric_range: 8 192
 * to 4 194 304 FP32-compensated eigenvalues, 18 432 to 4 194 304 in FP64) (same
 * as shk_exact_at_points_dev with the spectrum as its own query points, raw).
 * If precision is SHK_PREC_F32C and some eta lies outside [1e-18, 1e18], the
 * whole call runs in FP64.
 * n_streams = 0 (recommended): for the exact family, ONE segmented launch on
 * `stream` covers every spectrum (one workgroup per spectrum and 128-query
 * tile, largest spectra first; no partial-sum scratch); for ChebCode, as
 * n_streams = 4.  n_streams = 1..8: the library enqueues the single-spectrum
 * kernels, largest spectrum first, on that many streams of its own, each
 * spectrum on the stream with the least work so far.  Either way the batch
 * is ordered after earlier work on `stream`, and later work on `stream`
 * waits for all of it.  If times_ms is not NULL (2*n_spectra floats) the
 * call synchronises and writes each spectrum's start and end in ms from the
 * batch start (stream pool: device events; segmented launch: device wall
 * clock at its first workgroup's start and last workgroup's end, from the
 * launch's earliest workgroup start), otherwise it returns without
 * synchronising.  For a fixed set of spectra run repeatedly, shk_batch_*
 * (below) keeps the table on the device and replays the launch from a
 * hipGraph. */
shk_status shk_all_points_many_dev(const double *d_eigenvalues, const int64_t *offsets,
                                   size_t n_spectra, const double *etas, int method,
                                   int precision, int scaled, double *d_out_re,
                                   double *d_out_im, int n_streams, float *times_ms,
                                   void *stream);

/* Raw sums at device query points (exact kernel or ChebCode per method). */
shk_status shk_at_points_dev(const double *d_query, size_t nq, const double *d_eigenvalues,
                             size_t p, double eta, int method, double cutoff_ratio,
                             double *d_out_re, double *d_out_im, void *stream);

/* ---- ChebCode one-shot transforms in a caller's workspace ------------------
 *
 * The ChebCode methods of shk_all_points_dev / shk_at_points_dev
 * (compute_all_stieltjes_chebcode_preset, src/stieltjes/chebcode.rs:1015-1120,
 * and the at-points arm, src/stieltjes/mod.rs:461-479) with every size kept
 * on the device: no host wait and no allocation, so the whole call can be
 * captured in a HIP graph and replayed on new eigenvalues of the same count.
 * The reference has no such form; these entries are for callers that
 * capture or cannot let a call wait on the host.
 *
 * The eigenvalues must be ascending (as symmetric eigensolvers return them).
 * On ascending input the results equal shk_all_points_dev's /
 * shk_at_points_dev's for the same method bit for bit.  Input that is not
 * ascending gives NaN outputs, and shk_chebcode_workspace_status reports it
 * after the fact (NaN input gives NaN outputs too; the device entries do not
 * scan for it).  method: a ChebCode preset (13..16); the auto methods are
 * refused (SHK_ERR_BAD_METHOD).
 *
 * The workspace (device memory, 256-byte aligned, at least
 * shk_chebcode_workspace_bytes(p, method) bytes) holds the tree at its largest
 * possible node count (2p plus the top walk's slots), so it is larger than
 * the tree the call builds: about 620 bytes per eigenvalue at the FAST order
 * and 770 at n = 11, plus 0.6 MB (p = 1e6: 0.62 / 0.77 GB).  Calls that share
 * a workspace must not overlap (the same stream orders them). */
shk_status shk_chebcode_workspace_bytes(size_t p, int method, size_t *bytes);

/* Scaled (1/p) when scaled != 0, else raw sums, at every eigenvalue. */
shk_status shk_chebcode_all_points_ws_dev(const double *d_eigenvalues, size_t p, double eta,
                                          int method, int scaled, void *d_workspace,
                                          size_t workspace_bytes, double *d_out_re,
                                          double *d_out_im, void *stream);

/* Raw sums at nq device query points (the tree at grid_leaf_cap's leaf size,
 * as shk_at_points_dev builds it). */
shk_status shk_chebcode_at_points_ws_dev(const double *d_query, size_t nq,
                                         const double *d_eigenvalues, size_t p, double eta,
                                         int method, void *d_workspace, size_t workspace_bytes,
                                         double *d_out_re, double *d_out_im, void *stream);

/* *ascending = 1 if the last call on this workspace found its eigenvalues
 * ascending (its outputs are the transform), 0 if not (its outputs are NaN).
 * Waits for the stream. */
shk_status shk_chebcode_workspace_status(const void *d_workspace, void *stream,
                                         int *ascending);

/* ---- ChebCodeBatch (src/stieltjes/chebcode.rs:1148-1362) ----------------- */

typedef struct shk_chebcode shk_chebcode;

/* ChebCodeBatch::build_mode: build the tree once on the GPU.  Input may be
 * unsorted (a permutation is kept, :1186-1217).  f32_far != 0 selects the
 * FAST f32 far field (FastMode::F32). */
shk_status shk_chebcode_build(const double *eigenvalues, size_t p, double theta, int n,
                              size_t leaf_cap, int f32_far, shk_chebcode **out);

/* ChebCodeBatch::build_preset with a ChebCode method id (13..16). */
shk_status shk_chebcode_build_preset(const double *eigenvalues, size_t p, int method,
                                     shk_chebcode **out);

/* Number of tree nodes, and the eigenvalue count. */
shk_status shk_chebcode_info(const shk_chebcode *h, size_t *node_count, size_t *p);

/* Copy the tree out (pre-order; any pointer may be NULL): node intervals,
 * source index ranges, right-child links (-1 = leaf), Chebyshev nodes and
 * weights (node_count*n each). */
shk_status shk_chebcode_export(const shk_chebcode *h, double *lo, double *hi, int64_t *lo_idx,
                               int64_t *hi_idx, int32_t *right, double *nodes, double *w);

/* ChebCodeBatch::evaluate (:1280-1306): RAW sums at the eigenvalues, in the
 * order they were given to the build. */
shk_status shk_chebcode_evaluate(const shk_chebcode *h, double eta, double *out_re,
                                 double *out_im);

/* ChebCodeBatch::evaluate_points (:1241-1277): RAW sums at query points. */
shk_status shk_chebcode_evaluate_points(const shk_chebcode *h, const double *query, size_t nq,
                                        double eta, double *out_re, double *out_im);

/* ChebCodeBatch::evaluate_many (:1312-1361): one RAW result row per eta,
 * out_re/out_im are n_eta*p row-major. */
shk_status shk_chebcode_evaluate_many(const shk_chebcode *h, const double *etas, size_t n_eta,
                                      double *out_re, double *out_im);

shk_status shk_chebcode_destroy(shk_chebcode *h);

/* Device-resident ChebCode: build from device eigenvalues on `stream`
 * (returns once the tree's node count is known; later work on any stream
 * waits for the build), and evaluate at the tree's own eigenvalues (input
 * order) into device buffers, multiplied by `scale`.  Same semantics as
 * shk_chebcode_build / shk_chebcode_evaluate. */
shk_status shk_chebcode_build_dev(const double *d_eigenvalues, size_t p, double theta, int n,
                                  size_t leaf_cap, int f32_far, void *stream, shk_chebcode **out);
shk_status shk_chebcode_evaluate_dev(const shk_chebcode *h, double eta, double scale,
                                     double *d_out_re, double *d_out_im, void *stream);

/* ChebCodeBatch::evaluate_many (chebcode.rs:1312-1361) on device buffers:
 * one RAW result row per eta (host array etas), d_out_re / d_out_im are
 * n_eta*p row-major on the device, in build input order; two etas per
 * traversal (acceptance at the smaller), enqueued on `stream` without
 * synchronising.  Bit for bit shk_chebcode_evaluate_many without the host
 * copies. */
shk_status shk_chebcode_evaluate_many_dev(const shk_chebcode *h, const double *etas,
                                          size_t n_eta, double *d_out_re, double *d_out_im,
                                          void *stream);

/* Evaluation at the tree's own eigenvalues (input order, times scale) with
 * local expansions instead of the per-query panel sums of
 * FlatChebTree::contribution_f64 (chebcode.rs:607-711): no counterpart in
 * the reference, whose notes leave a Chebyshev M2L to future work
 * (docs/internals.md:580-599).  Target nodes are the topmost nodes no wider
 * than eta/4; each walks the tree once with the reference's acceptance test
 * taken at its query nearest to the panel (so a panel is used only where the
 * reference would use it for every query of the target), evaluates the
 * accepted panels at its n Chebyshev points (M2L), and each query
 * interpolates them (L2P) and adds the exact sums of the leaves the walk
 * reached.  A leaf wider than eta/4 shares one such walk among its queries
 * and sums the accepted panels at each query; a target whose walk outgrows
 * its lists (16 leaves, 64 panels, a 256-node front) keeps the reference's
 * per-query walk.  Trees with an f64 far field and n = 11 or 12
 * (SHK_ERR_BAD_ARG otherwise).  Error: that of the reference traversal or
 * smaller, plus the interpolation error ~ (8 + sqrt 65)^-n (5e-14 at
 * n = 11).  Faster than shk_chebcode_evaluate_dev from about 2e5
 * eigenvalues (on MP spectra at eta = 1/sqrt(p): 1.1-1.2x at 2e5, 1.6-2.4x
 * at 1e6; 0.8x at 5e4, where the latency of one target's walk bounds the
 * launch).  Scratch from the calling thread's cached buffers.  Enqueued on
 * `stream` without synchronising. */
shk_status shk_chebcode_evaluate_local_dev(const shk_chebcode *h, double eta, double scale,
                                           double *d_out_re, double *d_out_im, void *stream);

/* Work of one traversal pass, with the evaluation kernels' node decisions:
 * *near_pairs = sum over queries of the sizes of the leaves each visits
 * (exact pair evaluations), *far_terms = n per accepted panel.  d_query =
 * NULL (with nq = 0) counts the all-points pass at the tree's own sources.
 * Counterpart of the near/far statistics the reference's traversal implies
 * (chebcode.rs:607-828); used to express evaluation time as FLOP rates.
 * Synchronises `stream`. */
shk_status shk_chebcode_count_work(const shk_chebcode *h, const double *d_query, size_t nq,
                                   double eta, int64_t *near_pairs, int64_t *far_terms,
                                   void *stream);

/* ---- batched spectra (no counterpart in the reference) ------------------- */

/* Many independent spectra as one prepared launch replayed from a hipGraph.
 * The reference batches only across eta on one spectrum
 * (ChebCodeBatch::evaluate_many, chebcode.rs:1312); each spectrum's result
 * here is its own all-points transform (compute_all_stieltjes with the exact
 * family), in FP32-compensated arithmetic, scaled by 1/p_s when scaled != 0.
 *
 * d_eigenvalues: all spectra concatenated on the device (kept alive by the
 * caller for the batch's life); offsets: host array of n_spectra+1 starts
 * (offsets[0] = 0, every spectrum non-empty); etas: host array of one eta per
 * spectrum.  Outputs share the layout of d_eigenvalues.
 *
 * The launch is the segmented kernel of shk_all_points_many_dev (one
 * workgroup per spectrum and query tile, largest spectra first, no
 * partial-sum scratch); the batch keeps its spectrum table on the device and
 * captures the launch in a hipGraph on first use (again when the outputs or
 * the tile height change).  Every run stamps each workgroup's device wall
 * clock for shk_batch_spectrum_times.  With the default tile height r = 2
 * every result equals the split-tile single-spectrum kernel's bit for bit
 * (shk_all_points_dev's outside its symmetric range, see above);
 * r = 1 or 4 (64- or 256-query tiles) agree within the FP32-compensated
 * bound.  r = 8 runs the symmetric kernel instead (tiles of 512, each
 * unordered pair once): the tile pairs of all spectra, in groups of spectra
 * whose partial sums fit the symmetric scratch limit
 * (shk_set_symmetric_scratch_limit, read when r = 8 is first prepared), each
 * group's launch followed by an in-order fold; every result then equals the
 * single-spectrum symmetric transform's bit for bit (shk_all_points_dev's
 * inside its symmetric range).  r = 8 is refused when one spectrum's partials
 * (16 * ceil(p/512) * p bytes) alone exceed that limit.  A batch is used from
 * one thread at a time; its runs share the stamp buffer, so a run's times
 * are read before the next run. */
typedef struct shk_batch shk_batch;

shk_status shk_batch_create(const double *d_eigenvalues, const int64_t *offsets, size_t n_spectra,
                            const double *etas, int scaled, shk_batch **out);
/* Enqueue the graph on `stream`; returns without synchronising. */
shk_status shk_batch_run(shk_batch *b, double *d_out_re, double *d_out_im, void *stream);
/* Time each configuration r in {1, 2, 4, 8} on this batch (three replays
 * each, synchronising `stream`; r = 8 only where it is accepted) and keep
 * the fastest; the chosen r and its time are returned. */
shk_status shk_batch_tune(shk_batch *b, double *d_out_re, double *d_out_im, void *stream,
                          int *best_r, double *best_ms);
/* Set the configuration r: 1, 2 or 4 queries per lane of the split-tile
 * kernel, or 8 for the symmetric kernel. */
shk_status shk_batch_set_config(shk_batch *b, int r);
/* Per-spectrum start/end of the last run in microseconds from that run's
 * first workgroup start (device wall clock); the caller synchronises first.
 * Fails before the first run. */
shk_status shk_batch_spectrum_times(const shk_batch *b, double *start_us, double *end_us);
shk_status shk_batch_destroy(shk_batch *b);

/* ---- reverse mode (no counterpart in the reference) ---------------------- */

/* The reference differentiates the transform only with respect to the query
 * point (stieltjes_with_deriv_sum, src/stieltjes/autovec.rs:38-59, exposed
 * as stieltjes_transform_with_deriv, src/python.rs:758-781).  These entries
 * give a differentiable pair on device data:
 *
 * shk_exact_at_points_dev: raw exact sums S_m = sum_j 1/((x_m - lam_j) - i eta)
 * at device query points, in SHK_PREC_F64 or SHK_PREC_F32C (the at-points
 * counterpart of compute_stieltjes_at_points with the exact family,
 * src/stieltjes/mod.rs:394-494, with a precision choice). */
shk_status shk_exact_at_points_dev(const double *d_query, size_t nq, const double *d_eigenvalues,
                                   size_t p, double eta, int precision, double *d_out_re,
                                   double *d_out_im, void *stream);

/* Adjoint of shk_exact_at_points_dev.  Given d_grad_re[m], d_grad_im[m] =
 * dL/dRe S_m, dL/dIm S_m of a real loss L, writes dL/dlam_j (p values),
 * dL/dx_m (nq values) and dL/deta (one device double); each output pointer
 * may be NULL to skip it.  With d = x_m - lam_j, D = d^2 + eta^2:
 *   dL/dlam_j = sum_m (G_re (d^2 - eta^2) + 2 G_im eta d) / D^2
 *   dS_m/dx   = sum_j ((eta^2 - d^2) - 2 i eta d) / D^2   (autovec.rs:38-59)
 *   dL/dx_m   = G_re Re dS/dx + G_im Im dS/dx
 *   dL/deta   = sum_m G_re Im dS/dx - G_im Re dS/dx       (dS/deta = -i dS/dx)
 * For the all-points transform (queries = eigenvalues) the caller adds the
 * query and eigenvalue gradients. */
shk_status shk_exact_adjoint_dev(const double *d_query, size_t nq, const double *d_eigenvalues,
                                 size_t p, double eta, const double *d_grad_re,
                                 const double *d_grad_im, int precision,
                                 double *d_grad_eigenvalues, double *d_grad_query,
                                 double *d_grad_eta, void *stream);

/* Adjoint of the all-points transform (raw sums S_i = sum_j
 * 1/((lam_i - lam_j) - i eta), diagonal included).  Given d_grad_re[i],
 * d_grad_im[i] = dL/dRe S_i, dL/dIm S_i, writes the total dL/dlam_i (the
 * eigenvalue as query and as source: the sum of shk_exact_adjoint_dev's
 * eigenvalue and query gradients at queries = eigenvalues) and dL/deta (one
 * device double; NULL skips it).  A pair (i, j) adds T to dL/dlam_i and -T to
 * dL/dlam_j, so each unordered pair is evaluated once (symmetric kernels,
 * p <= 4 194 304; the pair (i, i), whose T is 0, is left out of the
 * gradient sums, so no 1/eta^2 terms cancel).  Larger p add the two sweeps
 * of shk_exact_adjoint_dev in FP64.  SHK_PREC_F32C gives FP32-compensated
 * arithmetic for 1e-9 <= eta <= 1e18 and FP64 otherwise (an eta gradient
 * alone, d_grad_eigenvalues = NULL, runs the query sweep).
 * Deterministic.  The
 * reference differentiates with respect to the query only (autovec.rs:38-59)
 * and has no counterpart. */
shk_status shk_exact_adjoint_all_points_dev(const double *d_eigenvalues, size_t p, double eta,
                                            const double *d_grad_re, const double *d_grad_im,
                                            int precision, double *d_grad_eigenvalues,
                                            double *d_grad_eta, void *stream);

/* El Karoui bulk fit, FP64.  The residual of the Marchenko-Pastur equation
 * as written in src/deconvolution/mod.rs:6 for a discrete population
 * H = sum_k w_k delta_{t_k}:
 *   a_m = 1 - c - c z_m g_m,  r_m = g_m - sum_k w_k / (t_k a_m - z_m),
 *   loss = (1/M) sum_m |r_m|^2,  z_m = x_m + i eta,
 * where g_m is the sample Stieltjes transform in the reference inversion's
 * convention (Im g > 0, scaled by 1/p: g = -conj(S)/p for the raw sums S,
 * src/deconvolution/mod.rs:210-213).  Writes the loss and the gradients
 * with respect to Re g, Im g (M values each), the weights (K values) and eta
 * through z only (one device double; the part through g is the adjoint's).
 * Every output pointer may be NULL.  The reference inverts the equation
 * pointwise (mod.rs:116-259) and has no fit. */
shk_status shk_el_karoui_fit_dev(const double *d_x, const double *d_g_re, const double *d_g_im,
                                 size_t m, double eta, double c, const double *d_atoms,
                                 const double *d_weights, size_t k, double *d_loss,
                                 double *d_grad_g_re, double *d_grad_g_im,
                                 double *d_grad_weights, double *d_grad_eta, void *stream);

/* One El Karoui fit step on device, in one call: the raw exact sums at
 * z_m = x_m + i eta (precision SHK_PREC_F64 or SHK_PREC_F32C), g = -conj(S)/p,
 * the loss of shk_el_karoui_fit_dev, and its gradients with respect to the
 * eigenvalues (p values), eta (one device double, through z and through S)
 * and the weights (k values).  The same arithmetic as composing
 * shk_exact_at_points_dev, shk_el_karoui_fit_dev and shk_exact_adjoint_dev,
 * without the host round trips between them.  Gradient pointers may be NULL.
 * No reference counterpart (see shk_el_karoui_fit_dev). */
shk_status shk_el_karoui_fit_step_dev(const double *d_eigenvalues, size_t p, const double *d_x,
                                      size_t m, double eta, double c, const double *d_atoms,
                                      const double *d_weights, size_t k, int precision,
                                      double *d_loss, double *d_grad_eigenvalues,
                                      double *d_grad_eta, double *d_grad_weights, void *stream);

/* ---- measurement ---------------------------------------------------------- */

/* When enabled, the library brackets the dominant kernel of each exact or
 * treecode evaluation with HIP events on the stream it is launched on and
 * accumulates the elapsed time (read after the stream has synchronised). */
shk_status shk_profile_enable(int enable);
shk_status shk_profile_read(double *kernel_ms_total, int64_t *launches);
shk_status shk_profile_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* SHRINKERS_STIELTJES_H */
