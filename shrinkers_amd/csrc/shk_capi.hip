// C ABI of libshrinkers_hip.so: validation, method dispatch, device memory
// and the measurement hooks.  Declarations and contracts are in
// include/shrinkers_stieltjes.h; each entry point cites the reference
// function it replaces.
#include "shk_internal.hpp"
#include "shrinkers_stieltjes.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <atomic>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace shk {

namespace {
thread_local std::string g_err;
// Scratch budget of the symmetric kernels' partial sums (bytes).
std::atomic<size_t> g_sym_budget{shk::kSymScratchBudget};

std::atomic<int> g_prof_on{0};
std::mutex g_prof_mu;
double g_prof_ms = 0.0;
int64_t g_prof_n = 0;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_pending;
}  // namespace

void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

ProfileMark profile_begin(hipStream_t stream) {
    ProfileMark m;
    if (!g_prof_on.load(std::memory_order_relaxed)) return m;
    // no marks inside a graph capture (the graph's launch is marked instead)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return m;
    if (hipEventCreate(&m.a) != hipSuccess || hipEventCreate(&m.b) != hipSuccess) {
        if (m.a) (void)hipEventDestroy(m.a);
        return ProfileMark{};
    }
    (void)hipEventRecord(m.a, stream);
    return m;
}

void profile_end(const ProfileMark &m, hipStream_t stream) {
    if (!m.a) return;
    (void)hipEventRecord(m.b, stream);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_pending.emplace_back(m.a, m.b);
}

}  // namespace shk

using namespace shk;

namespace {

struct MethodName {
    const char *name;
    int id;
};

// Python names and aliases: src/python.rs:321-350.
const MethodName kNames[] = {
    {"naive", SHK_NAIVE},
    {"autovec", SHK_AUTOVEC},
    {"blocked", SHK_BLOCKED},
    {"blocked_autovec", SHK_BLOCKED_AUTOVEC},
    {"blocked_tiled", SHK_BLOCKED_TILED},
    {"blocked_windowed", SHK_BLOCKED_WINDOWED},
    {"blocked_hybrid", SHK_BLOCKED_HYBRID},
    {"adaptive", SHK_ADAPTIVE},
    {"fft5", SHK_FFT5},
    {"fft3", SHK_FFT3},
    {"fft2", SHK_FFT2},
    {"nufft", SHK_NUFFT},
    {"fft_fast", SHK_NUFFT},
    {"fmm", SHK_TREECODE},
    {"treecode", SHK_TREECODE},
    {"chebcode", SHK_CHEBCODE},
    {"chebyshev", SHK_CHEBCODE},
    {"chebcode_fast", SHK_CHEBCODE_FAST},
    {"chebf", SHK_CHEBCODE_FAST},
    {"chebcode_xtreme", SHK_CHEBCODE_XTREME},
    {"chebx", SHK_CHEBCODE_XTREME},
    {"chebcode_balanced", SHK_CHEBCODE_BALANCED},
    {"chebb", SHK_CHEBCODE_BALANCED},
    {"ewald", SHK_EWALD},
    {"dst", SHK_DST},
    {"auto", SHK_AUTO},
    {"hodlr", SHK_HODLR},
    {"speed_auto", SHK_SPEED_AUTO},
    {"speed", SHK_SPEED_AUTO},
    {"accuracy_auto", SHK_ACCURACY_AUTO},
    {"accuracy", SHK_ACCURACY_AUTO},
};

// StieltjesMethod::name: src/config.rs:92-118.
const char *kCanonical[] = {"naive",         "autovec",         "blocked",
                            "blocked_autovec", "blocked_tiled", "blocked_windowed",
                            "blocked_hybrid", "adaptive",       "fft5",
                            "fft3",          "fft2",            "nufft",
                            "fmm",           "chebcode",        "chebcode_fast",
                            "chebcode_xtreme", "chebcode_balanced", "ewald",
                            "dst",           "auto",            "hodlr",
                            "speed_auto",    "accuracy_auto"};
constexpr int kNumMethods = 23;

enum Family { FAM_EXACT, FAM_CHEB, FAM_UNSUPPORTED, FAM_BAD };

Family family_of(int m) {
    switch (m) {
        case SHK_NAIVE:
        case SHK_AUTOVEC:
        case SHK_BLOCKED:
        case SHK_BLOCKED_AUTOVEC:
        case SHK_BLOCKED_TILED:
        case SHK_BLOCKED_WINDOWED:
        case SHK_BLOCKED_HYBRID:
            return FAM_EXACT;
        case SHK_CHEBCODE:
        case SHK_CHEBCODE_FAST:
        case SHK_CHEBCODE_XTREME:
        case SHK_CHEBCODE_BALANCED:
            return FAM_CHEB;
        case SHK_ADAPTIVE:
        case SHK_FFT5:
        case SHK_FFT3:
        case SHK_FFT2:
        case SHK_NUFFT:
        case SHK_TREECODE:
        case SHK_EWALD:
        case SHK_DST:
        case SHK_HODLR:
            return FAM_UNSUPPORTED;
        default:
            return FAM_BAD;
    }
}

// ChebPreset constants: src/stieltjes/chebcode.rs:942-973.
ChebParams preset_of(int m) {
    switch (m) {
        case SHK_CHEBCODE_FAST:
            return ChebParams{1.0, 8, 32, 1};
        case SHK_CHEBCODE_XTREME:
            return ChebParams{0.25, 11, 16, 0};
        case SHK_CHEBCODE_BALANCED:
            return ChebParams{0.55, 11, 32, 0};
        case SHK_CHEBCODE:
        default:
            return ChebParams{0.5, 11, 32, 0};
    }
}

// grid_leaf_cap: src/stieltjes/mod.rs:361-370.
int64_t grid_leaf_cap(int64_t n, int64_t p, int64_t nq, int64_t preset_leaf) {
    const int64_t kMaxGridLeaf = 8192;
    if (nq == 0 || p == 0) return preset_leaf;
    const double optimal = (double)n * std::sqrt(2.0 * (double)p / (double)nq);
    int64_t o = (int64_t)std::llround(optimal);
    const int64_t hi = std::max(kMaxGridLeaf, preset_leaf);
    return std::min(std::max(o, preset_leaf), hi);
}

shk_status fail(shk_status s, const std::string &msg) {
    set_error(msg);
    return s;
}

bool all_finite(const double *v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

double max_abs(const double *v, size_t n) {
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m = std::max(m, std::fabs(v[i]));
    return m;
}

shk_status check_eta(double eta) {
    if (!(std::isfinite(eta) && eta > 0.0))
        return fail(SHK_ERR_BAD_ETA, "eta must be a positive finite number, got " +
                                         std::to_string(eta));
    return SHK_OK;
}

std::string method_label(int m) {
    if (m >= 0 && m < kNumMethods) return kCanonical[m];
    return std::to_string(m);
}

shk_status check_method(int m) {
    switch (family_of(m)) {
        case FAM_EXACT:
        case FAM_CHEB:
            return SHK_OK;
        case FAM_UNSUPPORTED:
            return fail(SHK_ERR_UNSUPPORTED,
                        "method '" + method_label(m) +
                            "' has no GPU implementation in this library (supported: the "
                            "exact family and the chebcode presets)");
        default:
            return fail(SHK_ERR_BAD_METHOD, "unknown method id " + std::to_string(m));
    }
}

// Device readiness: a gfx950 device must be current.
shk_status ensure_device() {
    static std::once_flag once;
    static shk_status state = SHK_OK;
    static std::string msg;
    std::call_once(once, [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            state = SHK_ERR_NO_DEVICE;
            msg = "no HIP device is visible";
            return;
        }
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
            state = SHK_ERR_NO_DEVICE;
            msg = "hipGetDeviceProperties failed";
            return;
        }
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            state = SHK_ERR_NO_DEVICE;
            msg = std::string("libshrinkers_hip is built for gfx950, device is ") +
                  prop.gcnArchName;
            return;
        }
        // Scratch and trees come from the default stream-ordered pool; keep
        // freed blocks in the pool for reuse instead of unmapping them at
        // every synchronisation.
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    });
    if (state != SHK_OK) set_error(msg);
    return state;
}

// Per-thread stream for the host entry points (default priority).
hipStream_t host_stream() {
    thread_local hipStream_t s = nullptr;
    if (!s) {
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    }
    return s;
}

// Device copies of a host-array call's inputs and outputs, in buffers cached
// per thread: the k-th buffer a call takes is cache slot CACHE_HOST_0 + k, so
// a call allocates nothing once the buffers are large enough.  A call's
// buffers are released together when it returns.
thread_local int g_host_slots = 0;

struct DevBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    int slot = -1;
    ~DevBuf() {
        if (slot >= 0) {
            (void)cache_release(slot, s);
            --g_host_slots;
        }
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

// Pinned host staging for a host-array call: the caller's arrays are copied
// into it (and out of it) with memcpy, and the device copies run from pinned
// memory.  One buffer per thread, grown to the largest call; a call uses it
// only between its start and its final stream synchronisation.
int stage_acquire(size_t bytes, char **out) {
    thread_local char *base = nullptr;
    thread_local size_t cap = 0;
    if (cap < bytes) {
        if (base) SHK_HIP(hipHostFree(base));
        base = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4 + 4096;
        void *p = nullptr;
        SHK_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
        base = static_cast<char *>(p);
        cap = want;
    }
    *out = base;
    return 0;
}

inline size_t stage_round(size_t b) { return (b + 255) / 256 * 256; }

// The host<->device copies of one host-array call, run through the thread's
// pinned staging buffer.  `reserve` sizes the buffer for every upload and
// download of the call at once (it must not move while copies are queued);
// `finish` synchronises the stream and copies the downloads out to the
// caller's arrays.  A call that returns early still waits for its queued
// copies, so the next call never reuses staging memory a copy is using.
class HostStage {
  public:
    explicit HostStage(hipStream_t s) : s_(s) {}
    HostStage(const HostStage &) = delete;
    HostStage &operator=(const HostStage &) = delete;
    ~HostStage() {
        if (queued_) (void)hipStreamSynchronize(s_);
    }
    int reserve(std::initializer_list<size_t> sizes) {
        size_t total = 0;
        for (size_t b : sizes) total += stage_round(b);
        cap_ = total;
        used_ = 0;
        return stage_acquire(total, &base_);
    }
    int upload(void *dev, const void *host, size_t bytes) {
        char *at = take(bytes);
        if (!at) return overflow();
        std::memcpy(at, host, bytes);
        queued_ = true;
        SHK_HIP(hipMemcpyAsync(dev, at, bytes, hipMemcpyHostToDevice, s_));
        return 0;
    }
    // With host == nullptr the bytes stay in staging; `staged` (if given)
    // points at them, valid after `finish` until the call returns.
    int download(void *host, const void *dev, size_t bytes, const char **staged = nullptr) {
        char *at = take(bytes);
        if (!at || n_out_ == kMaxOut) return overflow();
        queued_ = true;
        SHK_HIP(hipMemcpyAsync(at, dev, bytes, hipMemcpyDeviceToHost, s_));
        out_[n_out_++] = {host, at, bytes};
        if (staged) *staged = at;
        return 0;
    }
    int finish() {
        SHK_HIP(hipStreamSynchronize(s_));
        queued_ = false;
        for (int k = 0; k < n_out_; ++k)
            if (out_[k].host) std::memcpy(out_[k].host, out_[k].staged, out_[k].bytes);
        n_out_ = 0;
        return 0;
    }

  private:
    static constexpr int kMaxOut = 8;
    struct Out {
        void *host;
        const char *staged;
        size_t bytes;
    };
    char *take(size_t bytes) {
        const size_t r = stage_round(bytes);
        if (!base_ || used_ + r > cap_) return nullptr;
        char *at = base_ + used_;
        used_ += r;
        return at;
    }
    static int overflow() {
        set_error("host-array call: staging copies exceed the reserved size");
        return SHK_ERR_HIP;
    }
    hipStream_t s_;
    char *base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    Out out_[kMaxOut];
    int n_out_ = 0;
    bool queued_ = false;
};

int dev_alloc(DevBuf &b, size_t bytes, hipStream_t s) {
    if (g_host_slots >= kCacheHostSlots) {
        set_error("host-array call: more device buffers than cache slots");
        return SHK_ERR_HIP;
    }
    const int slot = CACHE_HOST_0 + g_host_slots;
    if (int st = cache_acquire(slot, bytes ? bytes : 8, s, &b.p)) return st;
    b.s = s;
    b.slot = slot;
    ++g_host_slots;
    return 0;
}

// Exact-family kind for an all-points call (src/stieltjes/mod.rs:709-790,
// dispatch_exact_blocked :562-594).
ExactKind allpoints_kind(int m, double cutoff_ratio) {
    const bool cut = cutoff_ratio > 0.0;
    switch (m) {
        case SHK_NAIVE:
        case SHK_AUTOVEC:
            return EXACT_PLAIN;  // the per-point kernels ignore the cutoff
        case SHK_BLOCKED_HYBRID:
            return cut ? EXACT_HYBRID : EXACT_PLAIN;
        default:
            return cut ? EXACT_CUTOFF : EXACT_PLAIN;
    }
}

// Exact-family kind at arbitrary query points (mod.rs:425-446, 480-491).
ExactKind atpoints_kind(int m, double cutoff_ratio) {
    const bool cut = cutoff_ratio > 0.0;
    if (m == SHK_NAIVE || m == SHK_AUTOVEC) return EXACT_PLAIN;
    return cut ? EXACT_CUTOFF : EXACT_PLAIN;
}

int run_allpoints(const double *d_ev, size_t p, double eta, int method, double cutoff_ratio,
                  int precision, double scale, double *d_re, double *d_im, hipStream_t s) {
    const int fam = family_of(method);
    if (fam == FAM_CHEB) {
        // transient tree: its arrays are the thread's cached buffers, so the
        // evaluation stays asynchronous and destruction needs no sync
        ChebTree *t = nullptr;
        SHK_TRY(cheb_build(d_ev, p, preset_of(method), &t, s, 1));
        const int st = cheb_eval(t, nullptr, p, eta, scale, 1, d_re, d_im, s);
        cheb_destroy(t);
        return st;
    }
    const ExactKind kind = allpoints_kind(method, cutoff_ratio);
    if (precision == SHK_PREC_F32C && kind == EXACT_PLAIN && f32c_eta_ok(eta)) {
        if (p >= kSymMinP && p <= kSymMaxP)
            return exact_sym(d_ev, p, eta, scale, 0, d_re, d_im, s);
        return exact_f32c(d_ev, p, d_ev, p, eta, scale, d_re, d_im, 0, s);
    }
    if (kind == EXACT_PLAIN && p >= kSym64MinP && p <= kSymMaxP)
        return exact_sym(d_ev, p, eta, scale, 1, d_re, d_im, s);
    double *outs[2] = {d_re, d_im};
    return exact_f64(d_ev, p, d_ev, p, eta, kind, cutoff_ratio * eta, scale, outs, s);
}

int run_atpoints(const double *d_q, size_t nq, const double *d_ev, size_t p, double eta,
                 int method, double cutoff_ratio, double *d_re, double *d_im, hipStream_t s) {
    if (family_of(method) == FAM_CHEB) {
        ChebParams prm = preset_of(method);
        prm.leaf_cap = grid_leaf_cap(prm.n, (int64_t)p, (int64_t)nq, prm.leaf_cap);
        ChebTree *t = nullptr;
        SHK_TRY(cheb_build(d_ev, p, prm, &t, s, 1));
        const int st = cheb_eval(t, d_q, nq, eta, 1.0, 0, d_re, d_im, s);
        cheb_destroy(t);
        return st;
    }
    double *outs[2] = {d_re, d_im};
    return exact_f64(d_q, nq, d_ev, p, eta, atpoints_kind(method, cutoff_ratio),
                     cutoff_ratio * eta, 1.0, outs, s);
}

}  // namespace

namespace shk {
int device_ready() { return ensure_device(); }
size_t sym_scratch_budget() { return g_sym_budget.load(std::memory_order_relaxed); }
}  // namespace shk

// ---- per-thread device buffer cache ----------------------------------------

namespace {

struct CacheEntry {
    void *p = nullptr;
    size_t cap = 0;
    hipEvent_t last = nullptr;     // recorded at the end of the last use
    hipStream_t last_stream = nullptr;
    bool used = false;
};

struct ThreadCache {
    // per device id; the buffers live until cache_free_all or process end
    // (no destructor: the HIP runtime may already be gone at thread exit)
    std::vector<std::array<CacheEntry, kCacheSlots>> dev;
    CacheEntry *entry(int slot) {
        int d = 0;
        if (hipGetDevice(&d) != hipSuccess || d < 0) return nullptr;
        if ((size_t)d >= dev.size()) dev.resize((size_t)d + 1);
        return &dev[(size_t)d][(size_t)slot];
    }
};

ThreadCache &thread_cache() {
    thread_local ThreadCache tc;
    return tc;
}

}  // namespace

namespace shk {

int cache_acquire(int slot, size_t bytes, hipStream_t s, void **out) {
    *out = nullptr;
    CacheEntry *c = (slot >= 0 && slot < kCacheSlots) ? thread_cache().entry(slot) : nullptr;
    if (!c) {
        set_error("buffer cache: no current device");
        return SHK_ERR_HIP;
    }
    // A cached buffer cannot serve a graph: replays would share it with
    // later eager calls, and its event would be recorded inside the capture.
    // Calls that need scratch are refused under capture (before their first
    // kernel); shk_all_points_noalloc_dev serves captured callers.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SHK_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {
        set_error("this call needs scratch memory, which the library does not allocate under "
                  "graph capture (a captured call cannot share the cached buffers); "
                  "shk_all_points_noalloc_dev runs the all-points transform captured");
        return SHK_ERR_BAD_ARG;
    }
    if (!c->last) SHK_HIP(hipEventCreateWithFlags(&c->last, hipEventDisableTiming));
    if (c->used && c->last_stream != s) SHK_HIP(hipStreamWaitEvent(s, c->last, 0));
    if (c->cap < bytes) {
        if (c->p) {
            SHK_HIP(hipEventSynchronize(c->last));
            SHK_HIP(hipFree(c->p));
            c->p = nullptr;
            c->cap = 0;
        }
        const size_t cap = bytes + bytes / 4 + 4096;  // headroom for slowly growing sizes
        SHK_HIP(hipMalloc(&c->p, cap));
        c->cap = cap;
    }
    *out = c->p;
    return 0;
}

int cache_release(int slot, hipStream_t s) {
    CacheEntry *c = (slot >= 0 && slot < kCacheSlots) ? thread_cache().entry(slot) : nullptr;
    if (!c || !c->last) return 0;
    SHK_HIP(hipEventRecord(c->last, s));
    c->last_stream = s;
    c->used = true;
    return 0;
}

int cache_free_all() {
    for (int slot = 0; slot < kCacheSlots; ++slot) {
        CacheEntry *c = thread_cache().entry(slot);
        if (!c || !c->p) continue;
        if (c->used) SHK_HIP(hipEventSynchronize(c->last));
        SHK_HIP(hipFree(c->p));
        c->p = nullptr;
        c->cap = 0;
        c->used = false;
    }
    return 0;
}

int host_mailbox(int **mail) {
    // per device id; no destructor (the HIP runtime may already be gone at
    // thread exit), the pinned pages go with the process
    thread_local std::vector<int *> boxes;
    int d = 0;
    SHK_HIP(hipGetDevice(&d));
    if ((size_t)d >= boxes.size()) boxes.resize((size_t)d + 1, nullptr);
    if (!boxes[(size_t)d]) {
        void *p = nullptr;
        SHK_HIP(hipHostMalloc(&p, sizeof(int) * kMailInts, hipHostMallocCoherent));
        std::memset(p, 0, sizeof(int) * kMailInts);
        boxes[(size_t)d] = static_cast<int *>(p);
    }
    *mail = boxes[(size_t)d];
    return 0;
}

int mailbox_wait(volatile int *mail, int seq, hipStream_t stream, int spin_us) {
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(const_cast<int *>(mail), __ATOMIC_ACQUIRE) != seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(spin_us)) {
            SHK_HIP(hipStreamSynchronize(stream));
            if (__atomic_load_n(const_cast<int *>(mail), __ATOMIC_ACQUIRE) != seq) {
                set_error("host mailbox: the kernel finished without writing it");
                return SHK_ERR_HIP;
            }
            break;
        }
    }
    return 0;
}

}  // namespace shk

namespace {

// Streams and events the multi-spectrum entry point reuses: created on
// first use on the calling thread (default priority), kept until the thread
// ends.
constexpr int kMaxManyStreams = 8;
constexpr int kDefaultManyStreams = 4;  // n_streams = 0 on the stream-pool path
struct ManyPools {
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done;    // one per stream, no timing
    std::vector<hipEvent_t> events;  // batch start + per-spectrum start/end
    int reserve(int n_streams, size_t n_events) {
        while ((int)streams.size() < n_streams) {
            hipStream_t st = nullptr;
            hipEvent_t ev = nullptr;
            if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
                set_error("could not create a stream for the multi-spectrum launch");
                return SHK_ERR_HIP;
            }
            streams.push_back(st);
            done.push_back(ev);
        }
        while (events.size() < n_events) {
            hipEvent_t ev = nullptr;
            if (hipEventCreate(&ev) != hipSuccess) {
                set_error("could not create a timing event");
                return SHK_ERR_HIP;
            }
            events.push_back(ev);
        }
        return 0;
    }
    // Pinned host copy of the segmented launch's spectrum table.  `staged`
    // marks the end of the last copy out of it; it is waited on before the
    // table is rewritten, so back-to-back calls never overwrite a pending copy.
    SegSpectrum *staging = nullptr;
    size_t staging_cap = 0;
    hipEvent_t staged = nullptr;
    int reserve_table(size_t n) {
        if (!staged && hipEventCreateWithFlags(&staged, hipEventDisableTiming) != hipSuccess) {
            set_error("could not create an event for the segmented launch");
            return SHK_ERR_HIP;
        }
        if (hipEventSynchronize(staged) != hipSuccess) {
            set_error("waiting for the previous spectrum-table copy failed");
            return SHK_ERR_HIP;
        }
        if (staging_cap >= n) return 0;
        if (staging) (void)hipHostFree(staging);
        staging = nullptr;
        staging_cap = 0;
        const size_t cap = std::max<size_t>(n, 1024);
        if (hipHostMalloc((void **)&staging, cap * sizeof(SegSpectrum), hipHostMallocDefault) !=
            hipSuccess) {
            set_error("could not allocate the pinned spectrum table");
            return SHK_ERR_HIP;
        }
        staging_cap = cap;
        return 0;
    }
};

ManyPools &many_pools() {
    thread_local ManyPools pools;
    return pools;
}

#define SHK_CHECK(expr)                     \
    do {                                    \
        int s_ = (expr);                    \
        if (s_ != 0) return (shk_status)s_; \
    } while (0)

#define SHK_HIP_S(expr)                                                         \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
            return SHK_ERR_HIP;                                                 \
        }                                                                       \
    } while (0)

}  // namespace

struct shk_chebcode {
    ChebTree *tree = nullptr;
};

namespace {
// Speed presets on the GPU, from measured one-shot times on MI355X (MP
// spectra, eta = 1/sqrt(p), `all_points_dev` / `at_points_dev`): the exact
// all-points transform is faster than ChebCode FAST up to about p = 20 000
// (FP64 0.128 vs 0.130 ms there, FP32-compensated 0.081; 0.015 vs 0.084 ms
// at 1 000), and the exact at-points transform up to about 2e8 pairs (200
// points x 5e4: 0.012 vs 0.125 ms; 2 000 x 5e4: 0.056 vs 0.137 ms; 4 096 x
// 2e5: 0.42 vs 0.16 ms).  Both are also the more accurate.  The
// reference's presets come from its own CPU measurements the same way
// (src/config/pareto_autogen.rs).
constexpr size_t kSpeedExactMaxP = 20000;
constexpr double kSpeedExactMaxPairs = 2e8;

int resolve_auto_at_points(int method, size_t nq, size_t p) {
    if (method == SHK_AUTO || method == SHK_SPEED_AUTO)
        return ((double)nq * (double)p <= kSpeedExactMaxPairs) ? SHK_BLOCKED
                                                                 : SHK_CHEBCODE_FAST;
    return shk_resolve_auto_method(method, 0, p);
}

}  // namespace

extern "C" {

const char *shk_last_error(void) { return last_error(); }

const char *shk_version(void) { return "shrinkers_amd 0.1.0 (gfx950)"; }

shk_status shk_release_cached_memory(void) { return (shk_status)cache_free_all(); }

shk_status shk_method_from_name(const char *name, int *method_out) {
    if (!name || !method_out) return fail(SHK_ERR_BAD_ARG, "NULL argument");
    for (const auto &e : kNames) {
        if (std::strcmp(e.name, name) == 0) {
            *method_out = e.id;
            return SHK_OK;
        }
    }
    return fail(SHK_ERR_BAD_METHOD, std::string("unknown method '") + name + "'");
}

const char *shk_method_name(int method) {
    if (method < 0 || method >= kNumMethods) return nullptr;
    return kCanonical[method];
}

double shk_default_eta(size_t p) { return 0.1 / std::sqrt((double)p); }

shk_status shk_set_symmetric_scratch_limit(size_t bytes) {
    g_sym_budget.store(bytes ? bytes : kSymScratchBudget, std::memory_order_relaxed);
    return SHK_OK;
}

shk_status shk_symmetric_range(int precision, size_t *min_p, size_t *max_p) {
    if (!min_p || !max_p) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    *min_p = precision == SHK_PREC_F64 ? kSym64MinP : kSymMinP;
    *max_p = kSymMaxP;
    return SHK_OK;
}

int shk_resolve_auto_method(int method, int parallel, size_t p) {
    (void)parallel;
    switch (method) {
        case SHK_AUTO:
        case SHK_SPEED_AUTO:
            return (p <= kSpeedExactMaxP) ? SHK_BLOCKED : SHK_CHEBCODE_FAST;
        case SHK_ACCURACY_AUTO:
            return SHK_BLOCKED;
        default:
            return method;
    }
}

shk_status shk_compute_all_stieltjes(const double *eigenvalues, size_t p, double eta, int method,
                                     double cutoff_ratio, int parallel, int precision,
                                     double *out_re, double *out_im) {
    if (p == 0) return SHK_OK;  // mod.rs:611-613
    if (!eigenvalues || !out_re || !out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (!all_finite(eigenvalues, p))
        return fail(SHK_ERR_NONFINITE, "eigenvalues must be finite (found NaN or infinity)");
    SHK_CHECK(check_eta(eta));
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C here");
    method = shk_resolve_auto_method(method, parallel, p);
    SHK_CHECK(check_method(method));
    if (precision == SHK_PREC_F32C && max_abs(eigenvalues, p) > kF32cMaxAbs)
        precision = SHK_PREC_F64;
    SHK_CHECK(ensure_device());
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    DevBuf ev, re, im;
    SHK_CHECK(dev_alloc(ev, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(re, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(im, sizeof(double) * p, s));
    const size_t bp = sizeof(double) * p;
    HostStage hs(s);
    SHK_CHECK(hs.reserve({bp, bp, bp}));
    SHK_CHECK(hs.upload(ev.p, eigenvalues, bp));
    SHK_CHECK(run_allpoints(ev.as<double>(), p, eta, method, cutoff_ratio, precision,
                            1.0 / (double)p, re.as<double>(), im.as<double>(), s));
    SHK_CHECK(hs.download(out_re, re.p, bp));
    SHK_CHECK(hs.download(out_im, im.p, bp));
    SHK_CHECK(hs.finish());
    return SHK_OK;
}

shk_status shk_compute_all_stieltjes_f32(const float *eigenvalues, size_t p, float eta,
                                         int method, double cutoff_ratio, float *out_re,
                                         float *out_im) {
    if (p == 0) return SHK_OK;  // mod.rs:811-814
    if (!eigenvalues || !out_re || !out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    for (size_t i = 0; i < p; ++i)
        if (!std::isfinite(eigenvalues[i]))
            return fail(SHK_ERR_NONFINITE, "eigenvalues must be finite (found NaN or infinity)");
    SHK_CHECK(check_eta((double)eta));
    method = shk_resolve_auto_method(method, 0, p);
    SHK_CHECK(check_method(method));
    SHK_CHECK(ensure_device());
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    DevBuf ev32, ev, re, im, re32, im32;
    SHK_CHECK(dev_alloc(ev32, sizeof(float) * p, s));
    SHK_CHECK(dev_alloc(ev, sizeof(double) * p, s));
    const size_t b32 = sizeof(float) * p, b64 = sizeof(double) * p;
    HostStage hs(s);
    SHK_CHECK(hs.reserve({b32, b64, b64}));
    SHK_CHECK(hs.upload(ev32.p, eigenvalues, b32));
    SHK_CHECK(widen_f32(ev32.as<float>(), ev.as<double>(), p, s));
    const double eta_d = (double)eta;
    const double scale = 1.0 / (double)p;
    const ExactKind kind = allpoints_kind(method, cutoff_ratio);
    float amax = 0.0f;
    for (size_t i = 0; i < p; ++i) amax = std::max(amax, std::fabs(eigenvalues[i]));
    if (family_of(method) == FAM_EXACT && kind == EXACT_PLAIN && f32c_eta_ok(eta_d) &&
        (double)amax <= kF32cMaxAbs) {
        // Blocked/BlockedTiled/Auto arm (mod.rs:817-830): compensated f32.
        SHK_CHECK(dev_alloc(re32, sizeof(float) * p, s));
        SHK_CHECK(dev_alloc(im32, sizeof(float) * p, s));
        SHK_CHECK(exact_f32c(ev.as<double>(), p, ev.as<double>(), p, eta_d, scale, re32.p,
                             im32.p, 1, s));
        SHK_CHECK(hs.download(out_re, re32.p, b32));
        SHK_CHECK(hs.download(out_im, im32.p, b32));
        SHK_CHECK(hs.finish());
        return SHK_OK;
    }
    // Other methods, and eta or magnitudes outside the FP32-compensated
    // range: the f64 path, converted (mod.rs:832-849).
    SHK_CHECK(dev_alloc(re, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(im, sizeof(double) * p, s));
    SHK_CHECK(run_allpoints(ev.as<double>(), p, eta_d, method, cutoff_ratio, SHK_PREC_F64, scale,
                            re.as<double>(), im.as<double>(), s));
    const char *sr = nullptr, *si = nullptr;
    SHK_CHECK(hs.download(nullptr, re.p, b64, &sr));
    SHK_CHECK(hs.download(nullptr, im.p, b64, &si));
    SHK_CHECK(hs.finish());
    const double *hr = reinterpret_cast<const double *>(sr);
    const double *hi = reinterpret_cast<const double *>(si);
    for (size_t i = 0; i < p; ++i) {
        out_re[i] = (float)hr[i];
        out_im[i] = (float)hi[i];
    }
    return SHK_OK;
}

shk_status shk_compute_stieltjes_at_points(const double *query_points, size_t nq,
                                           const double *eigenvalues, size_t p, double eta,
                                           int method, double cutoff_ratio, int parallel,
                                           double *out_re, double *out_im) {
    if (nq == 0) return SHK_OK;
    if (!query_points || !out_re || !out_im || (p && !eigenvalues))
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (p == 0) {  // an empty sum
        for (size_t i = 0; i < nq; ++i) out_re[i] = out_im[i] = 0.0;
        return SHK_OK;
    }
    if (!all_finite(eigenvalues, p) || !all_finite(query_points, nq))
        return fail(SHK_ERR_NONFINITE, "eigenvalues must be finite (found NaN or infinity)");
    SHK_CHECK(check_eta(eta));
    method = resolve_auto_at_points(method, nq, p);
    SHK_CHECK(check_method(method));
    SHK_CHECK(ensure_device());
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    DevBuf q, ev, re, im;
    SHK_CHECK(dev_alloc(q, sizeof(double) * nq, s));
    SHK_CHECK(dev_alloc(ev, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(re, sizeof(double) * nq, s));
    SHK_CHECK(dev_alloc(im, sizeof(double) * nq, s));
    const size_t bq = sizeof(double) * nq, bp = sizeof(double) * p;
    HostStage hs(s);
    SHK_CHECK(hs.reserve({bq, bp, bq, bq}));
    SHK_CHECK(hs.upload(q.p, query_points, bq));
    SHK_CHECK(hs.upload(ev.p, eigenvalues, bp));
    SHK_CHECK(run_atpoints(q.as<double>(), nq, ev.as<double>(), p, eta, method, cutoff_ratio,
                           re.as<double>(), im.as<double>(), s));
    SHK_CHECK(hs.download(out_re, re.p, bq));
    SHK_CHECK(hs.download(out_im, im.p, bq));
    SHK_CHECK(hs.finish());
    return SHK_OK;
}

shk_status shk_compute_all_stieltjes_with_deriv(const double *eigenvalues, size_t p, double eta,
                                                double *val_re, double *val_im, double *der_re,
                                                double *der_im) {
    if (p == 0) return SHK_OK;
    if (!eigenvalues || !val_re || !val_im || !der_re || !der_im)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (!all_finite(eigenvalues, p))
        return fail(SHK_ERR_NONFINITE, "eigenvalues must be finite (found NaN or infinity)");
    SHK_CHECK(check_eta(eta));
    SHK_CHECK(ensure_device());
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    DevBuf ev, o;
    SHK_CHECK(dev_alloc(ev, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(o, sizeof(double) * 4 * p, s));
    const size_t bp = sizeof(double) * p;
    HostStage hs(s);
    SHK_CHECK(hs.reserve({bp, bp, bp, bp, bp}));
    SHK_CHECK(hs.upload(ev.p, eigenvalues, bp));
    double *base = o.as<double>();
    double *outs[4] = {base, base + p, base + 2 * p, base + 3 * p};
    SHK_CHECK(exact_f64(ev.as<double>(), p, ev.as<double>(), p, eta, EXACT_DERIV, 0.0, 1.0, outs,
                        s));
    double *host[4] = {val_re, val_im, der_re, der_im};
    for (int k = 0; k < 4; ++k) SHK_CHECK(hs.download(host[k], outs[k], bp));
    SHK_CHECK(hs.finish());
    return SHK_OK;
}

shk_status shk_compute_stieltjes_scaled_ray(const double *eigenvalues, size_t p, double scale,
                                            double *out_re, double *out_im) {
    if (p == 0) return SHK_OK;
    if (!eigenvalues || !out_re || !out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (!all_finite(eigenvalues, p))
        return fail(SHK_ERR_NONFINITE, "eigenvalues must be finite (found NaN or infinity)");
    if (!std::isfinite(scale)) return fail(SHK_ERR_BAD_ARG, "scale must be finite");
    SHK_CHECK(ensure_device());
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    DevBuf ev, re, im;
    SHK_CHECK(dev_alloc(ev, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(re, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(im, sizeof(double) * p, s));
    const size_t bp = sizeof(double) * p;
    HostStage hs(s);
    SHK_CHECK(hs.reserve({bp, bp, bp}));
    SHK_CHECK(hs.upload(ev.p, eigenvalues, bp));
    double *outs[2] = {re.as<double>(), im.as<double>()};
    SHK_CHECK(exact_f64(ev.as<double>(), p, ev.as<double>(), p, scale, EXACT_RAY, 0.0, 1.0, outs,
                        s));
    SHK_CHECK(hs.download(out_re, re.p, bp));
    SHK_CHECK(hs.download(out_im, im.p, bp));
    SHK_CHECK(hs.finish());
    return SHK_OK;
}

shk_status shk_all_points_dev(const double *d_eigenvalues, size_t p, double eta, int method,
                              double cutoff_ratio, int precision, int scaled, double *d_out_re,
                              double *d_out_im, void *stream) {
    if (p == 0) return SHK_OK;
    if (!d_eigenvalues || !d_out_re || !d_out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    method = shk_resolve_auto_method(method, 0, p);
    SHK_CHECK(check_method(method));
    SHK_CHECK(ensure_device());
    const double scale = scaled ? 1.0 / (double)p : 1.0;
    return (shk_status)run_allpoints(d_eigenvalues, p, eta, method, cutoff_ratio, precision,
                                     scale, d_out_re, d_out_im, (hipStream_t)stream);
}

shk_status shk_all_points_many_dev(const double *d_eigenvalues, const int64_t *offsets,
                                   size_t n_spectra, const double *etas, int method,
                                   int precision, int scaled, double *d_out_re,
                                   double *d_out_im, int n_streams, float *times_ms,
                                   void *stream) {
    if (n_spectra == 0) return SHK_OK;
    if (!d_eigenvalues || !offsets || !etas || !d_out_re || !d_out_im)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (offsets[0] != 0) return fail(SHK_ERR_BAD_ARG, "offsets[0] must be 0");
    if (n_streams < 0 || n_streams > kMaxManyStreams)
        return fail(SHK_ERR_BAD_ARG, "n_streams must be in [0, " +
                                         std::to_string(kMaxManyStreams) + "]");
    for (size_t k = 0; k < n_spectra; ++k) {
        if (offsets[k + 1] < offsets[k]) return fail(SHK_ERR_BAD_ARG, "offsets must not decrease");
        SHK_CHECK(check_eta(etas[k]));
    }
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    if (precision == SHK_PREC_F32C)
        for (size_t k = 0; k < n_spectra; ++k)
            if (!f32c_eta_ok(etas[k])) precision = SHK_PREC_F64;
    {
        // the auto presets by the largest spectrum: the segmented exact launch
        // while every spectrum is in the exact family's fast range
        size_t p_max = 0;
        for (size_t k = 0; k < n_spectra; ++k)
            p_max = std::max(p_max, (size_t)(offsets[k + 1] - offsets[k]));
        method = shk_resolve_auto_method(method, 0, p_max);
    }
    SHK_CHECK(check_method(method));
    SHK_CHECK(ensure_device());
    const bool exact = family_of(method) == FAM_EXACT;
    const ExactKind kind = allpoints_kind(method, 0.0);
    ManyPools &pool = many_pools();
    // Largest spectra first (in the segmented grid and on the streams alike).
    std::vector<size_t> order(n_spectra);
    for (size_t k = 0; k < n_spectra; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return offsets[a + 1] - offsets[a] > offsets[b + 1] - offsets[b];
    });
    hipStream_t caller = (hipStream_t)stream;
    {
        // The spectrum table travels through host staging memory that the
        // next call overwrites, and ChebCode spectra wait on the host once
        // per build: refused under capture, before anything is enqueued.
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        SHK_HIP_S(hipStreamIsCapturing(caller, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(SHK_ERR_BAD_ARG,
                        "the multi-spectrum call cannot be captured in a graph; a prepared batch "
                        "(shk_batch_create, shk_batch_run) replays its launch from a graph");
    }
    if (n_streams == 0 && exact && kind == EXACT_PLAIN) {
        // One segmented launch on the caller's stream: no library streams,
        // no partial sums.
        SHK_CHECK(pool.reserve_table(n_spectra));
        int n_table = 0, n_tiles = 0;
        SHK_CHECK(exact_many_plan(offsets, order.data(), n_spectra, etas, scaled, pool.staging,
                                  &n_table, &n_tiles));
        std::vector<long long> stamps;
        if (n_table > 0) {
            SegSpectrum *d_table = nullptr;
            long long *d_stamps = nullptr;
            SHK_HIP_S(hipMallocAsync((void **)&d_table, sizeof(SegSpectrum) * n_table, caller));
            if (times_ms)
                SHK_HIP_S(hipMallocAsync((void **)&d_stamps, sizeof(long long) * 2 * n_tiles,
                                         caller));
            SHK_HIP_S(hipMemcpyAsync(d_table, pool.staging, sizeof(SegSpectrum) * n_table,
                                     hipMemcpyHostToDevice, caller));
            SHK_HIP_S(hipEventRecord(pool.staged, caller));
            const int st = exact_many(d_eigenvalues, d_table, n_table, n_tiles,
                                      precision == SHK_PREC_F32C, d_out_re, d_out_im, d_stamps,
                                      caller);
            (void)hipFreeAsync(d_table, caller);
            if (!st && times_ms) {
                stamps.resize(2 * (size_t)n_tiles);
                hipError_t ce = hipMemcpyAsync(stamps.data(), d_stamps,
                                               sizeof(long long) * stamps.size(),
                                               hipMemcpyDeviceToHost, caller);
                if (ce == hipSuccess) ce = hipStreamSynchronize(caller);
                (void)hipFreeAsync(d_stamps, caller);
                SHK_HIP_S(ce);
            } else if (d_stamps) {
                (void)hipFreeAsync(d_stamps, caller);
            }
            if (st) return (shk_status)st;
        }
        if (times_ms) {
            // Each spectrum: its first workgroup's start and last workgroup's
            // end, in ms from the earliest workgroup start of the launch.
            int dev = 0, khz = 0;
            SHK_HIP_S(hipGetDevice(&dev));
            SHK_HIP_S(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            if (khz <= 0) return fail(SHK_ERR_HIP, "device reports no wall-clock rate");
            long long t0 = 0;
            if (!stamps.empty()) t0 = *std::min_element(stamps.begin(), stamps.end());
            for (size_t k = 0; k < n_spectra; ++k) times_ms[2 * k] = times_ms[2 * k + 1] = 0.f;
            int e = 0;
            for (size_t idx = 0; idx < n_spectra && e < n_table; ++idx) {
                const size_t k = order[idx];
                if (offsets[k + 1] <= offsets[k]) continue;
                const int a = pool.staging[e].tile0;
                const int b = (e + 1 < n_table) ? pool.staging[e + 1].tile0 : n_tiles;
                long long lo = stamps[2 * (size_t)a], hi = stamps[2 * (size_t)a + 1];
                for (int t = a; t < b; ++t) {
                    lo = std::min(lo, stamps[2 * (size_t)t]);
                    hi = std::max(hi, stamps[2 * (size_t)t + 1]);
                }
                times_ms[2 * k] = (float)((double)(lo - t0) / khz);
                times_ms[2 * k + 1] = (float)((double)(hi - t0) / khz);
                ++e;
            }
        }
        return SHK_OK;
    }
    if (n_streams == 0) n_streams = kDefaultManyStreams;
    SHK_CHECK(pool.reserve(n_streams, times_ms ? 2 * n_spectra + 1 : 1));
    // Each spectrum to the stream with the least p^2 so far; then each stream
    // gets one partial-sum workspace of the largest size its exact-family
    // spectra need (reused in stream order).
    std::vector<double> load(n_streams, 0.0);
    std::vector<int> stream_of(n_spectra, -1);
    std::vector<size_t> ws_need(n_streams, 0);
    for (size_t idx = 0; idx < n_spectra; ++idx) {
        const size_t k = order[idx];
        const size_t p = (size_t)(offsets[k + 1] - offsets[k]);
        if (p == 0) continue;
        const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[j] += (double)p * (double)p;
        stream_of[k] = j;
        if (exact) ws_need[j] = std::max(ws_need[j], exact_workspace_doubles(p, p, kind));
    }
    hipEvent_t t0 = pool.events[0];
    SHK_HIP_S(hipEventRecord(t0, caller));
    std::vector<double *> ws(n_streams, nullptr);
    for (int j = 0; j < n_streams; ++j) {
        SHK_HIP_S(hipStreamWaitEvent(pool.streams[j], t0, 0));
        if (ws_need[j])
            SHK_HIP_S(hipMallocAsync((void **)&ws[j], sizeof(double) * ws_need[j],
                                     pool.streams[j]));
    }
    int st = 0;
    for (size_t idx = 0; idx < n_spectra && !st; ++idx) {
        const size_t k = order[idx];
        const int j = stream_of[k];
        if (j < 0) continue;
        const size_t p = (size_t)(offsets[k + 1] - offsets[k]);
        hipStream_t s = pool.streams[j];
        if (times_ms) SHK_HIP_S(hipEventRecord(pool.events[1 + 2 * k], s));
        const double scale = scaled ? 1.0 / (double)p : 1.0;
        const double *ev = d_eigenvalues + offsets[k];
        double *re = d_out_re + offsets[k], *im = d_out_im + offsets[k];
        if (exact && precision == SHK_PREC_F32C && kind == EXACT_PLAIN) {
            st = exact_f32c(ev, p, ev, p, etas[k], scale, re, im, 0, s, ws[j]);
        } else if (exact) {
            double *outs[2] = {re, im};
            st = exact_f64(ev, p, ev, p, etas[k], kind, 0.0, scale, outs, s, ws[j]);
        } else {
            st = run_allpoints(ev, p, etas[k], method, 0.0, precision, scale, re, im, s);
        }
        if (!st && times_ms) SHK_HIP_S(hipEventRecord(pool.events[2 + 2 * k], s));
    }
    for (int j = 0; j < n_streams; ++j)
        if (ws[j]) (void)hipFreeAsync(ws[j], pool.streams[j]);
    if (st) return (shk_status)st;
    for (int j = 0; j < n_streams; ++j) {
        SHK_HIP_S(hipEventRecord(pool.done[j], pool.streams[j]));
        SHK_HIP_S(hipStreamWaitEvent(caller, pool.done[j], 0));
    }
    if (times_ms) {
        for (int j = 0; j < n_streams; ++j) SHK_HIP_S(hipEventSynchronize(pool.done[j]));
        for (size_t k = 0; k < n_spectra; ++k) {
            float a = 0.f, b = 0.f;
            if (offsets[k + 1] > offsets[k]) {
                SHK_HIP_S(hipEventElapsedTime(&a, t0, pool.events[1 + 2 * k]));
                SHK_HIP_S(hipEventElapsedTime(&b, t0, pool.events[2 + 2 * k]));
            }
            times_ms[2 * k] = a;
            times_ms[2 * k + 1] = b;
        }
    }
    return SHK_OK;
}

shk_status shk_at_points_dev(const double *d_query, size_t nq, const double *d_eigenvalues,
                             size_t p, double eta, int method, double cutoff_ratio,
                             double *d_out_re, double *d_out_im, void *stream) {
    if (nq == 0 || p == 0) return SHK_OK;
    if (!d_query || !d_eigenvalues || !d_out_re || !d_out_im)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    method = resolve_auto_at_points(method, nq, p);
    SHK_CHECK(check_method(method));
    SHK_CHECK(ensure_device());
    return (shk_status)run_atpoints(d_query, nq, d_eigenvalues, p, eta, method, cutoff_ratio,
                                    d_out_re, d_out_im, (hipStream_t)stream);
}

// ---- ChebCode one-shot transforms in a caller's workspace --------------------

namespace {
// The workspace entries take the ChebCode presets only (no auto method: its
// family depends on the size).
shk_status ws_method(int method) {
    if (family_of(method) != FAM_CHEB)
        return fail(SHK_ERR_BAD_METHOD,
                    "the workspace entries take a ChebCode preset (chebcode, chebcode_fast, "
                    "chebcode_xtreme, chebcode_balanced)");
    return SHK_OK;
}
}  // namespace

shk_status shk_chebcode_workspace_bytes(size_t p, int method, size_t *bytes) {
    if (!bytes) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    *bytes = 0;
    if (p == 0) return SHK_OK;
    SHK_CHECK(ws_method(method));
    if (p > (size_t)INT32_MAX / 2) return fail(SHK_ERR_BAD_ARG, "p must be below 2^30");
    *bytes = cheb_ws_bytes(p, preset_of(method).n);
    return SHK_OK;
}

shk_status shk_chebcode_all_points_ws_dev(const double *d_eigenvalues, size_t p, double eta,
                                          int method, int scaled, void *d_workspace,
                                          size_t workspace_bytes, double *d_out_re,
                                          double *d_out_im, void *stream) {
    if (p == 0) return SHK_OK;
    if (!d_eigenvalues || !d_workspace || !d_out_re || !d_out_im)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    SHK_CHECK(ws_method(method));
    SHK_CHECK(ensure_device());
    const double scale = scaled ? 1.0 / (double)p : 1.0;
    return (shk_status)cheb_oneshot_ws(d_eigenvalues, p, preset_of(method), nullptr, p, eta, scale,
                                       1, d_workspace, workspace_bytes, d_out_re, d_out_im,
                                       (hipStream_t)stream);
}

shk_status shk_chebcode_at_points_ws_dev(const double *d_query, size_t nq,
                                         const double *d_eigenvalues, size_t p, double eta,
                                         int method, void *d_workspace, size_t workspace_bytes,
                                         double *d_out_re, double *d_out_im, void *stream) {
    if (nq == 0 || p == 0) return SHK_OK;
    if (!d_query || !d_eigenvalues || !d_workspace || !d_out_re || !d_out_im)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    SHK_CHECK(ws_method(method));
    SHK_CHECK(ensure_device());
    ChebParams prm = preset_of(method);
    prm.leaf_cap = grid_leaf_cap(prm.n, (int64_t)p, (int64_t)nq, prm.leaf_cap);
    return (shk_status)cheb_oneshot_ws(d_eigenvalues, p, prm, d_query, nq, eta, 1.0, 0,
                                       d_workspace, workspace_bytes, d_out_re, d_out_im,
                                       (hipStream_t)stream);
}

shk_status shk_chebcode_workspace_status(const void *d_workspace, void *stream, int *ascending) {
    if (!d_workspace || !ascending) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(ensure_device());
    return (shk_status)cheb_ws_status(d_workspace, (hipStream_t)stream, ascending);
}

shk_status shk_exact_at_points_dev(const double *d_query, size_t nq, const double *d_eigenvalues,
                                   size_t p, double eta, int precision, double *d_out_re,
                                   double *d_out_im, void *stream) {
    if (nq == 0 || p == 0) return SHK_OK;
    if (!d_query || !d_eigenvalues || !d_out_re || !d_out_im)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    SHK_CHECK(ensure_device());
    hipStream_t s = (hipStream_t)stream;
    if (precision == SHK_PREC_F32C && f32c_eta_ok(eta))
        return (shk_status)exact_f32c(d_query, nq, d_eigenvalues, p, eta, 1.0, d_out_re, d_out_im,
                                      0, s);
    double *outs[2] = {d_out_re, d_out_im};
    return (shk_status)exact_f64(d_query, nq, d_eigenvalues, p, eta, EXACT_PLAIN, 0.0, 1.0, outs,
                                 s);
}

shk_status shk_exact_adjoint_dev(const double *d_query, size_t nq, const double *d_eigenvalues,
                                 size_t p, double eta, const double *d_grad_re,
                                 const double *d_grad_im, int precision,
                                 double *d_grad_eigenvalues, double *d_grad_query,
                                 double *d_grad_eta, void *stream) {
    if (!d_grad_eigenvalues && !d_grad_query && !d_grad_eta) return SHK_OK;
    if (nq == 0 || p == 0) {
        // empty sums: every gradient is zero (a fill kernel, not a memset, so
        // the call can be captured in a graph)
        SHK_CHECK(ensure_device());
        hipStream_t s = (hipStream_t)stream;
        if (d_grad_eigenvalues) SHK_CHECK(zero_fill(d_grad_eigenvalues, p, s));
        if (d_grad_query) SHK_CHECK(zero_fill(d_grad_query, nq, s));
        if (d_grad_eta) SHK_CHECK(zero_fill(d_grad_eta, 1, s));
        return SHK_OK;
    }
    if (!d_query || !d_eigenvalues || !d_grad_re || !d_grad_im)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    SHK_CHECK(ensure_device());
    return (shk_status)exact_adjoint(d_query, nq, d_eigenvalues, p, eta, d_grad_re, d_grad_im,
                                     precision == SHK_PREC_F32C && adj_f32c_eta_ok(eta),
                                     d_grad_eigenvalues, d_grad_query, d_grad_eta,
                                     (hipStream_t)stream);
}

shk_status shk_exact_adjoint_all_points_dev(const double *d_eigenvalues, size_t p, double eta,
                                            const double *d_grad_re, const double *d_grad_im,
                                            int precision, double *d_grad_eigenvalues,
                                            double *d_grad_eta, void *stream) {
    if (!d_grad_eigenvalues && !d_grad_eta) return SHK_OK;
    if (p == 0) {
        SHK_CHECK(ensure_device());
        if (d_grad_eta) SHK_CHECK(zero_fill(d_grad_eta, 1, (hipStream_t)stream));
        return SHK_OK;
    }
    if (!d_eigenvalues || !d_grad_re || !d_grad_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    SHK_CHECK(ensure_device());
    hipStream_t s = (hipStream_t)stream;
    // The eigenvalue gradient through the symmetric kernels wherever they run
    // (p <= kSymMaxP): they leave the pair (i, i) out of the gradient sums
    // (its T is 0).  In the two sweeps the diagonal pair adds -G_re/eta^2 to
    // the source sweep and +G_re/eta^2 to the query sweep, each among its
    // neighbours' terms, and the two cancel only to the rounding of 1/eta^2:
    // in f32 runs that swamps the result at small eta, so beyond kSymMaxP the
    // two sweeps run in FP64.  (The eta gradient alone has no such
    // cancellation: its diagonal term -G_im/eta^2 is part of the result.)
    const bool sym = d_grad_eigenvalues && p <= kSymMaxP;
    const bool f64 = precision == SHK_PREC_F64 || !adj_f32c_eta_ok(eta) ||
                     (d_grad_eigenvalues && !sym);
    if (sym)
        return (shk_status)exact_adjoint_sym(d_eigenvalues, p, eta, d_grad_re, d_grad_im, f64,
                                             d_grad_eigenvalues, d_grad_eta, s);
    if (!d_grad_eigenvalues)  // the eta gradient alone: the query sweep
        return (shk_status)exact_adjoint(d_eigenvalues, p, d_eigenvalues, p, eta, d_grad_re,
                                         d_grad_im, !f64, nullptr, nullptr, d_grad_eta, s);
    // two sweeps: the eigenvalue gradient, plus the query gradient
    double *tmp = nullptr;
    CachedBuf cache(CACHE_ADJ_TMP, s);
    if (int e = cache.get(p, &tmp)) return (shk_status)e;
    SHK_CHECK(exact_adjoint(d_eigenvalues, p, d_eigenvalues, p, eta, d_grad_re, d_grad_im, !f64,
                            d_grad_eigenvalues, tmp, d_grad_eta, s));
    return (shk_status)add_into(d_grad_eigenvalues, tmp, p, s);
}

shk_status shk_el_karoui_fit_dev(const double *d_x, const double *d_g_re, const double *d_g_im,
                                 size_t m, double eta, double c, const double *d_atoms,
                                 const double *d_weights, size_t k, double *d_loss,
                                 double *d_grad_g_re, double *d_grad_g_im,
                                 double *d_grad_weights, double *d_grad_eta, void *stream) {
    if (m == 0) return fail(SHK_ERR_EMPTY, "the fit needs at least one grid point");
    if (k == 0) return fail(SHK_ERR_EMPTY, "the fit needs at least one atom");
    if (!d_x || !d_g_re || !d_g_im || !d_atoms || !d_weights)
        return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    if (!std::isfinite(c) || !(c > 0.0))
        return fail(SHK_ERR_BAD_ARG, "c must be a positive finite number, got " + std::to_string(c));
    SHK_CHECK(ensure_device());
    return (shk_status)ek_fit(d_x, d_g_re, d_g_im, m, eta, c, d_atoms, d_weights, k, d_loss,
                              d_grad_g_re, d_grad_g_im, d_grad_weights, d_grad_eta,
                              (hipStream_t)stream);
}

shk_status shk_el_karoui_fit_step_dev(const double *d_eigenvalues, size_t p, const double *d_x,
                                      size_t m, double eta, double c, const double *d_atoms,
                                      const double *d_weights, size_t k, int precision,
                                      double *d_loss, double *d_grad_eigenvalues,
                                      double *d_grad_eta, double *d_grad_weights, void *stream) {
    if (p == 0) return fail(SHK_ERR_EMPTY, "the fit needs at least one eigenvalue");
    if (m == 0) return fail(SHK_ERR_EMPTY, "the fit needs at least one grid point");
    if (k == 0) return fail(SHK_ERR_EMPTY, "the fit needs at least one atom");
    if (!d_eigenvalues || !d_x || !d_atoms || !d_weights) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (p > (size_t)INT32_MAX || m > ((size_t)1 << 24) || k > ((size_t)1 << 24))
        return fail(SHK_ERR_BAD_ARG, "size above the supported range");
    SHK_CHECK(check_eta(eta));
    if (!std::isfinite(c) || !(c > 0.0))
        return fail(SHK_ERR_BAD_ARG, "c must be a positive finite number, got " + std::to_string(c));
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return fail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    SHK_CHECK(ensure_device());
    return (shk_status)ek_fit_step(d_eigenvalues, p, d_x, m, eta, c, d_atoms, d_weights, k,
                                   precision == SHK_PREC_F32C && adj_f32c_eta_ok(eta), d_loss,
                                   d_grad_eigenvalues, d_grad_eta, d_grad_weights,
                                   (hipStream_t)stream);
}

shk_status shk_chebcode_build(const double *eigenvalues, size_t p, double theta, int n,
                              size_t leaf_cap, int f32_far, shk_chebcode **out) {
    if (!out) return fail(SHK_ERR_BAD_ARG, "NULL handle pointer");
    *out = nullptr;
    if (p == 0) return fail(SHK_ERR_EMPTY, "eigenvalues must be non-empty");
    if (!eigenvalues) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (!all_finite(eigenvalues, p))
        return fail(SHK_ERR_NONFINITE, "eigenvalues must be finite (found NaN or infinity)");
    if (n < 2) return fail(SHK_ERR_BAD_ARG, "ChebCode needs at least 2 Chebyshev nodes");
    SHK_CHECK(ensure_device());
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    DevBuf ev;
    SHK_CHECK(dev_alloc(ev, sizeof(double) * p, s));
    // A plain copy: staging it would make the call wait for the whole build.
    SHK_HIP_S(hipMemcpyAsync(ev.p, eigenvalues, sizeof(double) * p, hipMemcpyHostToDevice, s));
    ChebParams prm{theta, n, (int64_t)leaf_cap, f32_far ? 1 : 0};
    ChebTree *t = nullptr;
    SHK_CHECK(cheb_build(ev.as<double>(), p, prm, &t, s));
    *out = new shk_chebcode{t};
    return SHK_OK;
}

shk_status shk_chebcode_build_preset(const double *eigenvalues, size_t p, int method,
                                     shk_chebcode **out) {
    if (family_of(method) != FAM_CHEB)
        return fail(SHK_ERR_BAD_METHOD, "not a chebcode preset: " + method_label(method));
    const ChebParams prm = preset_of(method);
    return shk_chebcode_build(eigenvalues, p, prm.theta, prm.n, (size_t)prm.leaf_cap, prm.f32_far,
                              out);
}

shk_status shk_chebcode_info(const shk_chebcode *h, size_t *node_count, size_t *p) {
    if (!h || !h->tree) return fail(SHK_ERR_BAD_ARG, "NULL handle");
    if (node_count) *node_count = cheb_node_count(h->tree);
    if (p) *p = cheb_p(h->tree);
    return SHK_OK;
}

shk_status shk_chebcode_export(const shk_chebcode *h, double *lo, double *hi, int64_t *lo_idx,
                               int64_t *hi_idx, int32_t *right, double *nodes, double *w) {
    if (!h || !h->tree) return fail(SHK_ERR_BAD_ARG, "NULL handle");
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    return (shk_status)cheb_export(h->tree, lo, hi, lo_idx, hi_idx, right, nodes, w, s);
}

shk_status shk_chebcode_evaluate(const shk_chebcode *h, double eta, double *out_re,
                                 double *out_im) {
    if (!h || !h->tree || !out_re || !out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    const size_t p = cheb_p(h->tree);
    DevBuf re, im;
    SHK_CHECK(dev_alloc(re, sizeof(double) * p, s));
    SHK_CHECK(dev_alloc(im, sizeof(double) * p, s));
    SHK_CHECK(cheb_eval(h->tree, nullptr, p, eta, 1.0, 1, re.as<double>(), im.as<double>(), s));
    HostStage hs(s);
    SHK_CHECK(hs.reserve({sizeof(double) * p, sizeof(double) * p}));
    SHK_CHECK(hs.download(out_re, re.p, sizeof(double) * p));
    SHK_CHECK(hs.download(out_im, im.p, sizeof(double) * p));
    SHK_CHECK(hs.finish());
    return SHK_OK;
}

shk_status shk_chebcode_evaluate_points(const shk_chebcode *h, const double *query, size_t nq,
                                        double eta, double *out_re, double *out_im) {
    if (!h || !h->tree) return fail(SHK_ERR_BAD_ARG, "NULL handle");
    if (nq == 0) return SHK_OK;
    if (!query || !out_re || !out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (!all_finite(query, nq))
        return fail(SHK_ERR_NONFINITE, "query points must be finite (found NaN or infinity)");
    SHK_CHECK(check_eta(eta));
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    DevBuf q, re, im;
    SHK_CHECK(dev_alloc(q, sizeof(double) * nq, s));
    SHK_CHECK(dev_alloc(re, sizeof(double) * nq, s));
    SHK_CHECK(dev_alloc(im, sizeof(double) * nq, s));
    const size_t bq = sizeof(double) * nq;
    HostStage hs(s);
    SHK_CHECK(hs.reserve({bq, bq, bq}));
    SHK_CHECK(hs.upload(q.p, query, bq));
    SHK_CHECK(cheb_eval(h->tree, q.as<double>(), nq, eta, 1.0, 0, re.as<double>(),
                        im.as<double>(), s));
    SHK_CHECK(hs.download(out_re, re.p, bq));
    SHK_CHECK(hs.download(out_im, im.p, bq));
    SHK_CHECK(hs.finish());
    return SHK_OK;
}

shk_status shk_chebcode_evaluate_many(const shk_chebcode *h, const double *etas, size_t n_eta,
                                      double *out_re, double *out_im) {
    if (!h || !h->tree) return fail(SHK_ERR_BAD_ARG, "NULL handle");
    if (n_eta == 0) return SHK_OK;
    if (!etas || !out_re || !out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    for (size_t k = 0; k < n_eta; ++k) SHK_CHECK(check_eta(etas[k]));
    hipStream_t s = host_stream();
    if (!s) return fail(SHK_ERR_HIP, "could not create a HIP stream");
    const size_t p = cheb_p(h->tree);
    DevBuf buf;
    SHK_CHECK(dev_alloc(buf, sizeof(double) * 4 * p, s));
    double *b = buf.as<double>();
    for (size_t k = 0; k < n_eta; k += 2) {
        const double e0 = etas[k];
        const double e1 = (k + 1 < n_eta) ? etas[k + 1] : e0;
        SHK_CHECK(cheb_eval_x2(h->tree, e0, e1, b, b + p, b + 2 * p, b + 3 * p, s));
        const size_t bp = sizeof(double) * p;
        HostStage hs(s);
        SHK_CHECK(hs.reserve({bp, bp, bp, bp}));
        SHK_CHECK(hs.download(out_re + k * p, b, bp));
        SHK_CHECK(hs.download(out_im + k * p, b + p, bp));
        if (k + 1 < n_eta) {
            SHK_CHECK(hs.download(out_re + (k + 1) * p, b + 2 * p, bp));
            SHK_CHECK(hs.download(out_im + (k + 1) * p, b + 3 * p, bp));
        }
        SHK_CHECK(hs.finish());
    }
    return SHK_OK;
}

shk_status shk_chebcode_evaluate_local_dev(const shk_chebcode *h, double eta, double scale,
                                           double *d_out_re, double *d_out_im, void *stream) {
    if (!h || !h->tree || !d_out_re || !d_out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    return (shk_status)cheb_eval_local(h->tree, eta, scale, d_out_re, d_out_im,
                                       (hipStream_t)stream);
}

shk_status shk_chebcode_evaluate_many_dev(const shk_chebcode *h, const double *etas,
                                          size_t n_eta, double *d_out_re, double *d_out_im,
                                          void *stream) {
    if (!h || !h->tree) return fail(SHK_ERR_BAD_ARG, "NULL handle");
    if (n_eta == 0) return SHK_OK;
    if (!etas || !d_out_re || !d_out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    for (size_t k = 0; k < n_eta; ++k) SHK_CHECK(check_eta(etas[k]));
    const size_t p = cheb_p(h->tree);
    hipStream_t s = (hipStream_t)stream;
    // two rows per traversal (chebcode.rs:1312-1361); an odd last eta is
    // paired with itself and both lanes write the same row
    for (size_t k = 0; k < n_eta; k += 2) {
        const bool pair = k + 1 < n_eta;
        double *r0 = d_out_re + k * p, *i0 = d_out_im + k * p;
        double *r1 = pair ? r0 + p : r0, *i1 = pair ? i0 + p : i0;
        SHK_CHECK(cheb_eval_x2(h->tree, etas[k], pair ? etas[k + 1] : etas[k], r0, i0, r1, i1, s));
    }
    return SHK_OK;
}

shk_status shk_chebcode_build_dev(const double *d_eigenvalues, size_t p, double theta, int n,
                                  size_t leaf_cap, int f32_far, void *stream, shk_chebcode **out) {
    if (!out) return fail(SHK_ERR_BAD_ARG, "NULL handle pointer");
    *out = nullptr;
    if (p == 0) return fail(SHK_ERR_EMPTY, "eigenvalues must be non-empty");
    if (!d_eigenvalues) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (n < 2) return fail(SHK_ERR_BAD_ARG, "ChebCode needs at least 2 Chebyshev nodes");
    SHK_CHECK(ensure_device());
    ChebParams prm{theta, n, (int64_t)leaf_cap, f32_far ? 1 : 0};
    ChebTree *t = nullptr;
    SHK_CHECK(cheb_build(d_eigenvalues, p, prm, &t, (hipStream_t)stream));
    *out = new shk_chebcode{t};
    return SHK_OK;
}

shk_status shk_chebcode_evaluate_dev(const shk_chebcode *h, double eta, double scale,
                                     double *d_out_re, double *d_out_im, void *stream) {
    if (!h || !h->tree || !d_out_re || !d_out_im) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    SHK_CHECK(check_eta(eta));
    return (shk_status)cheb_eval(h->tree, nullptr, cheb_p(h->tree), eta, scale, 1, d_out_re,
                                 d_out_im, (hipStream_t)stream);
}

shk_status shk_chebcode_count_work(const shk_chebcode *h, const double *d_query, size_t nq,
                                   double eta, int64_t *near_pairs, int64_t *far_terms,
                                   void *stream) {
    if (!h || !h->tree || !near_pairs || !far_terms) return fail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (d_query == nullptr && nq != 0)
        return fail(SHK_ERR_BAD_ARG, "nq must be 0 when d_query is NULL (the tree's own sources)");
    SHK_CHECK(check_eta(eta));
    return (shk_status)cheb_count(h->tree, d_query, nq, eta, near_pairs, far_terms,
                                  (hipStream_t)stream);
}

shk_status shk_chebcode_destroy(shk_chebcode *h) {
    if (!h) return SHK_OK;
    cheb_destroy(h->tree);
    delete h;
    return SHK_OK;
}

shk_status shk_profile_enable(int enable) {
    g_prof_on.store(enable ? 1 : 0);
    return SHK_OK;
}

shk_status shk_profile_read(double *kernel_ms_total, int64_t *launches) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &ab : g_prof_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(ab.second) == hipSuccess &&
            hipEventElapsedTime(&ms, ab.first, ab.second) == hipSuccess) {
            g_prof_ms += ms;
            g_prof_n += 1;
        }
        (void)hipEventDestroy(ab.first);
        (void)hipEventDestroy(ab.second);
    }
    g_prof_pending.clear();
    if (kernel_ms_total) *kernel_ms_total = g_prof_ms;
    if (launches) *launches = g_prof_n;
    return SHK_OK;
}

shk_status shk_profile_reset(void) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &ab : g_prof_pending) {
        (void)hipEventSynchronize(ab.second);
        (void)hipEventDestroy(ab.first);
        (void)hipEventDestroy(ab.second);
    }
    g_prof_pending.clear();
    g_prof_ms = 0.0;
    g_prof_n = 0;
    return SHK_OK;
}

}  // extern "C"
