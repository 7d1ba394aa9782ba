// All-points exact transform that allocates nothing: the symmetric kernels
// (each unordered pair once, shk_sym.hip) run band by band on a file-scope
// device array instead of a cached or pool buffer, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).
//
// The array is `scratch`, kNoAllocScratchBytes (64 MiB) per device, held from
// library load until the process ends.  In host code its name is only a
// host-side stand-in: its device address is looked up once per device with
// hipGetSymbolAddress (the first call on that device) and that pointer is
// what the kernels receive.  Every call on a device shares the array, and a
// call leaves its contents behind; nothing in it needs clearing (every
// partial slot and running sum a call reads was written earlier in that
// call).  Calls that overlap on one device race on the array.
#include "shk_internal.hpp"
#include "shrinkers_stieltjes.h"

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

namespace shk {
namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr size_t kNoAllocScratchDoubles = kNoAllocScratchBytes / sizeof(double);
// Largest p: 64 MiB holds bands of 7 rows of tiles at this size (74 bands).
constexpr size_t kNoAllocMaxP = (size_t)1 << 18;

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

std::mutex g_mu;
std::vector<double *> g_dev_ptr;  // the array's device address, per device id

// The array's address on the current device (looked up on first use).
int scratch_on_device(double **out) {
    int dev = 0;
    SHK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    if ((size_t)dev >= g_dev_ptr.size()) g_dev_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_dev_ptr[(size_t)dev]) {
        void *p = nullptr;
        SHK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("no-allocation entry: the scratch array has no device address");
            return 7;
        }
        g_dev_ptr[(size_t)dev] = static_cast<double *>(p);
    }
    *out = g_dev_ptr[(size_t)dev];
    return 0;
}

shk_status nfail(shk_status s, const std::string &msg) {
    set_error(msg);
    return s;
}

}  // namespace
}  // namespace shk

extern "C" shk_status shk_all_points_noalloc_dev(const double *d_eigenvalues, size_t p, double eta,
                                                 int precision, int scaled, double *d_out_re,
                                                 double *d_out_im, void *stream) {
    using namespace shk;
    if (p == 0) return SHK_OK;
    if (!d_eigenvalues || !d_out_re || !d_out_im) return nfail(SHK_ERR_BAD_ARG, "NULL pointer");
    if (!(std::isfinite(eta) && eta > 0.0))
        return nfail(SHK_ERR_BAD_ETA,
                     "eta must be a positive finite number, got " + std::to_string(eta));
    if (precision != SHK_PREC_F64 && precision != SHK_PREC_F32C)
        return nfail(SHK_ERR_BAD_ARG, "precision must be SHK_PREC_F64 or SHK_PREC_F32C");
    if (p > kNoAllocMaxP)
        return nfail(SHK_ERR_BAD_ARG, "no-allocation entry: p = " + std::to_string(p) +
                                          " exceeds its maximum of " +
                                          std::to_string(kNoAllocMaxP));
    if (const int dv = device_ready()) return (shk_status)dv;
    double *buf = nullptr;
    if (const int st = scratch_on_device(&buf)) return (shk_status)st;
    const int f64 = precision == SHK_PREC_F64 || !f32c_eta_ok(eta);
    const double scale = scaled ? 1.0 / (double)p : 1.0;
    return (shk_status)exact_sym_on(d_eigenvalues, p, eta, scale, f64, d_out_re, d_out_im, buf,
                                    kNoAllocScratchDoubles, (hipStream_t)stream);
}
