// fl_eval_noalloc: the kernel sum of fl_eval for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  The source-chunk partial sums that fl_eval keeps in
// a caller workspace live here in one fixed device array per GPU, `scratch`,
// of NOALLOC_SCRATCH_BYTES.  No allocation happens on this path: no hipMalloc,
// hipFree or allocating library call.
//
// The array's host-side name is not a device address, so each device's copy
// is resolved once, with that device current, by hipGetSymbolAddress, and only
// that pointer is handed to the kernels.  Every element of the scratch that a
// call reads is written earlier in the same call (each source chunk writes its
// own partial sums), so nothing needs clearing between calls.  All calls on
// one GPU share its array: calls that overlap on one GPU (different streams,
// or different host threads) race on it, and the caller must order them.
#include "fl_internal.h"

#include <mutex>

namespace fl {
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = FL_EVAL_NOALLOC_SCRATCH_BYTES;   // per GPU, held until the process ends
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// This device's copy of `scratch` (resolved on the first call per device).
int scratch_for_current_device(void** out)
{
    int dev = 0;
    FL_HIP(hipGetDevice(&dev));
    FL_CHECK(dev >= 0 && dev < MAX_DEVICES, "device index %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        FL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(FL_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FL_OK;
}

}  // namespace
}  // namespace fl

extern "C" {

int fl_eval_noalloc(const fl_kernel* kernel, int64_t M, int64_t N, const double* tgt, const double* src,
                    const double* eps, int64_t eps_stride, const double* f_re, const double* f_im, double mu_re,
                    double mu_im, double* u_re, double* u_im, int precision, void* stream)
{
    fl::clear_error();
    FL_CHECK(M >= 0 && N >= 0, "negative size");
    const int nf = f_im ? 2 : 1;
    const size_t need = fl::eval_workspace_size(M, N, nf, precision);
    if (need > fl::NOALLOC_SCRATCH_BYTES)
        return fl::set_error(FL_EWORKSPACE,
                             "fl_eval_noalloc: this sum (M = %lld, N = %lld) needs %zu bytes of scratch, above the "
                             "library's fixed %zu; call fl_eval with a workspace",
                             (long long)M, (long long)N, need, fl::NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    if (int e = fl::scratch_for_current_device(&ws)) return e;
    return fl_eval(kernel, M, N, tgt, src, eps, eps_stride, f_re, f_im, mu_re, mu_im, u_re, u_im, precision, ws,
                   fl::NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"
