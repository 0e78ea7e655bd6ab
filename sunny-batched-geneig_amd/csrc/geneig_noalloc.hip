// geneig_noalloc.hip -- the batched FP64 n = 54 solve for callers that cannot
// allocate (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// geneig_full_n54_batched_launch keeps its stage workspace per (device,
// stream) and allocates it on first use.  This entry runs the same kernels
// with the workspace in a file-scope device array, `scratch`, of
// kScratchBytes: it calls no hipMalloc / hipFree and creates no stream or
// event.  Every process that loads the library holds the array on each GPU
// it launches on, until the process ends.  All calls on one GPU share that
// GPU's array, so calls that overlap on one GPU (different streams, or
// threads) race; calls ordered on one stream do not.  The stages clear the
// array's control words with kernel stores before they are read, so nothing
// a previous call left there is used.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "geneig_err.hpp"
#include "geneig_ws.hpp"
#include "../../include/geneig.h"
#include "../../include/geneig_solver.h"

namespace geneig_host {
// geneig_fixed.hip: the batched FP64 n = 54 solve into a caller-provided
// workspace (one region, the caller's stream only), and its size.
hipError_t full_n54_into(const void* H, const void* S, double* W, void* U, int* info, int batch,
                         hipStream_t stream, unsigned char* ws, size_t bytes);
size_t full_n54_into_bytes(int batch);
}  // namespace geneig_host

namespace {

// 512 MiB: the workspace of 4,485 pairs at n = 54 (about 120 KB each), under
// 0.2 % of an MI355X's HBM.
constexpr size_t kScratchBytes = 512ull << 20;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::map<int, unsigned char*> g_ptr;   // device -> device address of its copy of `scratch`

// The current device's address of `scratch`, looked up once per device with
// that device current.  Host code passes only this address to kernels: the
// array's own name is a host-side stand-in, not a device address.
hipError_t scratch_ptr(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ptr.find(dev);
    if (it == g_ptr.end()) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        it = g_ptr.emplace(dev, static_cast<unsigned char*>(p)).first;
    }
    *out = it->second;
    return hipSuccess;
}

int max_batch() {
    static const int m = [] {
        int lo = 0, hi = (int)(kScratchBytes / 1024);   // pairs take far more than 1 KB each
        while (lo < hi) {                               // largest batch whose workspace fits
            const int mid = lo + (hi - lo + 1) / 2;
            if (geneig_host::full_n54_into_bytes(mid) <= kScratchBytes) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }();
    return m;
}

}  // namespace

extern "C" {

int geneig_full_n54_batched_noalloc_max_batch(void) { return max_batch(); }

int geneig_full_n54_batched_launch_noalloc(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                                           hipDoubleComplex* d_U, int* d_info, int batch_size, hipStream_t stream) {
    if (batch_size <= 0) return hipSuccess;
    if (batch_size > max_batch()) {
        geneig_host::fail(GENEIG_E_ARG,
                          "geneig_full_n54_batched_launch_noalloc: batch %d needs %zu B of workspace; the "
                          "library's fixed array holds %zu B (at most %d pairs); split the batch",
                          batch_size, geneig_host::full_n54_into_bytes(batch_size), kScratchBytes, max_batch());
        return hipErrorInvalidValue;
    }
    unsigned char* ws = nullptr;
    hipError_t e = scratch_ptr(&ws);
    if (e != hipSuccess) return e;
    return geneig_host::full_n54_into(d_H, d_S, d_W, d_U, d_info, batch_size, stream, ws, kScratchBytes);
}

}  // extern "C"
