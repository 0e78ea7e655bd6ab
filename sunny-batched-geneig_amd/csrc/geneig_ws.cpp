// geneig_ws.cpp -- per-(device, stream) workspace for the solve kernels.
#include "geneig_ws.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <map>
#include <mutex>
#include <utility>

namespace geneig_host {

namespace {

struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, Buf> g_bufs;

size_t budget_bytes() {
    static const size_t b = [] {
        const char* e = std::getenv("GENEIG_WORKSPACE_MB");
        // 16 GiB by default: the QL stage runs 64 pairs per wave as serial
        // chains, so it needs as many pairs per dispatch as possible to hide
        // their latency (131,072 FP64 pairs at n = 54 fit one chunk), and
        // 16 GiB is under 6% of an MI355X's HBM.
        long mb = e ? std::atol(e) : 16384;
        if (mb < 16) mb = 16;
        return (size_t)mb << 20;
    }();
    return b;
}

}  // namespace

int chunk_pairs(int batch, unsigned long long pair_bytes) {
    if (batch <= 0) return 0;
    const unsigned long long c = std::max<unsigned long long>(1, budget_bytes() / pair_bytes);
    return (int)std::min<unsigned long long>((unsigned long long)batch, c);
}

int ql_log_entries() {
    const char* e = std::getenv("GENEIG_QL_LOG_ENTRIES");
    if (!e || !*e) return INT_MAX;
    const long v = std::atol(e);
    return v > 0 && v < INT_MAX ? (int)v : INT_MAX;
}

hipError_t workspace(hipStream_t stream, size_t bytes, void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_mu);
    Buf& b = g_bufs[{dev, stream}];
    if (b.bytes < bytes) {
        if (b.ptr) {
            // the old buffer may still be read by work queued on this stream
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return e;
            (void)hipFree(b.ptr);
            b.ptr = nullptr;
            b.bytes = 0;
        }
        e = hipMalloc(&b.ptr, bytes);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            return e;
        }
        b.bytes = bytes;
    }
    *out = b.ptr;
    return hipSuccess;
}

}  // namespace geneig_host
