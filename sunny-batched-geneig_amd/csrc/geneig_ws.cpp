// geneig_ws.cpp -- per-(device, stream) workspace for the solve kernels.
#include "geneig_ws.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace geneig_host {

namespace {

struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool captured = false;          // the stream was seen under capture
    std::vector<std::pair<void*, size_t>> retired;   // outgrown buffers a captured graph may still use
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

int part_count(int batch, int chunk, int want) {
    if (want < 1) return 1;
    if (want > kMaxParts) want = kMaxParts;
    if (chunk < batch) return 1;
    while (want > 1 && batch / want < kMinPartPairs) --want;
    return want;
}

PlainLayout plain_layout(int batch, unsigned long long pair_bytes, int want_parts) {
    PlainLayout L;
    if (batch <= 0) return L;
    L.chunk = chunk_pairs(batch, pair_bytes);
    L.parts = part_count(batch, L.chunk, want_parts);
    if (L.parts > 1) {
        L.part_slots = (batch + L.parts - 1) / L.parts;
        L.region = region_bytes(L.part_slots, pair_bytes);
        L.bytes = L.region * L.parts;
    } else {
        L.part_slots = L.chunk;
        L.region = region_bytes(L.chunk, pair_bytes);
        L.bytes = L.region;
    }
    return L;
}

MixedLayout mixed_layout(int batch, unsigned long long pair_bytes_f, unsigned long long pair_bytes_d,
                         int want_parts, unsigned long long extra_pair_bytes) {
    MixedLayout L;
    if (batch <= 0) return L;
    const unsigned long long pf = pair_bytes_f + extra_pair_bytes;
    L.list_bytes = round256(sizeof(int) * ((size_t)batch + 1));
    L.chunk_f = chunk_pairs(batch, pf);
    L.chunk_d = chunk_pairs(batch, pair_bytes_d);
    L.rounds = (batch + L.chunk_d - 1) / L.chunk_d;
    L.parts = part_count(batch, L.chunk_f, want_parts);
    L.part_slots = L.parts > 1 ? (batch + L.parts - 1) / L.parts : L.chunk_f;
    L.region_f = region_bytes(L.part_slots, pf);
    L.region_d = region_bytes(L.chunk_d, pair_bytes_d);
    L.bytes = L.list_bytes + std::max(L.region_f * L.parts, L.region_d);
    return L;
}

namespace {

struct Parts {
    hipStream_t s[kMaxParts] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kMaxParts] = {};
    int made = 0;
};

std::mutex g_parts_mu;
std::map<std::pair<int, hipStream_t>, Parts> g_parts;

// Streams and events of `stream` on the current device, at least `parts` of
// them (g_parts_mu held).
hipError_t parts_of(hipStream_t stream, int parts, Parts** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    Parts& p = g_parts[{dev, stream}];
    if (!p.fork) {
        e = hipEventCreateWithFlags(&p.fork, hipEventDisableTiming);
        if (e != hipSuccess) {
            p.fork = nullptr;
            return e;
        }
    }
    while (p.made < parts) {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) return e;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipStreamDestroy(s);
            return e;
        }
        p.s[p.made] = s;
        p.join[p.made] = ev;
        ++p.made;
    }
    *out = &p;
    return hipSuccess;
}

}  // namespace

hipError_t prepare_parts(hipStream_t stream, int parts) {
    if (parts <= 1) return hipSuccess;
    if (parts > kMaxParts) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_parts_mu);
    Parts* p = nullptr;
    return parts_of(stream, parts, &p);
}

hipError_t fork_streams(hipStream_t stream, int parts, hipStream_t* subs) {
    if (parts < 1 || parts > kMaxParts) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_parts_mu);
    Parts* p = nullptr;
    hipError_t e = parts_of(stream, parts, &p);
    if (e != hipSuccess) return e;
    e = hipEventRecord(p->fork, stream);
    if (e != hipSuccess) return e;
    for (int i = 0; i < parts; ++i) {
        e = hipStreamWaitEvent(p->s[i], p->fork, 0);
        if (e != hipSuccess) return e;
        subs[i] = p->s[i];
    }
    return hipSuccess;
}

hipError_t join_streams(hipStream_t stream, int parts, const hipStream_t* subs) {
    if (parts < 1 || parts > kMaxParts) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_parts_mu);
    Parts* p = nullptr;
    hipError_t e = parts_of(stream, parts, &p);
    if (e != hipSuccess) return e;
    for (int i = 0; i < parts; ++i) {
        if (subs[i] != p->s[i]) return hipErrorInvalidValue;
        e = hipEventRecord(p->join[i], p->s[i]);
        if (e != hipSuccess) return e;
        e = hipStreamWaitEvent(stream, p->join[i], 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t workspace(hipStream_t stream, size_t bytes, void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    e = hipStreamIsCapturing(stream, &cs);
    if (e != hipSuccess) return e;
    const bool capturing = (cs == hipStreamCaptureStatusActive);
    std::lock_guard<std::mutex> lk(g_mu);
    Buf& b = g_bufs[{dev, stream}];
    if (capturing) b.captured = true;
    if (b.bytes < bytes) {
        if (capturing) return hipErrorStreamCaptureUnsupported;   // reserve before capturing
        if (b.ptr) {
            if (b.captured) {
                b.retired.push_back({b.ptr, b.bytes});
            } else {
                // the old buffer may still be read by work queued on this stream
                e = hipStreamSynchronize(stream);
                if (e != hipSuccess) return e;
                (void)hipFree(b.ptr);
            }
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

hipError_t release(hipStream_t stream, size_t* freed) {
    if (freed) *freed = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_bufs.find({dev, stream});
    if (it == g_bufs.end()) return hipSuccess;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    Buf& b = it->second;
    if (freed) *freed = b.bytes;
    if (b.ptr) (void)hipFree(b.ptr);
    for (auto& r : b.retired) {
        (void)hipFree(r.first);
        if (freed) *freed += r.second;
    }
    g_bufs.erase(it);
    return hipSuccess;
}

}  // namespace geneig_host
