// geneig_multi.cpp -- mixed-size batches: one runtime-n solver per size, all
// sizes solved by one graph launch with one branch per size.
//
// Capture: the caller-independent capture stream records a fork event, every
// size's branch stream waits on it, runs its solver's chunked K1 -> K2 -> K3
// launches and records a join event the capture stream waits on.  Branch
// workspaces are reserved before capture, so nothing is allocated inside it
// and the graph holds kernel nodes only (no memset nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "../../include/geneig_multi.h"
#include "geneig_err.hpp"

using geneig_host::fail;

struct geneig_multi {
    int device = 0;
    std::vector<int> sizes;
    std::vector<geneig_solver_t> solvers;
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> join;
    hipStream_t cap = nullptr;
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // arguments the current graph was captured with
    std::vector<int> counts;
    std::vector<const void*> args;
    long long captures = 0, replays = 0;
    hipStream_t last_stream = nullptr;   // where the current graph last ran

    void drop_graph() {
        // a replay may still be running on the stream it was launched on
        if (exec && replays) (void)hipStreamSynchronize(last_stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~geneig_multi() {
        drop_graph();
        for (auto s : solvers) geneig_solver_destroy(s);
        for (auto s : branch)
            if (s) (void)hipStreamDestroy(s);
        for (auto e : join)
            if (e) (void)hipEventDestroy(e);
        if (cap) (void)hipStreamDestroy(cap);
        if (fork) (void)hipEventDestroy(fork);
    }
};

namespace {

int check_args(geneig_multi_t m, const int* counts, const void* const* A, const void* const* B, double* const* W,
               void* const* U, int* const* info) {
    if (!m) return fail(GENEIG_E_ARG, "null multi-size handle");
    if (!counts || !A || !B || !W || !U || !info) return fail(GENEIG_E_ARG, "null argument array");
    for (size_t k = 0; k < m->sizes.size(); ++k) {
        if (counts[k] < 0) return fail(GENEIG_E_ARG, "bin %zu: negative count %d", k, counts[k]);
        if (counts[k] > 0 && (!A[k] || !B[k] || !W[k] || !U[k] || !info[k]))
            return fail(GENEIG_E_ARG, "bin %zu (n=%d): null buffer", k, m->sizes[k]);
    }
    return GENEIG_E_OK;
}

std::vector<const void*> pack(size_t K, const void* const* A, const void* const* B, double* const* W,
                              void* const* U, int* const* info) {
    std::vector<const void*> v;
    v.reserve(5 * K);
    for (size_t k = 0; k < K; ++k) {
        v.push_back(A[k]);
        v.push_back(B[k]);
        v.push_back(W[k]);
        v.push_back(U[k]);
        v.push_back(info[k]);
    }
    return v;
}

int capture(geneig_multi_t m, const int* counts, const void* const* A, const void* const* B, double* const* W,
            void* const* U, int* const* info) {
    m->drop_graph();
    const size_t K = m->sizes.size();
    for (size_t k = 0; k < K; ++k) {
        const int rc = geneig_solver_reserve(m->solvers[k], counts[k], m->branch[k]);
        if (rc) return rc;
    }
    // Branches in order of decreasing work (n^3 per pair times the count): the
    // graph's nodes are submitted in capture order, so the bin that takes
    // longest claims the CUs first instead of trailing the small bins.
    std::vector<size_t> order(K);
    for (size_t k = 0; k < K; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const double na = m->sizes[a], nb = m->sizes[b];
        return na * na * na * counts[a] > nb * nb * nb * counts[b];
    });
    GENEIG_HIP_TRY(hipStreamBeginCapture(m->cap, hipStreamCaptureModeThreadLocal));
    int rc = GENEIG_E_OK;
    hipError_t e = hipEventRecord(m->fork, m->cap);
    for (size_t i = 0; i < K && e == hipSuccess && rc == GENEIG_E_OK; ++i) {
        const size_t k = order[i];
        if (counts[k] == 0) continue;
        e = hipStreamWaitEvent(m->branch[k], m->fork, 0);
        if (e != hipSuccess) break;
        rc = geneig_solver_launch(m->solvers[k], A[k], B[k], W[k], U[k], info[k], counts[k], m->branch[k]);
        if (rc) break;
        e = hipEventRecord(m->join[k], m->branch[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(m->cap, m->join[k], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(m->cap, &g);
    if (rc != GENEIG_E_OK || e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        return fail(GENEIG_E_HIP, "graph capture failed: %s",
                    hipGetErrorString(e != hipSuccess ? e : e2));
    }
    m->graph = g;
    GENEIG_HIP_TRY(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    m->counts.assign(counts, counts + K);
    m->args = pack(K, A, B, W, U, info);
    ++m->captures;
    return GENEIG_E_OK;
}

}  // namespace

extern "C" {

int geneig_multi_create(geneig_multi_t* out, int num_sizes, const int* sizes, int device_id, int tuning_mode,
                        int cache_mode) {
    if (!out) return fail(GENEIG_E_ARG, "null output handle");
    *out = nullptr;
    if (num_sizes <= 0 || !sizes) return fail(GENEIG_E_ARG, "need at least one size");
    std::unique_ptr<geneig_multi> m(new geneig_multi);
    m->device = device_id;
    GENEIG_HIP_TRY(hipSetDevice(device_id));
    for (int k = 0; k < num_sizes; ++k) {
        geneig_solver_t s = nullptr;
        const int rc = geneig_solver_create(&s, sizes[k], device_id, 0, 0, tuning_mode, cache_mode);
        if (rc) return rc;
        m->solvers.push_back(s);
        m->sizes.push_back(sizes[k]);
        hipStream_t st = nullptr;
        GENEIG_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        m->branch.push_back(st);
        hipEvent_t ev = nullptr;
        GENEIG_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        m->join.push_back(ev);
    }
    GENEIG_HIP_TRY(hipStreamCreateWithFlags(&m->cap, hipStreamNonBlocking));
    GENEIG_HIP_TRY(hipEventCreateWithFlags(&m->fork, hipEventDisableTiming));
    *out = m.release();
    return GENEIG_E_OK;
}

void geneig_multi_destroy(geneig_multi_t m) { delete m; }

int geneig_multi_launch(geneig_multi_t m, const int* counts, const void* const* d_A, const void* const* d_B,
                        double* const* d_W, void* const* d_U, int* const* d_info, hipStream_t stream) {
    int rc = check_args(m, counts, d_A, d_B, d_W, d_U, d_info);
    if (rc) return rc;
    const size_t K = m->sizes.size();
    const bool same = m->exec && std::vector<int>(counts, counts + K) == m->counts &&
                      pack(K, d_A, d_B, d_W, d_U, d_info) == m->args;
    if (!same) {
        rc = capture(m, counts, d_A, d_B, d_W, d_U, d_info);
        if (rc) return rc;
    }
    GENEIG_HIP_TRY(hipGraphLaunch(m->exec, stream));
    m->last_stream = stream;
    ++m->replays;
    return GENEIG_E_OK;
}

int geneig_multi_launch_serial(geneig_multi_t m, const int* counts, const void* const* d_A,
                               const void* const* d_B, double* const* d_W, void* const* d_U, int* const* d_info,
                               hipStream_t stream) {
    int rc = check_args(m, counts, d_A, d_B, d_W, d_U, d_info);
    if (rc) return rc;
    for (size_t k = 0; k < m->sizes.size(); ++k) {
        if (counts[k] == 0) continue;
        rc = geneig_solver_launch(m->solvers[k], d_A[k], d_B[k], d_W[k], d_U[k], d_info[k], counts[k], stream);
        if (rc) return rc;
    }
    return GENEIG_E_OK;
}

int geneig_multi_num_sizes(geneig_multi_t m) { return m ? (int)m->sizes.size() : 0; }
geneig_solver_t geneig_multi_solver(geneig_multi_t m, int k) {
    return (m && k >= 0 && k < (int)m->solvers.size()) ? m->solvers[k] : nullptr;
}
long long geneig_multi_captures(geneig_multi_t m) { return m ? m->captures : 0; }
long long geneig_multi_replays(geneig_multi_t m) { return m ? m->replays : 0; }

}  // extern "C"
