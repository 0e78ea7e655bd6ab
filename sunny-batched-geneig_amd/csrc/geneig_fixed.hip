// geneig_fixed.hip -- static n = 54 entry points (hipcc, gfx950).
//
// MI355X counterpart of the reference's src/geneig_fixed.cu: the same six
// extern "C" launchers (chol, reduce, reduce+eig, full, batched full, finish)
// and their query functions, with the pipeline of geneig_device.inc behind
// them.  Every launcher returns a hipError_t value (0 = success) where the
// reference returned void; callers that ignore the result keep working.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <set>

#include "geneig_device.inc"
#include "geneig_real.inc"
#include "geneig_ws.hpp"
#include "../../include/geneig.h"

namespace {

constexpr int kN = 54;
using C54 = geneig::Cfg<kN>;
using W54 = geneig::Ws<kN>;
// K1/K3 pack this many independent pairs (one wave each) into a workgroup; 1
// keeps the LDS allocation per workgroup at one pair so three workgroups fit
// a CU.
constexpr int kBPB = 1;
// Concurrent parts of a batched FP64 solve: the one-pair-per-lane QL and the
// memory-heavy K3 of one part overlap the others' K1.  Three measured fastest
// at 131,072 pairs (64.0 ms against 64.7 on one stream, 65.0 with two, 64.5
// with four, where five streams share the four hardware queues).
#ifndef GENEIG_PLAIN_PARTS
#define GENEIG_PLAIN_PARTS 3
#endif
constexpr int kPlainParts = GENEIG_PLAIN_PARTS;

template <int MODE>
__global__ void __launch_bounds__(64 * kBPB) k1_kernel(const geneig::cplx* __restrict__ A,
                                                        const geneig::cplx* __restrict__ B,
                                                        double* __restrict__ W, geneig::cplx* __restrict__ U,
                                                        int* __restrict__ info, unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ geneig::cplx vecs[kBPB * C54::VEC];
    geneig::k1_batched<kN, kBPB, MODE>(A, B, W, U, info, ws, batch, batch, smem, vecs);
}

// The FP64 Cholesky alone (two waves per SIMD); the batched FP64 solve runs it
// ahead of K1, which then starts from the stashed factor (MODE_MIXED).
__global__ void __launch_bounds__(64, GENEIG_CHOL64_WAVES_PER_EU) chol64_kernel(const geneig::cplx* __restrict__ B,
                                                                           double* __restrict__ W,
                                                                           geneig::cplx* __restrict__ U,
                                                                           int* __restrict__ info, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ geneig::cplx vecs[C54::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_CHOL64, double>(B, B, W, U, info, nullptr, batch, batch, smem, vecs);
}

// The Cholesky stage of the in-place solve: B factored into B (lower
// triangle), NaN into the eigenvector output A when B is not positive definite.
__global__ void __launch_bounds__(64, GENEIG_CHOL64_WAVES_PER_EU) chol_ip_kernel(const geneig::cplx* __restrict__ B,
                                                                            double* __restrict__ W,
                                                                            geneig::cplx* __restrict__ A,
                                                                            int* __restrict__ info, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ geneig::cplx vecs[C54::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_CHOL_IP, double>(B, B, W, A, info, nullptr, batch, batch, smem, vecs);
}

// Fused spin-wave solve: the Cholesky stage builds its pair's H(q) + diag I
// column per lane from the bond table (spinwave_column) instead of reading B,
// and stashes L in U; A = eta is built inside K1 (MODE_SW).
struct SpinwaveModel {
    int nbonds;
    const int* shift;      // nbonds
    const int* r;          // nbonds x 3
    const double* J;       // nbonds x 27
    const double* q;       // batch x 3
    const double* diag;    // batch, or null
};

__global__ void __launch_bounds__(64, 1) chol_sw_kernel(SpinwaveModel m, double* __restrict__ W,
                                                                            geneig::cplx* __restrict__ U,
                                                                            int* __restrict__ info, int off,
                                                                            int batch) {
    __shared__ geneig::cplx colv[C54::VEC];
    const int lane = threadIdx.x;
    const long long p = (long long)off + blockIdx.x;
    if (blockIdx.x >= (unsigned)batch) return;
    const size_t mat = (size_t)kN * kN;
    geneig::cplx a[kN];
    const double d =
        geneig::spinwave_column<kN>(a, lane, m.nbonds, m.shift, m.r, m.J, m.q + 3 * p, m.diag ? m.diag[p] : 0.0);
    const int bad = geneig::chol_core<kN, 4, false>(a, d, U + mat * p, colv, lane);
    if (bad) {
        if (lane == 0) info[p] = kN + bad;
        geneig::write_nan<kN>(W + (size_t)kN * p, U + mat * p, lane, true);
        return;
    }
    if (lane == 0) info[p] = 0;
}

// H(q) + diag I as the fused solve builds it, written out (for checking the
// builder against an independent construction).
__global__ void __launch_bounds__(64) sw_build_kernel(SpinwaveModel m, geneig::cplx* __restrict__ H, int batch) {
    const int lane = threadIdx.x;
    const long long p = blockIdx.x;
    if (p >= batch) return;
    geneig::cplx a[kN];
    geneig::spinwave_column<kN>(a, lane, m.nbonds, m.shift, m.r, m.J, m.q + 3 * p, m.diag ? m.diag[p] : 0.0);
    if (lane < kN) {
        geneig::cplx* col = H + (size_t)kN * kN * p + (size_t)kN * lane;
#pragma unroll
        for (int i = 0; i < kN; ++i) col[i] = a[i];
    }
}

// K2 in three kernels: the QL chases of 64 pairs per wave with the rotations
// logged (K2a), their replay into Z (K2b), and the one-pair-per-workgroup QL
// for the rare pair whose log overflowed (K2x).
constexpr unsigned kK2aLds = geneig::k2a_lds_bytes<kN, double>();
constexpr unsigned kK3Lds = geneig::k3_lds_bytes<kN, double>();
constexpr unsigned kK1Lds = geneig::k1_lds_bytes<kN, double>();
// dynamic LDS of the larger of K1 and K3 plus the static broadcast vectors
constexpr unsigned kWorkLds = kK1Lds > kK3Lds ? kK1Lds : kK3Lds;
constexpr unsigned kVecLds = sizeof(geneig::cplx) * C54::VEC;

__global__ void __launch_bounds__(64) k2a_kernel(int* __restrict__ info, unsigned char* ws, int batch, int cap,
                                                  int* __restrict__ ovf) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::k2a_batched<kN, double>(info, ws, batch, batch, reinterpret_cast<double*>(smem), cap, ovf);
}

__global__ void __launch_bounds__(64 * C54::R, GENEIG_K2B_WAVES_PER_EU) k2b_kernel(const int* __restrict__ info,
                                                                                 unsigned char* ws, int batch) {
    __shared__ geneig::cplx buf[geneig::k2b_buf_entries<kN, double>()];
    geneig::k2b_batched<kN, double>(info, ws, batch, batch, buf);
}

__global__ void __launch_bounds__(64 * C54::R, GENEIG_K2_WAVES_PER_EU) k2x_kernel(int* __restrict__ info, unsigned char* ws,
                                                                                int chunk, const int* __restrict__ ovf) {
    __shared__ geneig::cplx rotlog[C54::R * C54::K2_VEC];
    geneig::k2x_batched<kN>(info, ws, chunk, rotlog, ovf);
}

// Clears n ints with kernel stores (a captured launch then holds no memset node).
__global__ void zero_kernel(int* p, int n) { geneig::zero_ints(p, n); }

template <int MODE>
__global__ void __launch_bounds__(64 * kBPB, GENEIG_K3_WAVES_PER_EU) k3_kernel(const geneig::cplx* __restrict__ B,
                                                        double* __restrict__ W, geneig::cplx* __restrict__ U,
                                                        const int* __restrict__ info, unsigned char* ws,
                                                        int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ geneig::cplx vecs[kBPB * C54::VEC];
    geneig::k3_batched<kN, kBPB, MODE>(B, W, U, info, ws, batch, batch, smem, vecs);
}

// The spin-wave pencil in real arithmetic (geneig_real.inc): K1r and K3r
// around the FP64 QL stages.
#ifndef GENEIG_K1R_WAVES_PER_EU
#define GENEIG_K1R_WAVES_PER_EU 2
#endif
#ifndef GENEIG_K3R_WAVES_PER_EU
#define GENEIG_K3R_WAVES_PER_EU 2
#endif
using RC54 = geneig::RealCfg<kN>;

__global__ void __launch_bounds__(64, GENEIG_K1R_WAVES_PER_EU) k1r_kernel(const geneig::cplx* __restrict__ S,
                                                                          double* __restrict__ W,
                                                                          geneig::cplx* __restrict__ U,
                                                                          int* __restrict__ info, unsigned char* ws,
                                                                          int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ double vecs[2 * RC54::VEC];
    const long long pair = blockIdx.x;
    if (pair >= batch) return;
    const size_t mat = (size_t)kN * kN;
    geneig::k1r_pair<kN>(S + mat * pair, W + (size_t)kN * pair, U + mat * pair, info + pair, W54(ws, pair, batch),
                         reinterpret_cast<double*>(smem), vecs, vecs + RC54::VEC, (int)threadIdx.x);
}

__global__ void __launch_bounds__(64, GENEIG_K3R_WAVES_PER_EU) k3r_kernel(double* __restrict__ W,
                                                                          geneig::cplx* __restrict__ U,
                                                                          const int* __restrict__ info,
                                                                          unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ double ev[64];
    __shared__ int perm[64];
    const long long pair = blockIdx.x;
    if (pair >= batch) return;
    const size_t mat = (size_t)kN * kN;
    geneig::k3r_pair<kN>(W + (size_t)kN * pair, U + mat * pair, info + pair, W54(ws, pair, batch),
                         reinterpret_cast<double*>(smem), ev, perm, (int)threadIdx.x);
}

// hipFuncSetAttribute once per kernel per process, thread-safe (the reference
// used an unguarded static bool per launcher, src/geneig_fixed.cu:138-144).
std::mutex g_attr_mu;
std::set<const void*> g_attr_done;

hipError_t set_lds_once(const void* kernel, unsigned bytes) {
    std::lock_guard<std::mutex> lk(g_attr_mu);
    if (g_attr_done.count(kernel)) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) g_attr_done.insert(kernel);
    return e;
}

#define RETURN_IF(expr)                      \
    do {                                     \
        hipError_t _e = (expr);              \
        if (_e != hipSuccess) return _e;     \
    } while (0)

// Whether a mode runs the eigen stages (K2, K3) after K1.
template <int MODE>
constexpr bool eig_mode() {
    return MODE == geneig::MODE_FULL || MODE == geneig::MODE_REDUCE_EIG || MODE == geneig::MODE_FINISH ||
           MODE == geneig::MODE_IP;
}
// The batched FP64 solves (MODE_FULL; MODE_IP, in place) run the Cholesky as
// its own kernel (two waves per SIMD), then K1 from the factor.
template <int MODE>
constexpr bool split_chol() { return MODE == geneig::MODE_FULL || MODE == geneig::MODE_IP; }
template <int MODE>
constexpr int k1_mode() { return MODE == geneig::MODE_FULL ? geneig::MODE_MIXED : MODE; }

// Workspace layout of a launch of `batch` pairs (the batched FP64 solve runs
// kPlainParts concurrent parts when the batch is one chunk).
template <int MODE>
geneig_host::PlainLayout layout_of(int batch) {
    return geneig_host::plain_layout(batch, W54::PAIR_BYTES, split_chol<MODE>() ? kPlainParts : 1);
}

template <int MODE>
hipError_t set_attrs() {
    auto k1 = k1_kernel<k1_mode<MODE>()>;
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1), kK1Lds * kBPB));
    if (split_chol<MODE>()) {
        auto ch = MODE == geneig::MODE_IP ? chol_ip_kernel : chol64_kernel;
        RETURN_IF(set_lds_once(reinterpret_cast<const void*>(ch), geneig::chol64_lds_bytes<kN>()));
    }
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3_kernel<MODE>), kK3Lds * kBPB));
    return set_lds_once(reinterpret_cast<const void*>(k2a_kernel), kK2aLds);
}

// The stages for pairs [off, off + P) with solver region w (stage arrays for P
// slots, then the control block holding K2a's overflow list) on stream st.
template <int MODE>
hipError_t pipeline(const geneig::cplx* a, const geneig::cplx* b, double* W, geneig::cplx* u, int* info, int off,
                    int P, unsigned char* w, hipStream_t st) {
    auto k1 = k1_kernel<k1_mode<MODE>()>;
    const size_t mat = (size_t)kN * kN;
    const unsigned g13 = (unsigned)((P + kBPB - 1) / kBPB);
    double* w_off = W ? W + (size_t)kN * off : nullptr;
    int* ovf = geneig_host::region_ctl(w, P, W54::PAIR_BYTES);
    if (eig_mode<MODE>()) {
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, ovf, 1);
        RETURN_IF(hipGetLastError());
    }
    if (split_chol<MODE>()) {
        hipLaunchKernelGGL(MODE == geneig::MODE_IP ? chol_ip_kernel : chol64_kernel, dim3((unsigned)P), dim3(64),
                           geneig::chol64_lds_bytes<kN>(), st, b + mat * off, w_off, u + mat * off, info + off, P);
        RETURN_IF(hipGetLastError());
    }
    hipLaunchKernelGGL(k1, dim3(g13), dim3(64, kBPB), kK1Lds * kBPB, st, a + mat * off, b + mat * off, w_off,
                       u + mat * off, info + off, w, P);
    RETURN_IF(hipGetLastError());
    if (!eig_mode<MODE>()) return hipSuccess;
    hipLaunchKernelGGL(k2a_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), kK2aLds, st, info + off, w, P,
                       geneig_host::ql_log_entries(), ovf);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2b_kernel, dim3((unsigned)P), dim3(64 * C54::R), 0, st, info + off, w, P);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2x_kernel, dim3((unsigned)std::min(P, geneig::kListGridK2x)), dim3(64 * C54::R), 0, st,
                       info + off, w, P, (const int*)ovf);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k3_kernel<MODE>, dim3(g13), dim3(64, kBPB), kK3Lds * kBPB, st, b + mat * off, w_off,
                       u + mat * off, info + off, w, P);
    return hipGetLastError();
}

// The solve pipeline over a batch, chunked to bound the workspace.
// MODE_FINISH takes M in A and L in B.  fixed_ws: a caller-provided
// workspace of fixed_bytes (the allocation-free entry point); the launch then
// runs on the caller's stream alone and allocates nothing.
template <int MODE>
hipError_t run(const void* A, const void* B, double* W, void* U, int* info, int batch, hipStream_t stream,
               unsigned char* fixed_ws = nullptr, size_t fixed_bytes = 0) {
    if (batch <= 0) return hipSuccess;
    RETURN_IF(set_attrs<MODE>());
    geneig_host::PlainLayout L = fixed_ws ? geneig_host::plain_layout(batch, W54::PAIR_BYTES, 1) : layout_of<MODE>(batch);
    void* ws = nullptr;
    if (fixed_ws) {
        if (L.chunk < batch || L.bytes > fixed_bytes) return hipErrorInvalidValue;
        ws = fixed_ws;
    } else if (MODE != geneig::MODE_CHOL) {   // every mode past the Cholesky uses the workspace
        RETURN_IF(geneig_host::workspace(stream, L.bytes, &ws));
    }
    const auto* a = static_cast<const geneig::cplx*>(A);
    const auto* b = static_cast<const geneig::cplx*>(B);
    auto* u = static_cast<geneig::cplx*>(U);
    unsigned char* base = static_cast<unsigned char*>(ws);
    if (L.parts > 1) {
        hipStream_t subs[geneig_host::kMaxParts];
        RETURN_IF(geneig_host::fork_streams(stream, L.parts, subs));
        hipError_t e = hipSuccess;
        for (int i = 0; i < L.parts && e == hipSuccess; ++i) {
            const int off = (int)((long long)batch * i / L.parts);
            const int P = (int)((long long)batch * (i + 1) / L.parts) - off;
            e = pipeline<MODE>(a, b, W, u, info, off, P, base + L.region * i, subs[i]);
        }
        const hipError_t j = geneig_host::join_streams(stream, L.parts, subs);
        return e != hipSuccess ? e : j;
    }
    for (int off = 0; off < batch; off += L.chunk)
        RETURN_IF(pipeline<MODE>(a, b, W, u, info, off, std::min(L.chunk, batch - off), base, stream));
    return hipSuccess;
}

// The fused spin-wave solve: per part, the Cholesky stage from the model,
// then K1 (A = eta) .. K3 as the batched FP64 solve.
hipError_t pipeline_sw(const SpinwaveModel& m, double* W, geneig::cplx* u, int* info, int off, int P,
                       unsigned char* w, hipStream_t st) {
    auto k1 = k1_kernel<geneig::MODE_SW>;
    const size_t mat = (size_t)kN * kN;
    int* ovf = geneig_host::region_ctl(w, P, W54::PAIR_BYTES);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, ovf, 1);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(chol_sw_kernel, dim3((unsigned)P), dim3(64), 0, st, m, W, u, info, off, P);
    RETURN_IF(hipGetLastError());
    double* w_off = W + (size_t)kN * off;
    geneig::cplx* u_off = u + mat * off;
    hipLaunchKernelGGL(k1, dim3((unsigned)P), dim3(64, kBPB), kK1Lds * kBPB, st, u_off, u_off, w_off, u_off,
                       info + off, w, P);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2a_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), kK2aLds, st, info + off, w, P,
                       geneig_host::ql_log_entries(), ovf);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2b_kernel, dim3((unsigned)P), dim3(64 * C54::R), 0, st, info + off, w, P);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2x_kernel, dim3((unsigned)std::min(P, geneig::kListGridK2x)), dim3(64 * C54::R), 0, st,
                       info + off, w, P, (const int*)ovf);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k3_kernel<geneig::MODE_FULL>, dim3((unsigned)P), dim3(64, kBPB), kK3Lds * kBPB, st, u_off,
                       w_off, u_off, info + off, w, P);
    return hipGetLastError();
}

hipError_t run_sw(const SpinwaveModel& m, double* W, geneig::cplx* U, int* info, int batch, hipStream_t stream) {
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1_kernel<geneig::MODE_SW>), kK1Lds * kBPB));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3_kernel<geneig::MODE_FULL>), kK3Lds * kBPB));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k2a_kernel), kK2aLds));
    if (batch <= 0) return hipSuccess;
    const geneig_host::PlainLayout L = layout_of<geneig::MODE_FULL>(batch);
    void* ws = nullptr;
    RETURN_IF(geneig_host::workspace(stream, L.bytes, &ws));
    unsigned char* base = static_cast<unsigned char*>(ws);
    if (L.parts > 1) {
        hipStream_t subs[geneig_host::kMaxParts];
        RETURN_IF(geneig_host::fork_streams(stream, L.parts, subs));
        hipError_t e = hipSuccess;
        for (int i = 0; i < L.parts && e == hipSuccess; ++i) {
            const int off = (int)((long long)batch * i / L.parts);
            const int P = (int)((long long)batch * (i + 1) / L.parts) - off;
            e = pipeline_sw(m, W, U, info, off, P, base + L.region * i, subs[i]);
        }
        const hipError_t j = geneig_host::join_streams(stream, L.parts, subs);
        return e != hipSuccess ? e : j;
    }
    for (int off = 0; off < batch; off += L.chunk)
        RETURN_IF(pipeline_sw(m, W, U, info, off, std::min(L.chunk, batch - off), base, stream));
    return hipSuccess;
}

// The real-arithmetic spin-wave solve: per part K1r, the FP64 QL stages, K3r.
hipError_t pipeline_real(const geneig::cplx* s, double* W, geneig::cplx* u, int* info, int off, int P,
                         unsigned char* w, hipStream_t st) {
    const size_t mat = (size_t)kN * kN;
    int* ovf = geneig_host::region_ctl(w, P, W54::PAIR_BYTES);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, ovf, 1);
    RETURN_IF(hipGetLastError());
    double* w_off = W + (size_t)kN * off;
    geneig::cplx* u_off = u + mat * off;
    hipLaunchKernelGGL(k1r_kernel, dim3((unsigned)P), dim3(64), RC54::LDS, st, s + mat * off, w_off, u_off,
                       info + off, w, P);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2a_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), kK2aLds, st, info + off, w, P,
                       geneig_host::ql_log_entries(), ovf);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2b_kernel, dim3((unsigned)P), dim3(64 * C54::R), 0, st, info + off, w, P);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k2x_kernel, dim3((unsigned)std::min(P, geneig::kListGridK2x)), dim3(64 * C54::R), 0, st,
                       info + off, w, P, (const int*)ovf);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(k3r_kernel, dim3((unsigned)P), dim3(64), RC54::LDS, st, w_off, u_off, info + off, w, P);
    return hipGetLastError();
}

hipError_t run_real(const geneig::cplx* S, double* W, geneig::cplx* U, int* info, int batch, hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1r_kernel), RC54::LDS));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3r_kernel), RC54::LDS));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k2a_kernel), kK2aLds));
    const geneig_host::PlainLayout L = layout_of<geneig::MODE_FULL>(batch);
    void* ws = nullptr;
    RETURN_IF(geneig_host::workspace(stream, L.bytes, &ws));
    unsigned char* base = static_cast<unsigned char*>(ws);
    if (L.parts > 1) {
        hipStream_t subs[geneig_host::kMaxParts];
        RETURN_IF(geneig_host::fork_streams(stream, L.parts, subs));
        hipError_t e = hipSuccess;
        for (int i = 0; i < L.parts && e == hipSuccess; ++i) {
            const int off = (int)((long long)batch * i / L.parts);
            const int P = (int)((long long)batch * (i + 1) / L.parts) - off;
            e = pipeline_real(S, W, U, info, off, P, base + L.region * i, subs[i]);
        }
        const hipError_t j = geneig_host::join_streams(stream, L.parts, subs);
        return e != hipSuccess ? e : j;
    }
    for (int off = 0; off < batch; off += L.chunk)
        RETURN_IF(pipeline_real(S, W, U, info, off, std::min(L.chunk, batch - off), base, stream));
    return hipSuccess;
}

}  // namespace

namespace geneig_host {

// The batched FP64 n = 54 solve into a caller-provided workspace of `bytes`
// (one region, the caller's stream only, nothing allocated).  Returns
// hipErrorInvalidValue when the batch's workspace does not fit.
hipError_t full_n54_into(const void* H, const void* S, double* W, void* U, int* info, int batch,
                         hipStream_t stream, unsigned char* ws, size_t bytes) {
    if (!ws) return hipErrorInvalidValue;
    return run<geneig::MODE_FULL>(H, S, W, U, info, batch, stream, ws, bytes);
}

// Workspace bytes of full_n54_into for `batch` pairs.
size_t full_n54_into_bytes(int batch) { return plain_layout(batch, W54::PAIR_BYTES, 1).bytes; }

}  // namespace geneig_host

extern "C" {

int chol_n54_launch(const hipDoubleComplex* d_S, hipDoubleComplex* d_L, int* d_info,
                    hipStream_t stream) {
    return run<geneig::MODE_CHOL>(nullptr, d_S, nullptr, d_L, d_info, 1, stream);
}
unsigned int chol_n54_block_dim_x(void) { return 64; }
unsigned int chol_n54_shared_mem_bytes(void) { return kWorkLds + kVecLds; }

int geneig_reduce_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                             hipDoubleComplex* d_M, int* d_info, hipStream_t stream) {
    return run<geneig::MODE_REDUCE>(d_H, d_S, nullptr, d_M, d_info, 1, stream);
}
unsigned int geneig_reduce_n54_block_dim_x(void) { return 64; }
unsigned int geneig_reduce_n54_shared_mem_bytes(void) { return kWorkLds + kVecLds; }

int geneig_reduce_eig_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                                 double* d_W, hipDoubleComplex* d_V, int* d_info,
                                 hipStream_t stream) {
    return run<geneig::MODE_REDUCE_EIG>(d_H, d_S, d_W, d_V, d_info, 1, stream);
}
unsigned int geneig_reduce_eig_n54_block_dim_x(void) { return 64; }
unsigned int geneig_reduce_eig_n54_shared_mem_bytes(void) { return kWorkLds + kVecLds; }
// The factor lives in the U/V buffer, so the "Cholesky region" of LDS is the
// working matrix itself; the eigen-stage region is the two broadcast vectors.
unsigned int geneig_reduce_eig_n54_chol_smem_bytes(void) { return kWorkLds; }
unsigned int geneig_reduce_eig_n54_heev_smem_bytes(void) { return kVecLds; }
unsigned int geneig_reduce_eig_n54_heev_workspace(void) { return C54::VEC; }

int geneig_full_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                           hipDoubleComplex* d_U, int* d_info, hipStream_t stream) {
    return run<geneig::MODE_FULL>(d_H, d_S, d_W, d_U, d_info, 1, stream);
}
unsigned int geneig_full_n54_block_dim_x(void) { return 64; }
unsigned int geneig_full_n54_shared_mem_bytes(void) { return kWorkLds + kVecLds; }

int geneig_full_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                                   double* d_W, hipDoubleComplex* d_U, int* d_info,
                                   int batch_size, hipStream_t stream) {
    return run<geneig::MODE_FULL>(d_H, d_S, d_W, d_U, d_info, batch_size, stream);
}
unsigned int geneig_full_n54_batched_block_dim_x(void) { return 64; }
unsigned int geneig_full_n54_batched_shared_mem_bytes(void) { return (kWorkLds + kVecLds) * kBPB; }

int geneig_full_n54_batched_inplace_launch(hipDoubleComplex* d_H, hipDoubleComplex* d_S, double* d_W, int* d_info,
                                           int batch_size, hipStream_t stream) {
    // the eigenvectors go over H, the factor over S (U = H: K1 reads all of
    // a pair's H before K3 writes that pair's eigenvectors)
    return run<geneig::MODE_IP>(d_H, d_S, d_W, d_H, d_info, batch_size, stream);
}

int geneig_finish_n54_launch(const hipDoubleComplex* d_L, const hipDoubleComplex* d_M, double* d_W,
                             hipDoubleComplex* d_U, int* d_info, int batch_size,
                             hipStream_t stream) {
    return run<geneig::MODE_FINISH>(d_M, d_L, d_W, d_U, d_info, batch_size, stream);
}
unsigned int geneig_finish_n54_block_dim_x(void) { return 64; }
unsigned int geneig_finish_n54_shared_mem_bytes(void) { return (kWorkLds + kVecLds) * kBPB; }

int geneig_spinwave_n54_launch(int nbonds, const int* d_shift, const int* d_r, const double* d_J, const double* d_q,
                               const double* d_diag, double* d_W, hipDoubleComplex* d_U, int* d_info, int batch_size,
                               hipStream_t stream) {
    if (batch_size <= 0) return hipSuccess;
    if (nbonds < 0 || (nbonds > 0 && (!d_shift || !d_r || !d_J)) || !d_q) return hipErrorInvalidValue;
    const SpinwaveModel m{nbonds, d_shift, d_r, d_J, d_q, d_diag};
    return run_sw(m, d_W, reinterpret_cast<geneig::cplx*>(d_U), d_info, batch_size, stream);
}

int geneig_eta_paired_n54_batched_launch(const hipDoubleComplex* d_S, double* d_W, hipDoubleComplex* d_U,
                                         int* d_info, int batch_size, hipStream_t stream) {
    return run_real(reinterpret_cast<const geneig::cplx*>(d_S), d_W, reinterpret_cast<geneig::cplx*>(d_U), d_info,
                    batch_size, stream);
}

int geneig_spinwave_n54_build(int nbonds, const int* d_shift, const int* d_r, const double* d_J, const double* d_q,
                              const double* d_diag, hipDoubleComplex* d_H, int batch_size, hipStream_t stream) {
    if (batch_size <= 0) return hipSuccess;
    if (nbonds < 0 || (nbonds > 0 && (!d_shift || !d_r || !d_J)) || !d_q || !d_H) return hipErrorInvalidValue;
    const SpinwaveModel m{nbonds, d_shift, d_r, d_J, d_q, d_diag};
    hipLaunchKernelGGL(sw_build_kernel, dim3((unsigned)batch_size), dim3(64), 0, stream, m,
                       reinterpret_cast<geneig::cplx*>(d_H), batch_size);
    return hipGetLastError();
}

size_t geneig_n54_workspace_bytes(int batch_size, int form) {
    if (batch_size <= 0) return 0;
    if (form == 1) return geneig_host::mixed_n54_layout(batch_size).bytes;
    return layout_of<geneig::MODE_FULL>(batch_size).bytes;   // also the in-place solve's
}

int geneig_n54_reserve(int batch_size, int form, hipStream_t stream) {
    if (form != 0 && form != 1) return hipErrorInvalidValue;
    if (batch_size <= 0) return hipSuccess;
    const int parts = form == 1 ? geneig_host::mixed_n54_layout(batch_size).parts
                                : layout_of<geneig::MODE_FULL>(batch_size).parts;
    RETURN_IF(geneig_host::prepare_parts(stream, parts));
    void* ws = nullptr;
    return geneig_host::workspace(stream, geneig_n54_workspace_bytes(batch_size, form), &ws);
}

int geneig_workspace_release(hipStream_t stream, size_t* freed_bytes) {
    return geneig_host::release(stream, freed_bytes);
}

}  // extern "C"
