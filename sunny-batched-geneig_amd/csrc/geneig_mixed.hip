// geneig_mixed.hip -- precision form of the n = 54 batched solve (hipcc, gfx950).
//
// Stages per batch (all on the caller's stream):
//   C0   FP64 Cholesky of B, L stashed in U (zhegv numbering and NaN outputs
//        when B is not positive definite)
//   K1f  complex64 two-sided reduction, Householder tridiagonalization, Q
//   K2f  float implicit QL (K2a chases with a rotation log, K2b replay, K2x
//        for an overflowed log)
//   K3f  sort, Y = Q Z, X0 = L^{-H} Y (complex64 arithmetic, complex128 out)
//   RF   FP64 refinement on the matrix cores; pairs it cannot finish go to a
//        device list
//   FB   the listed pairs run through the FP64 pipeline (K1/K2/K3 over the list)
// K1f..K3f and FB are chunked to bound the workspace, exactly as the FP64 path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <set>

#include "geneig_device.inc"
#include "geneig_refine.inc"
#include "geneig_ws.hpp"
#include "../../include/geneig_mixed.h"

#ifndef GENEIG_K2F_WAVES_PER_EU
#define GENEIG_K2F_WAVES_PER_EU 4
#endif
// Concurrent parts of a precision-form launch: one part's matrix-core
// refinement overlaps another's complex64 stages.  Three measured fastest at
// 131,072 pairs (55.6 ms against 56.8 on one stream, 56.1 with two).
#ifndef GENEIG_MIXED_PARTS
#define GENEIG_MIXED_PARTS 3
#endif
constexpr int kMixedParts = GENEIG_MIXED_PARTS;

#ifndef GENEIG_REFINE_WAVES_PER_EU
#define GENEIG_REFINE_WAVES_PER_EU (geneig::refine_waves_per_eu<54>())
#endif

namespace {

constexpr int kN = 54;
constexpr int kMaxIter = 4;
constexpr double kTol = 1e-7;
using CD = geneig::Cfg<kN, double>;
using CF = geneig::Cfg<kN, float>;
using WD = geneig::Ws<kN, double>;
using WF = geneig::Ws<kN, float>;
using RC = geneig::RCfg<kN>;
using geneig::cplx;
using geneig::cplxf;

__global__ void __launch_bounds__(64, GENEIG_CHOL64_WAVES_PER_EU) chol64_kernel(const cplx* __restrict__ B, double* __restrict__ W,
                                                     cplx* __restrict__ U, int* __restrict__ info, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplx vecs[CD::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_CHOL64, double>(B, B, W, U, info, nullptr, batch, smem, vecs);
}

__global__ void __launch_bounds__(64) k1f_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                  double* __restrict__ W, cplx* __restrict__ U,
                                                  int* __restrict__ info, unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplxf vecs[CF::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_MIXED, float>(A, B, W, U, info, ws, batch, smem, vecs);
}

// K2 as the FP64 path runs it: chases of 64 pairs per wave with the rotations
// logged (K2a), replay into Z (K2b), one-pair QL for an overflowed log (K2x).
constexpr unsigned kK2aLdsF = geneig::k2a_lds_bytes<kN, float>();
constexpr unsigned kK2aLdsD = geneig::k2a_lds_bytes<kN, double>();
constexpr unsigned kK3LdsF = geneig::k3_lds_bytes<kN, float>();
constexpr unsigned kK1LdsF = geneig::k1_lds_bytes<kN, float>();
constexpr unsigned kK1LdsD = geneig::k1_lds_bytes<kN, double>();
constexpr unsigned kK3LdsD = geneig::k3_lds_bytes<kN, double>();

__global__ void __launch_bounds__(64) k2af_kernel(int* __restrict__ info, unsigned char* ws, int batch, int cap) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::k2a_batched<kN, float>(info, ws, batch, reinterpret_cast<float*>(smem), cap);
}

__global__ void __launch_bounds__(64 * CF::R, GENEIG_K2F_WAVES_PER_EU) k2bf_kernel(const int* __restrict__ info,
                                                                                 unsigned char* ws, int batch) {
    __shared__ cplxf buf[geneig::k2b_buf_entries<kN, float>()];
    geneig::k2b_batched<kN, float>(info, ws, batch, buf);
}

__global__ void __launch_bounds__(64 * CF::R, GENEIG_K2F_WAVES_PER_EU) k2xf_kernel(int* __restrict__ info,
                                                                                 unsigned char* ws, int batch) {
    __shared__ cplxf rotlog[CF::R * CF::K2_VEC];
    geneig::k2x_batched<kN, float>(info, ws, batch, rotlog);
}

__global__ void __launch_bounds__(64, GENEIG_K3F_WAVES_PER_EU) k3f_kernel(const cplx* __restrict__ B, double* __restrict__ W,
                                                  cplx* __restrict__ U, const int* __restrict__ info,
                                                  unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplxf vecs[CF::VEC];
    geneig::k3_batched<kN, 1, geneig::MODE_MIXED, float>(B, W, U, info, ws, batch, smem, vecs);
}

__global__ void __launch_bounds__(64 * RC::NT, GENEIG_REFINE_WAVES_PER_EU) refine_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                              double* __restrict__ W, cplx* __restrict__ U,
                                                              const int* __restrict__ info, int* __restrict__ fb_list,
                                                              int* __restrict__ fb_count, int* __restrict__ iters,
                                                              int batch, int maxit, double tol, int first) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::refine_batched<kN>(A, B, W, U, info, fb_list, fb_count, iters, batch, maxit, tol, smem, first);
}

// The FP64 pipeline over a list of pairs (the fallback).
__global__ void __launch_bounds__(64) k1l_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                  double* __restrict__ W, cplx* __restrict__ U,
                                                  int* __restrict__ info, unsigned char* ws, int count,
                                                  const int* __restrict__ list) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplx vecs[CD::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_FULL, double>(A, B, W, U, info, ws, count, smem, vecs, list);
}

__global__ void __launch_bounds__(64) k2al_kernel(int* __restrict__ info, unsigned char* ws, int count,
                                                   const int* __restrict__ list, int cap) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::k2a_batched<kN, double>(info, ws, count, reinterpret_cast<double*>(smem), cap, list);
}

__global__ void __launch_bounds__(64 * CD::R, GENEIG_K2B_WAVES_PER_EU) k2bl_kernel(const int* __restrict__ info,
                                                                                 unsigned char* ws, int count,
                                                                                 const int* __restrict__ list) {
    __shared__ cplx buf[geneig::k2b_buf_entries<kN, double>()];
    geneig::k2b_batched<kN, double>(info, ws, count, buf, list);
}

__global__ void __launch_bounds__(64 * CD::R, GENEIG_K2_WAVES_PER_EU) k2xl_kernel(int* __restrict__ info,
                                                                                unsigned char* ws, int count,
                                                                                const int* __restrict__ list) {
    __shared__ cplx rotlog[CD::R * CD::K2_VEC];
    geneig::k2x_batched<kN, double>(info, ws, count, rotlog, list);
}

__global__ void __launch_bounds__(64, GENEIG_K3_WAVES_PER_EU) k3l_kernel(const cplx* __restrict__ B, double* __restrict__ W,
                                                  cplx* __restrict__ U, const int* __restrict__ info,
                                                  unsigned char* ws, int count, const int* __restrict__ list) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplx vecs[CD::VEC];
    geneig::k3_batched<kN, 1, geneig::MODE_FULL, double>(B, W, U, info, ws, count, smem, vecs, list);
}

// Clears the fallback counter with a kernel store, so a captured sequence
// carries no memset node.
__global__ void zero_kernel(int* p) { *p = 0; }

std::mutex g_attr_mu;
std::set<const void*> g_attr_done;

hipError_t set_lds_once(const void* kernel, unsigned bytes) {
    std::lock_guard<std::mutex> lk(g_attr_mu);
    if (g_attr_done.count(kernel)) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) g_attr_done.insert(kernel);
    return e;
}

#define RETURN_IF(expr)                  \
    do {                                 \
        hipError_t _e = (expr);          \
        if (_e != hipSuccess) return _e; \
    } while (0)

hipError_t run_mixed(const cplx* A, const cplx* B, double* W, cplx* U, int* info, int batch, int maxit,
                     double tol, int* iters, int* h_fallback, hipStream_t stream) {
    if (h_fallback) *h_fallback = 0;
    if (batch < 0) return hipErrorInvalidValue;
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(chol64_kernel), geneig::chol64_lds_bytes<kN>()));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1f_kernel), kK1LdsF));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3f_kernel), kK3LdsF));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(refine_kernel), RC::LDS));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1l_kernel), kK1LdsD));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3l_kernel), kK3LdsD));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k2af_kernel), kK2aLdsF));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k2al_kernel), kK2aLdsD));
    if (batch == 0) return hipSuccess;
    if (maxit <= 0) maxit = kMaxIter;
    if (!(tol > 0.0)) tol = kTol;

    // workspace: [fallback counter + list | solver chunk area]
    const size_t list_bytes = ((sizeof(int) * ((size_t)batch + 1) + 255) / 256) * 256;
    const int chunk_f = geneig_host::chunk_pairs(batch, WF::PAIR_BYTES);
    const int chunk_d = geneig_host::chunk_pairs(batch, WD::PAIR_BYTES);
    // a batch that fits one complex64 chunk runs as kMixedParts concurrent
    // parts, each with its own region of the solver area and its own stream
    const int parts = geneig_host::part_count(batch, chunk_f, kMixedParts);
    const size_t region = (((size_t)((batch + parts - 1) / parts)) * WF::PAIR_BYTES + 255) / 256 * 256;
    const size_t solver_bytes = std::max(parts > 1 ? region * parts : (size_t)chunk_f * WF::PAIR_BYTES,
                                         (size_t)chunk_d * WD::PAIR_BYTES);
    void* ws = nullptr;
    RETURN_IF(geneig_host::workspace(stream, list_bytes + solver_bytes, &ws));
    int* fb_count = static_cast<int*>(ws);
    int* fb_list = fb_count + 1;
    unsigned char* solver = static_cast<unsigned char*>(ws) + list_bytes;
    const size_t mat = (size_t)kN * kN;

    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(1), 0, stream, fb_count);
    RETURN_IF(hipGetLastError());
    // the complex64 stages for pairs [off, off + P) with solver area w on stream st
    auto bulk = [&](int off, int P, unsigned char* w, hipStream_t st) -> hipError_t {
        hipLaunchKernelGGL(k1f_kernel, dim3((unsigned)P), dim3(64), kK1LdsF, st, A + mat * off, B + mat * off,
                           W + (size_t)kN * off, U + mat * off, info + off, w, P);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2af_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), kK2aLdsF, st, info + off, w, P,
                           geneig_host::ql_log_entries());
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2bf_kernel, dim3((unsigned)P), dim3(64 * CF::R), 0, st, info + off, w, P);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2xf_kernel, dim3((unsigned)P), dim3(64 * CF::R), 0, st, info + off, w, P);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k3f_kernel, dim3((unsigned)P), dim3(64), kK3LdsF, st, B + mat * off,
                           W + (size_t)kN * off, U + mat * off, info + off, w, P);
        return hipGetLastError();
    };
    if (parts > 1) {
        hipStream_t subs[geneig_host::kMaxParts];
        RETURN_IF(geneig_host::fork_streams(stream, parts, subs));
        hipError_t e = hipSuccess;
        for (int i = 0; i < parts && e == hipSuccess; ++i) {
            const int off = (int)((long long)batch * i / parts);
            const int P = (int)((long long)batch * (i + 1) / parts) - off;
            hipLaunchKernelGGL(chol64_kernel, dim3((unsigned)P), dim3(64), geneig::chol64_lds_bytes<kN>(), subs[i],
                               B + mat * off, W + (size_t)kN * off, U + mat * off, info + off, P);
            e = hipGetLastError();
            if (e == hipSuccess) e = bulk(off, P, solver + region * i, subs[i]);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(refine_kernel, dim3((unsigned)P), dim3(64 * RC::NT), RC::LDS, subs[i], A, B, W, U,
                                   info, fb_list, fb_count, iters, off + P, maxit, tol, off);
                e = hipGetLastError();
            }
        }
        const hipError_t j = geneig_host::join_streams(stream, parts, subs);
        RETURN_IF(e);
        RETURN_IF(j);
    } else {
        hipLaunchKernelGGL(chol64_kernel, dim3((unsigned)batch), dim3(64), geneig::chol64_lds_bytes<kN>(), stream, B,
                           W, U, info, batch);
        RETURN_IF(hipGetLastError());
        for (int off = 0; off < batch; off += chunk_f) RETURN_IF(bulk(off, std::min(chunk_f, batch - off), solver, stream));
        hipLaunchKernelGGL(refine_kernel, dim3((unsigned)batch), dim3(64 * RC::NT), RC::LDS, stream, A, B, W, U, info,
                           fb_list, fb_count, iters, batch, maxit, tol, 0);
        RETURN_IF(hipGetLastError());
    }

    int count = 0;
    RETURN_IF(hipMemcpyAsync(&count, fb_count, sizeof(int), hipMemcpyDeviceToHost, stream));
    RETURN_IF(hipStreamSynchronize(stream));
    if (count < 0 || count > batch) return hipErrorUnknown;
    if (h_fallback) *h_fallback = count;
    for (int off = 0; off < count; off += chunk_d) {
        const int P = std::min(chunk_d, count - off);
        hipLaunchKernelGGL(k1l_kernel, dim3((unsigned)P), dim3(64), kK1LdsD, stream, A, B, W, U, info, solver, P,
                           fb_list + off);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2al_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), kK2aLdsD, stream, info, solver,
                           P, fb_list + off, geneig_host::ql_log_entries());
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2bl_kernel, dim3((unsigned)P), dim3(64 * CD::R), 0, stream, info, solver, P,
                           fb_list + off);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2xl_kernel, dim3((unsigned)P), dim3(64 * CD::R), 0, stream, info, solver, P,
                           fb_list + off);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k3l_kernel, dim3((unsigned)P), dim3(64), kK3LdsD, stream, B, W, U, info, solver, P,
                           fb_list + off);
        RETURN_IF(hipGetLastError());
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

int geneig_mixed_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                                    hipDoubleComplex* d_U, int* d_info, int batch_size, int max_iter,
                                    double tol, int* d_iters, int* h_fallback, hipStream_t stream) {
    return run_mixed(reinterpret_cast<const cplx*>(d_H), reinterpret_cast<const cplx*>(d_S), d_W,
                     reinterpret_cast<cplx*>(d_U), d_info, batch_size, max_iter, tol, d_iters, h_fallback, stream);
}

int geneig_mixed_default_max_iter(void) { return kMaxIter; }
double geneig_mixed_default_tol(void) { return kTol; }

}  // extern "C"
