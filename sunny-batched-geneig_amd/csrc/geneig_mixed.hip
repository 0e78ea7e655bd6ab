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
//   FB   the listed pairs run through the FP64 pipeline (K1/K2/K3 over the list,
//        whose length the kernels read on the device: no host round trip)
// K1f..K3f and FB are chunked to bound the workspace, exactly as the FP64 path.
// The pencil (eta, S) of linear spin-wave theory, eta = diag(I_{n/2},
// -I_{n/2}), has its own launcher: A is never read (K1f and the fallback K1
// form eta in registers, the refinement scales X by eta instead of a matrix-
// core pass for A X), with the same arithmetic, so its outputs equal the
// general launcher's on the materialized pair.
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
    geneig::k1_batched<kN, 1, geneig::MODE_CHOL64, double>(B, B, W, U, info, nullptr, batch, batch, smem, vecs);
}

template <bool ETA>
__global__ void __launch_bounds__(64) k1f_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                  double* __restrict__ W, cplx* __restrict__ U,
                                                  int* __restrict__ info, unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplxf vecs[CF::VEC];
    geneig::k1_batched<kN, 1, ETA ? geneig::MODE_SW : geneig::MODE_MIXED, float>(A, B, W, U, info, ws, batch,
                                                                                  batch, smem, vecs);
}

// K2 as the FP64 path runs it: chases of 64 pairs per wave with the rotations
// logged (K2a), replay into Z (K2b), one-pair QL for an overflowed log (K2x).
constexpr unsigned kK2aLdsF = geneig::k2a_lds_bytes<kN, float>();
constexpr unsigned kK2aLdsD = geneig::k2a_lds_bytes<kN, double>();
constexpr unsigned kK3LdsF = geneig::k3_lds_bytes<kN, float>();
constexpr unsigned kK1LdsF = geneig::k1_lds_bytes<kN, float>();
constexpr unsigned kK1LdsD = geneig::k1_lds_bytes<kN, double>();
constexpr unsigned kK3LdsD = geneig::k3_lds_bytes<kN, double>();

__global__ void __launch_bounds__(64) k2af_kernel(int* __restrict__ info, unsigned char* ws, int batch, int cap,
                                                   int* __restrict__ ovf) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::k2a_batched<kN, float>(info, ws, batch, batch, reinterpret_cast<float*>(smem), cap, ovf);
}

__global__ void __launch_bounds__(64 * CF::R, GENEIG_K2F_WAVES_PER_EU) k2bf_kernel(const int* __restrict__ info,
                                                                                 unsigned char* ws, int batch) {
    __shared__ cplxf buf[geneig::k2b_buf_entries<kN, float>()];
    geneig::k2b_batched<kN, float>(info, ws, batch, batch, buf);
}

__global__ void __launch_bounds__(64 * CF::R, GENEIG_K2F_WAVES_PER_EU) k2xf_kernel(int* __restrict__ info,
                                                                                 unsigned char* ws, int chunk,
                                                                                 const int* __restrict__ ovf) {
    __shared__ cplxf rotlog[CF::R * CF::K2_VEC];
    geneig::k2x_batched<kN, float>(info, ws, chunk, rotlog, ovf);
}

__global__ void __launch_bounds__(64, GENEIG_K3F_WAVES_PER_EU) k3f_kernel(const cplx* __restrict__ B, double* __restrict__ W,
                                                  cplx* __restrict__ U, const int* __restrict__ info,
                                                  unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplxf vecs[CF::VEC];
    geneig::k3_batched<kN, 1, geneig::MODE_MIXED, float>(B, W, U, info, ws, batch, batch, smem, vecs);
}

// K3f of the paired form: only the positive half's columns (27 .. 53), the
// ones its refinement reads, are formed and written.
__global__ void __launch_bounds__(64, GENEIG_K3F_WAVES_PER_EU) k3f_half_kernel(const cplx* __restrict__ B,
                                                  double* __restrict__ W, cplx* __restrict__ U,
                                                  const int* __restrict__ info, unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplxf vecs[CF::VEC];
    geneig::k3_batched<kN, 1, geneig::MODE_MIXED, float, false, kN / 2>(B, W, U, info, ws, batch, batch, smem, vecs);
}

template <bool ETA>
__global__ void __launch_bounds__(64 * RC::NT, GENEIG_REFINE_WAVES_PER_EU) refine_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                              double* __restrict__ W, cplx* __restrict__ U,
                                                              const int* __restrict__ info, int* __restrict__ fb_list,
                                                              int* __restrict__ fb_count, int* __restrict__ iters,
                                                              int batch, int maxit, double tol, int first) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::refine_batched<kN, ETA>(A, B, W, U, info, fb_list, fb_count, iters, batch, maxit, tol, smem, first);
}

// The refinement of the spin-wave pencil's positive half (PAIRED): two waves
// per pair, n/2 columns of X in LDS, so six pairs per CU.
using RCP = geneig::RCfg<kN, true>;
#ifndef GENEIG_REFINE_PAIRED_WAVES_PER_EU
#define GENEIG_REFINE_PAIRED_WAVES_PER_EU (geneig::refine_waves_per_eu<kN, true>())
#endif
__global__ void __launch_bounds__(64 * RCP::NW, GENEIG_REFINE_PAIRED_WAVES_PER_EU)
    refine_paired_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B, double* __restrict__ W,
                         cplx* __restrict__ U, const int* __restrict__ info, int* __restrict__ fb_list,
                         int* __restrict__ fb_count, int* __restrict__ iters, int batch, int maxit, double tol,
                         int first) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::refine_batched<kN, true, true>(A, B, W, U, info, fb_list, fb_count, iters, batch, maxit, tol, smem,
                                           first);
}

// The FP64 pipeline over the fallback list: round `off` covers entries off ..
// off + chunk - 1 of the list, whose length the kernels read on the device
// (dispatch_slots), so the host never waits for it.
template <bool ETA>
__global__ void __launch_bounds__(64) k1l_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                  double* __restrict__ W, cplx* __restrict__ U,
                                                  int* __restrict__ info, unsigned char* ws, int chunk,
                                                  const int* __restrict__ list, const int* __restrict__ dcount,
                                                  int off) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplx vecs[CD::VEC];
    geneig::k1_batched<kN, 1, ETA ? geneig::MODE_FULL_SW : geneig::MODE_FULL, double, true>(
        A, B, W, U, info, ws,
        geneig::dispatch_slots(0, dcount, off, chunk), chunk, smem,
                                                         vecs, list + off);
}

__global__ void __launch_bounds__(64) k2al_kernel(int* __restrict__ info, unsigned char* ws, int chunk,
                                                   const int* __restrict__ list, const int* __restrict__ dcount,
                                                   int off, int cap, int* __restrict__ ovf) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::k2a_batched<kN, double, true>(info, ws, geneig::dispatch_slots(0, dcount, off, chunk), chunk,
                                    reinterpret_cast<double*>(smem), cap, ovf, list + off);
}

__global__ void __launch_bounds__(64 * CD::R, GENEIG_K2B_WAVES_PER_EU) k2bl_kernel(const int* __restrict__ info,
                                                                                 unsigned char* ws, int chunk,
                                                                                 const int* __restrict__ list,
                                                                                 const int* __restrict__ dcount,
                                                                                 int off) {
    __shared__ cplx buf[geneig::k2b_buf_entries<kN, double>()];
    geneig::k2b_batched<kN, double, true>(info, ws, geneig::dispatch_slots(0, dcount, off, chunk), chunk, buf, list + off);
}

__global__ void __launch_bounds__(64 * CD::R, GENEIG_K2_WAVES_PER_EU) k2xl_kernel(int* __restrict__ info,
                                                                                unsigned char* ws, int chunk,
                                                                                const int* __restrict__ list,
                                                                                int off,
                                                                                const int* __restrict__ ovf) {
    __shared__ cplx rotlog[CD::R * CD::K2_VEC];
    geneig::k2x_batched<kN, double>(info, ws, chunk, rotlog, ovf, list + off);
}

__global__ void __launch_bounds__(64, GENEIG_K3_WAVES_PER_EU) k3l_kernel(const cplx* __restrict__ B, double* __restrict__ W,
                                                  cplx* __restrict__ U, const int* __restrict__ info,
                                                  unsigned char* ws, int chunk, const int* __restrict__ list,
                                                  const int* __restrict__ dcount, int off) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplx vecs[CD::VEC];
    geneig::k3_batched<kN, 1, geneig::MODE_FULL, double, true>(B, W, U, info, ws,
                                                         geneig::dispatch_slots(0, dcount, off, chunk), chunk, smem,
                                                         vecs, list + off);
}

// Clears n ints with kernel stores, so a captured sequence carries no memset node.
__global__ void zero_kernel(int* p, int n) { geneig::zero_ints(p, n); }

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

// ETA: A = eta, not read (A may be any pointer; the kernels that would read it
// get B in its place).  PAIRED (with ETA): the refinement of the positive
// half of the spin-wave pencil's eigenpairs (refine_pair).
template <bool ETA, bool PAIRED = false>
hipError_t run_mixed(const cplx* A, const cplx* B, double* W, cplx* U, int* info, int batch, int maxit,
                     double tol, int* iters, int* h_fallback, hipStream_t stream) {
    static_assert(!PAIRED || ETA, "the paired refinement is for the pencil (eta, S)");
    if (ETA) A = B;
    if (h_fallback) *h_fallback = 0;
    if (batch <= 0) return hipSuccess;   // the library's rule: no work, no error
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(chol64_kernel), geneig::chol64_lds_bytes<kN>()));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1f_kernel<ETA>), kK1LdsF));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(PAIRED ? k3f_half_kernel : k3f_kernel), kK3LdsF));
    if (PAIRED)
        RETURN_IF(set_lds_once(reinterpret_cast<const void*>(refine_paired_kernel), RCP::LDS));
    else
        RETURN_IF(set_lds_once(reinterpret_cast<const void*>(refine_kernel<ETA>), RC::LDS));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1l_kernel<ETA>), kK1LdsD));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3l_kernel), kK3LdsD));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k2af_kernel), kK2aLdsF));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k2al_kernel), kK2aLdsD));
    if (maxit <= 0) maxit = kMaxIter;
    if (!(tol > 0.0)) tol = kTol;

    // workspace: [fallback counter + list | solver area]; a batch that fits
    // one complex64 chunk runs as kMixedParts concurrent parts, each with its
    // own region of the solver area and its own stream
    const geneig_host::MixedLayout L = geneig_host::mixed_n54_layout(batch);
    void* ws = nullptr;
    RETURN_IF(geneig_host::workspace(stream, L.bytes, &ws));
    int* fb_count = static_cast<int*>(ws);
    int* fb_list = fb_count + 1;
    unsigned char* solver = static_cast<unsigned char*>(ws) + L.list_bytes;
    const size_t mat = (size_t)kN * kN;

    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, fb_count, 1);
    RETURN_IF(hipGetLastError());
    // the complex64 stages for pairs [off, off + P) with solver region w on stream st
    auto bulk = [&](int off, int P, unsigned char* w, hipStream_t st) -> hipError_t {
        int* ovf = geneig_host::region_ctl(w, P, WF::PAIR_BYTES);
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, ovf, 1);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k1f_kernel<ETA>, dim3((unsigned)P), dim3(64), kK1LdsF, st, A + mat * off, B + mat * off,
                           W + (size_t)kN * off, U + mat * off, info + off, w, P);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2af_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), kK2aLdsF, st, info + off, w, P,
                           geneig_host::ql_log_entries(), ovf);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2bf_kernel, dim3((unsigned)P), dim3(64 * CF::R), 0, st, info + off, w, P);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2xf_kernel, dim3((unsigned)std::min(P, geneig::kListGridK2x)), dim3(64 * CF::R), 0, st,
                           info + off, w, P, (const int*)ovf);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(PAIRED ? k3f_half_kernel : k3f_kernel, dim3((unsigned)P), dim3(64), kK3LdsF, st,
                           B + mat * off, W + (size_t)kN * off, U + mat * off, info + off, w, P);
        return hipGetLastError();
    };
    // the refinement of pairs [first, end) on stream st, one workgroup per pair
    auto refine = [&](unsigned grid, hipStream_t st, int end, int first) -> hipError_t {
        if constexpr (PAIRED)
            hipLaunchKernelGGL(refine_paired_kernel, dim3(grid), dim3(64 * RCP::NW), RCP::LDS, st, A, B, W, U, info,
                               fb_list, fb_count, iters, end, maxit, tol, first);
        else
            hipLaunchKernelGGL(refine_kernel<ETA>, dim3(grid), dim3(64 * RC::NT), RC::LDS, st, A, B, W, U, info,
                               fb_list, fb_count, iters, end, maxit, tol, first);
        return hipGetLastError();
    };
    if (L.parts > 1) {
        hipStream_t subs[geneig_host::kMaxParts];
        RETURN_IF(geneig_host::fork_streams(stream, L.parts, subs));
        hipError_t e = hipSuccess;
        for (int i = 0; i < L.parts && e == hipSuccess; ++i) {
            const int off = (int)((long long)batch * i / L.parts);
            const int P = (int)((long long)batch * (i + 1) / L.parts) - off;
            hipLaunchKernelGGL(chol64_kernel, dim3((unsigned)P), dim3(64), geneig::chol64_lds_bytes<kN>(), subs[i],
                               B + mat * off, W + (size_t)kN * off, U + mat * off, info + off, P);
            e = hipGetLastError();
            if (e == hipSuccess) e = bulk(off, P, solver + L.region_f * i, subs[i]);
            if (e == hipSuccess) e = refine((unsigned)P, subs[i], off + P, off);
        }
        const hipError_t j = geneig_host::join_streams(stream, L.parts, subs);
        RETURN_IF(e);
        RETURN_IF(j);
    } else {
        hipLaunchKernelGGL(chol64_kernel, dim3((unsigned)batch), dim3(64), geneig::chol64_lds_bytes<kN>(), stream, B,
                           W, U, info, batch);
        RETURN_IF(hipGetLastError());
        for (int off = 0; off < batch; off += L.chunk_f)
            RETURN_IF(bulk(off, std::min(L.chunk_f, batch - off), solver, stream));
        RETURN_IF(refine((unsigned)batch, stream, batch, 0));
    }

    // The FP64 re-solve of the listed pairs, in rounds of chunk_d list
    // entries; every dispatch reads the list's length on the device and walks
    // it with a fixed grid (an empty list costs a few microseconds), so the
    // launch never waits for the host.
    int* fovf = geneig_host::region_ctl(solver, L.chunk_d, WD::PAIR_BYTES);
    const int cd = L.chunk_d;
    const unsigned g1 = (unsigned)std::min(cd, geneig::kListGridK1);
    const unsigned g2a = (unsigned)std::min((cd + 63) / 64, geneig::kListGridK2a);
    const unsigned g2x = (unsigned)std::min(cd, geneig::kListGridK2x);
    const int* lst = fb_list;
    const int* cnt = fb_count;
    for (int r = 0; r < L.rounds; ++r) {
        const int off = r * cd;
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, fovf, 1);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k1l_kernel<ETA>, dim3(g1), dim3(64), kK1LdsD, stream, A, B, W, U, info, solver, cd, lst, cnt,
                           off);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2al_kernel, dim3(g2a), dim3(64), kK2aLdsD, stream, info, solver, cd, lst, cnt, off,
                           geneig_host::ql_log_entries(), fovf);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2bl_kernel, dim3(g1), dim3(64 * CD::R), 0, stream, info, solver, cd, lst, cnt, off);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2xl_kernel, dim3(g2x), dim3(64 * CD::R), 0, stream, info, solver, cd, lst, off,
                           (const int*)fovf);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k3l_kernel, dim3(g1), dim3(64), kK3LdsD, stream, B, W, U, info, solver, cd, lst, cnt, off);
        RETURN_IF(hipGetLastError());
    }
    if (h_fallback) {
        // reported on request only: this is the launch's one wait for the device
        int count = 0;
        RETURN_IF(hipMemcpyAsync(&count, fb_count, sizeof(int), hipMemcpyDeviceToHost, stream));
        RETURN_IF(hipStreamSynchronize(stream));
        if (count < 0 || count > batch) return hipErrorUnknown;
        *h_fallback = count;
    }
    return hipSuccess;
}

}  // namespace

namespace geneig_host {
MixedLayout mixed_n54_layout(int batch) {
    return mixed_layout(batch, WF::PAIR_BYTES, WD::PAIR_BYTES, kMixedParts);
}
}  // namespace geneig_host

extern "C" {

int geneig_mixed_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                                    hipDoubleComplex* d_U, int* d_info, int batch_size, int max_iter,
                                    double tol, int* d_iters, int* h_fallback, hipStream_t stream) {
    return run_mixed<false>(reinterpret_cast<const cplx*>(d_H), reinterpret_cast<const cplx*>(d_S), d_W,
                            reinterpret_cast<cplx*>(d_U), d_info, batch_size, max_iter, tol, d_iters, h_fallback,
                            stream);
}

int geneig_mixed_eta_n54_batched_launch(const hipDoubleComplex* d_S, double* d_W, hipDoubleComplex* d_U, int* d_info,
                                        int batch_size, int max_iter, double tol, int* d_iters, int* h_fallback,
                                        hipStream_t stream) {
    return run_mixed<true>(nullptr, reinterpret_cast<const cplx*>(d_S), d_W, reinterpret_cast<cplx*>(d_U), d_info,
                           batch_size, max_iter, tol, d_iters, h_fallback, stream);
}

int geneig_mixed_eta_paired_n54_batched_launch(const hipDoubleComplex* d_S, double* d_W, hipDoubleComplex* d_U,
                                               int* d_info, int batch_size, int max_iter, double tol, int* d_iters,
                                               int* h_fallback, hipStream_t stream) {
    return run_mixed<true, true>(nullptr, reinterpret_cast<const cplx*>(d_S), d_W, reinterpret_cast<cplx*>(d_U),
                                 d_info, batch_size, max_iter, tol, d_iters, h_fallback, stream);
}

int geneig_mixed_default_max_iter(void) { return kMaxIter; }
double geneig_mixed_default_tol(void) { return kTol; }

}  // extern "C"
