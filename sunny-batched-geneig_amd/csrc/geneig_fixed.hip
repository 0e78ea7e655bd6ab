// geneig_fixed.hip -- static n = 54 entry points (hipcc, gfx950).
//
// MI355X counterpart of the reference's src/geneig_fixed.cu: the same six
// extern "C" launchers (chol, reduce, reduce+eig, full, batched full, finish)
// and their query functions, with the pipeline of geneig_device.inc behind
// them.  Every launcher returns a hipError_t value (0 = success) where the
// reference returned void; callers that ignore the result keep working.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <set>

#include "geneig_device.inc"
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
    geneig::k1_batched<kN, kBPB, MODE>(A, B, W, U, info, ws, batch, smem, vecs);
}

// The FP64 Cholesky alone (two waves per SIMD); the batched FP64 solve runs it
// ahead of K1, which then starts from the stashed factor (MODE_MIXED).
__global__ void __launch_bounds__(64, GENEIG_CHOL64_WAVES_PER_EU) chol64_kernel(const geneig::cplx* __restrict__ B,
                                                                           double* __restrict__ W,
                                                                           geneig::cplx* __restrict__ U,
                                                                           int* __restrict__ info, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ geneig::cplx vecs[C54::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_CHOL64, double>(B, B, W, U, info, nullptr, batch, smem, vecs);
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

__global__ void __launch_bounds__(64) k2a_kernel(int* __restrict__ info, unsigned char* ws, int batch, int cap) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::k2a_batched<kN, double>(info, ws, batch, reinterpret_cast<double*>(smem), cap);
}

__global__ void __launch_bounds__(64 * C54::R, GENEIG_K2B_WAVES_PER_EU) k2b_kernel(const int* __restrict__ info,
                                                                                 unsigned char* ws, int batch) {
    __shared__ geneig::cplx buf[geneig::k2b_buf_entries<kN, double>()];
    geneig::k2b_batched<kN, double>(info, ws, batch, buf);
}

__global__ void __launch_bounds__(64 * C54::R, GENEIG_K2_WAVES_PER_EU) k2x_kernel(int* __restrict__ info, unsigned char* ws, int batch) {
    __shared__ geneig::cplx rotlog[C54::R * C54::K2_VEC];
    geneig::k2x_batched<kN>(info, ws, batch, rotlog);
}

template <int MODE>
__global__ void __launch_bounds__(64 * kBPB, GENEIG_K3_WAVES_PER_EU) k3_kernel(const geneig::cplx* __restrict__ B,
                                                        double* __restrict__ W, geneig::cplx* __restrict__ U,
                                                        const int* __restrict__ info, unsigned char* ws,
                                                        int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ geneig::cplx vecs[kBPB * C54::VEC];
    geneig::k3_batched<kN, kBPB, MODE>(B, W, U, info, ws, batch, smem, vecs);
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

// The solve pipeline over a batch, chunked to bound the workspace.
// MODE_FINISH takes M in A and L in B.
template <int MODE>
hipError_t run(const void* A, const void* B, double* W, void* U, int* info, int batch, hipStream_t stream) {
    // the batched FP64 solve: the Cholesky as its own kernel (two waves per
    // SIMD), then K1 from the stashed factor; results are bit-identical (the
    // same operations per entry, only the update chunking differs)
    constexpr bool split_chol = (MODE == geneig::MODE_FULL);
    auto k1 = k1_kernel<split_chol ? geneig::MODE_MIXED : MODE>;
    auto k3 = k3_kernel<MODE>;
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1), kK1Lds * kBPB));
    if (split_chol)
        RETURN_IF(set_lds_once(reinterpret_cast<const void*>(chol64_kernel), geneig::chol64_lds_bytes<kN>()));

    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3), kK3Lds * kBPB));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k2a_kernel), kK2aLds));
    if (batch <= 0) return hipSuccess;
    constexpr bool eig = (MODE == geneig::MODE_FULL || MODE == geneig::MODE_REDUCE_EIG ||
                          MODE == geneig::MODE_FINISH);
    const int chunk = geneig_host::chunk_pairs(batch, W54::PAIR_BYTES);
    // a batched solve that fits one chunk runs as kPlainParts concurrent
    // parts (each with its own region of the workspace and its own stream)
    const int parts = split_chol ? geneig_host::part_count(batch, chunk, kPlainParts) : 1;
    const size_t region = [&](int P) { return ((size_t)P * W54::PAIR_BYTES + 255) / 256 * 256; }((batch + parts - 1) / parts);
    void* ws = nullptr;
    // every mode past the Cholesky uses the workspace (K1 stages W = L^{-1} A there)
    if (MODE != geneig::MODE_CHOL)
        RETURN_IF(geneig_host::workspace(stream, parts > 1 ? region * parts : (size_t)chunk * W54::PAIR_BYTES, &ws));
    const size_t mat = (size_t)kN * kN;
    const auto* a = static_cast<const geneig::cplx*>(A);
    const auto* b = static_cast<const geneig::cplx*>(B);
    auto* u = static_cast<geneig::cplx*>(U);
    // the stages for pairs [off, off + P) with workspace w on stream st
    auto pipeline = [&](int off, int P, unsigned char* w, hipStream_t st) -> hipError_t {
        const unsigned g13 = (unsigned)((P + kBPB - 1) / kBPB);
        if (split_chol) {
            hipLaunchKernelGGL(chol64_kernel, dim3((unsigned)P), dim3(64), geneig::chol64_lds_bytes<kN>(), st,
                               b + mat * off, W ? W + (size_t)kN * off : nullptr, u + mat * off, info + off, P);
            RETURN_IF(hipGetLastError());
        }
        hipLaunchKernelGGL(k1, dim3(g13), dim3(64, kBPB), kK1Lds * kBPB, st, a + mat * off,
                           b + mat * off, W ? W + (size_t)kN * off : nullptr, u + mat * off, info + off, w, P);
        RETURN_IF(hipGetLastError());
        if (!eig) return hipSuccess;
        hipLaunchKernelGGL(k2a_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), kK2aLds, st, info + off, w, P,
                           geneig_host::ql_log_entries());
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2b_kernel, dim3((unsigned)P), dim3(64 * C54::R), 0, st, info + off, w, P);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2x_kernel, dim3((unsigned)P), dim3(64 * C54::R), 0, st, info + off, w, P);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k3, dim3(g13), dim3(64, kBPB), kK3Lds * kBPB, st, b + mat * off,
                           W + (size_t)kN * off, u + mat * off, info + off, w, P);
        return hipGetLastError();
    };
    if (parts > 1) {
        hipStream_t subs[geneig_host::kMaxParts];
        RETURN_IF(geneig_host::fork_streams(stream, parts, subs));
        hipError_t e = hipSuccess;
        for (int i = 0; i < parts && e == hipSuccess; ++i) {
            const int off = (int)((long long)batch * i / parts);
            const int P = (int)((long long)batch * (i + 1) / parts) - off;
            e = pipeline(off, P, static_cast<unsigned char*>(ws) + region * i, subs[i]);
        }
        const hipError_t j = geneig_host::join_streams(stream, parts, subs);
        return e != hipSuccess ? e : j;
    }
    for (int off = 0; off < batch; off += chunk)
        RETURN_IF(pipeline(off, std::min(chunk, batch - off), static_cast<unsigned char*>(ws), stream));
    return hipSuccess;
}

}  // namespace

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

int geneig_finish_n54_launch(const hipDoubleComplex* d_L, const hipDoubleComplex* d_M, double* d_W,
                             hipDoubleComplex* d_U, int* d_info, int batch_size,
                             hipStream_t stream) {
    return run<geneig::MODE_FINISH>(d_M, d_L, d_W, d_U, d_info, batch_size, stream);
}
unsigned int geneig_finish_n54_block_dim_x(void) { return 64; }
unsigned int geneig_finish_n54_shared_mem_bytes(void) { return (kWorkLds + kVecLds) * kBPB; }

}  // extern "C"
