// geneig_fixed.hip -- static n = 54 entry points (hipcc, gfx950).
//
// MI355X counterpart of the reference's src/geneig_fixed.cu: the same six
// extern "C" launchers (chol, reduce, reduce+eig, full, batched full, finish)
// and their query functions, with the pipeline of geneig_device.inc behind
// them.  Every launcher returns a hipError_t value (0 = success) where the
// reference returned void; callers that ignore the result keep working.
#include <hip/hip_runtime.h>

#include <mutex>

#include "geneig_device.inc"
#include "../../include/geneig.h"

namespace {

constexpr int kN = 54;
using C54 = geneig::Cfg<kN>;
// The batched launcher packs this many independent pairs (one wave each)
// into a workgroup; 1 keeps the LDS allocation per workgroup at one pair so
// the hardware can place three workgroups on a CU.
constexpr int kBatchedBPB = 1;

template <int MODE>
__global__ void __launch_bounds__(64) geneig_single_kernel(const geneig::cplx* __restrict__ A,
                                                           const geneig::cplx* __restrict__ B,
                                                           double* __restrict__ W,
                                                           geneig::cplx* __restrict__ U,
                                                           int* __restrict__ info) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::solve_pair<kN, MODE>(A, B, W, U, info, smem, threadIdx.x);
}

template <int MODE, int BPB>
__global__ void __launch_bounds__(64 * BPB) geneig_batched_kernel(const geneig::cplx* __restrict__ A,
                                                                  const geneig::cplx* __restrict__ B,
                                                                  double* __restrict__ W,
                                                                  geneig::cplx* __restrict__ U,
                                                                  int* __restrict__ info, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::batched<kN, BPB, MODE>(A, B, W, U, info, batch, smem);
}

// hipFuncSetAttribute once per kernel per process, thread-safe (the reference
// used an unguarded static bool, src/geneig_fixed.cu:138-144).
template <typename K>
hipError_t set_lds_once(std::once_flag& flag, hipError_t& status, K kernel, unsigned bytes) {
    std::call_once(flag, [&] {
        status = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    });
    return status;
}

template <int MODE>
hipError_t launch_single(const void* A, const void* B, double* W, void* U, int* info,
                         hipStream_t stream) {
    static std::once_flag flag;
    static hipError_t st = hipSuccess;
    auto k = geneig_single_kernel<MODE>;
    hipError_t e = set_lds_once(flag, st, k, C54::BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(1), dim3(64), C54::BYTES, stream,
                       static_cast<const geneig::cplx*>(A), static_cast<const geneig::cplx*>(B), W,
                       static_cast<geneig::cplx*>(U), info);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_batched(const void* A, const void* B, double* W, void* U, int* info, int batch,
                          hipStream_t stream) {
    static std::once_flag flag;
    static hipError_t st = hipSuccess;
    auto k = geneig_batched_kernel<MODE, kBatchedBPB>;
    hipError_t e = set_lds_once(flag, st, k, C54::BYTES * kBatchedBPB);
    if (e != hipSuccess) return e;
    if (batch <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((batch + kBatchedBPB - 1) / kBatchedBPB);
    hipLaunchKernelGGL(k, dim3(grid), dim3(64, kBatchedBPB), C54::BYTES * kBatchedBPB, stream,
                       static_cast<const geneig::cplx*>(A), static_cast<const geneig::cplx*>(B), W,
                       static_cast<geneig::cplx*>(U), info, batch);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int chol_n54_launch(const hipDoubleComplex* d_S, hipDoubleComplex* d_L, int* d_info,
                    hipStream_t stream) {
    return launch_single<geneig::MODE_CHOL>(nullptr, d_S, nullptr, d_L, d_info, stream);
}
unsigned int chol_n54_block_dim_x(void) { return 64; }
unsigned int chol_n54_shared_mem_bytes(void) { return C54::BYTES; }

int geneig_reduce_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                             hipDoubleComplex* d_M, int* d_info, hipStream_t stream) {
    return launch_single<geneig::MODE_REDUCE>(d_H, d_S, nullptr, d_M, d_info, stream);
}
unsigned int geneig_reduce_n54_block_dim_x(void) { return 64; }
unsigned int geneig_reduce_n54_shared_mem_bytes(void) { return C54::BYTES; }

int geneig_reduce_eig_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                                 double* d_W, hipDoubleComplex* d_V, int* d_info,
                                 hipStream_t stream) {
    return launch_single<geneig::MODE_REDUCE_EIG>(d_H, d_S, d_W, d_V, d_info, stream);
}
unsigned int geneig_reduce_eig_n54_block_dim_x(void) { return 64; }
unsigned int geneig_reduce_eig_n54_shared_mem_bytes(void) { return C54::BYTES; }
// The factor lives in the U/V buffer, so the "Cholesky region" of LDS is the
// working matrix itself; the eigen-stage region adds the two broadcast vectors.
unsigned int geneig_reduce_eig_n54_chol_smem_bytes(void) { return C54::OFF_COL; }
unsigned int geneig_reduce_eig_n54_heev_smem_bytes(void) { return C54::BYTES - C54::OFF_COL; }
unsigned int geneig_reduce_eig_n54_heev_workspace(void) { return 2 * C54::NP; }

int geneig_full_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                           hipDoubleComplex* d_U, int* d_info, hipStream_t stream) {
    return launch_single<geneig::MODE_FULL>(d_H, d_S, d_W, d_U, d_info, stream);
}
unsigned int geneig_full_n54_block_dim_x(void) { return 64; }
unsigned int geneig_full_n54_shared_mem_bytes(void) { return C54::BYTES; }

int geneig_full_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                                   double* d_W, hipDoubleComplex* d_U, int* d_info,
                                   int batch_size, hipStream_t stream) {
    return launch_batched<geneig::MODE_FULL>(d_H, d_S, d_W, d_U, d_info, batch_size, stream);
}
unsigned int geneig_full_n54_batched_block_dim_x(void) { return 64; }
unsigned int geneig_full_n54_batched_shared_mem_bytes(void) { return C54::BYTES * kBatchedBPB; }

int geneig_finish_n54_launch(const hipDoubleComplex* d_L, const hipDoubleComplex* d_M, double* d_W,
                             hipDoubleComplex* d_U, int* d_info, int batch_size,
                             hipStream_t stream) {
    return launch_batched<geneig::MODE_FINISH>(d_M, d_L, d_W, d_U, d_info, batch_size, stream);
}
unsigned int geneig_finish_n54_block_dim_x(void) { return 64; }
unsigned int geneig_finish_n54_shared_mem_bytes(void) { return C54::BYTES * kBatchedBPB; }

}  // extern "C"
