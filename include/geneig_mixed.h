/* geneig_mixed.h -- precision form of the batched n = 54 generalized
 * Hermitian-definite eigensolve on gfx950.
 *
 * No counterpart in the reference: src/geneig_fixed.cu offers the FP64
 * pipeline only (geneig_full_n54_batched_launch, src/geneig_fixed.cu:474).
 * This entry point takes the same arguments and produces the same outputs
 * (ascending eigenvalues W, B-orthonormal eigenvectors U, zhegv-numbered
 * info), computed as:
 *   1. FP64 Cholesky of B (B not positive definite: info = n + j, NaN outputs,
 *      exactly as the FP64 path);
 *   2. complex64 reduction C = L^-1 A L^-H, tridiagonalization, QL and
 *      back-transform to an approximate X;
 *   3. FP64 refinement of (lambda, X) against the original A and B on the FP64
 *      matrix cores, until the largest correction is <= tol;
 *   4. pairs whose refinement diverges, stalls after max_iter steps, or meets
 *      coinciding eigenvalues are solved again by the FP64 pipeline.
 *
 * Step 4 reads the number of listed pairs on the device (its kernels walk
 * the list with fixed grids), so the launch does not wait for the device
 * unless h_fallback is given: then it waits once, at the end, to report the
 * count.  With h_fallback NULL the launch is asynchronous and can be captured
 * into a graph (after geneig_n54_reserve on the capturing stream).
 * max_iter <= 0 selects 4, tol <= 0 selects 1e-7 (the refined X then carries
 * errors of order tol^2).  d_iters (nullable, device, batch_size ints)
 * receives the refinement steps taken per pair (0: handed to step 4 before
 * refining).  h_fallback (nullable, host) receives the number of pairs step 4
 * solved again.  Returns a hipError_t value (0 = success).
 */
#ifndef GENEIG_MIXED_H
#define GENEIG_MIXED_H

#include <hip/hip_complex.h>
#include <hip/hip_runtime_api.h>

#include "geneig_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

int geneig_mixed_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                                    hipDoubleComplex* d_U, int* d_info, int batch_size, int max_iter,
                                    double tol, int* d_iters, int* h_fallback, hipStream_t stream);

/* The precision form for the pencil (eta, S) of linear spin-wave theory,
 * A = eta = diag(I_27, -I_27) (not passed, not read) and S = H(q), Hermitian
 * positive definite, n = 54: what Sunny's hegvd_batched! is handed
 * (benchmarks/sw08_baseline.jl:163-174).  The complex64 reduction and the FP64
 * fallback form eta in registers, and the refinement scales X by eta instead
 * of a matrix-core pass for A X; the arithmetic is otherwise the same, so W, U
 * and info equal geneig_mixed_n54_batched_launch's on (eta, S).  No reference
 * counterpart (the reference has no precision form). */
int geneig_mixed_eta_n54_batched_launch(const hipDoubleComplex* d_S, double* d_W, hipDoubleComplex* d_U,
                                        int* d_info, int batch_size, int max_iter, double tol, int* d_iters,
                                        int* h_fallback, hipStream_t stream);

/* The same pencil when S has the block form of a spin-wave Hamiltonian,
 * S = [[P, Q], [conj(Q), conj(P)]] (27 x 27 blocks, as Sunny's H(q) has):
 * its eigenpairs then come in pairs (lambda, x), (-lambda, Sigma conj(x)),
 * Sigma swapping the two halves of x, and the FP64 refinement works on the 27
 * positive ones only (half the matrix-core products per pair, twice the
 * pairs per CU) and writes their partners.  Every pair the refinement
 * finishes has the pairing exactly: W[53 - i] = -W[i], X(:, i) =
 * Sigma conj(X(:, 53 - i)); pairs it hands back are solved by the FP64 path
 * as there, and carry the pairing to rounding and phase.  The outputs meet
 * the same bars as geneig_mixed_eta_n54_batched_launch's, from which they
 * differ by rounding.  The block form is not checked: an S without it gets
 * wrong results.  No reference counterpart. */
int geneig_mixed_eta_paired_n54_batched_launch(const hipDoubleComplex* d_S, double* d_W, hipDoubleComplex* d_U,
                                               int* d_info, int batch_size, int max_iter, double tol,
                                               int* d_iters, int* h_fallback, hipStream_t stream);

/* Precision form for any n (1..100): a runtime solver (geneig_solver.h) whose
 * hipRTC module holds the precision-form kernels specialized for n (cached
 * like the FP64 module, variant 1 in the cache key).  geneig_solver_launch on
 * it runs the precision form with the default controls; this launch exposes
 * them.  Destroy with geneig_solver_destroy. */
int geneig_mixed_solver_create(geneig_solver_t* out, int n, int device_id, int cache_mode);
/* The same with the threads per pair of its Cholesky, K1 and K3 kernels
 * chosen: 64 (one wave per pair) or 256 (four waves per pair, the matrix
 * spread over the workgroup's registers); 0 picks 64 up to n = 64 and 256
 * above.  No reference counterpart (the reference has no precision form). */
int geneig_mixed_solver_create_bdx(geneig_solver_t* out, int n, int device_id, int block_dim_x, int cache_mode);
int geneig_mixed_solver_launch(geneig_solver_t s, const void* d_H, const void* d_S, double* d_W, void* d_U,
                               int* d_info, int batch_size, int max_iter, double tol, int* d_iters,
                               int* h_fallback, hipStream_t stream);

/* Default refinement controls used when the launch is given max_iter <= 0 / tol <= 0. */
int geneig_mixed_default_max_iter(void);
double geneig_mixed_default_tol(void);

#ifdef __cplusplus
}
#endif

#endif /* GENEIG_MIXED_H */
