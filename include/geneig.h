/*
 * geneig.h -- C ABI of the MI355X (gfx950) batched complex Hermitian-definite
 * generalized eigensolver  A x = lambda B x  (B Hermitian positive definite).
 *
 * Data contract (identical to the reference, src/nvrtc/kernel_source.hpp:200-207):
 *   A (H), B (S)  batch x n x n complex128, column-major, ld = n, pair stride n^2,
 *                 stored as full Hermitian matrices (the lower triangle of B and
 *                 the full A are read).
 *   W             batch x n float64, ascending eigenvalues, stride n.
 *   U (X)         batch x n x n complex128, B-orthonormal eigenvectors (U^H B U = I),
 *                 arbitrary column phase.  U must not alias A or B.
 *   info          batch int32: 0 ok; n + j when the leading minor of order j of B
 *                 is not positive definite (W, U of that pair are NaN); k > 0
 *                 when k off-diagonals of the tridiagonal did not converge
 *                 (LAPACK zhegv numbering).  Written for every pair, so the
 *                 caller no longer has to zero it.
 * Every entry point that launches work returns a hipError_t value (0 = success).
 * batch <= 0 is a no-op returning 0.  Device pointers only.
 *
 * Workspace: the solve stages hand their intermediates (tridiagonal, Q,
 * rotation log, Z) over in a device workspace the library keeps per
 * (device, stream): about 120 KB per pair at n = 54 (FP64), bounded by a
 * 16 GiB budget (GENEIG_WORKSPACE_MB overrides; larger batches run in
 * chunks).  It is allocated by the first launch on a stream that needs it and
 * grown by a later, larger one; geneig_n54_workspace_bytes gives the size,
 * geneig_n54_reserve allocates it ahead of time (a launch after it allocates
 * nothing and can be captured into a graph), geneig_workspace_release frees
 * it.  The reference's kernel used shared memory only and had no workspace.
 */
#ifndef GENEIG_H
#define GENEIG_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>
#include <hip/hip_complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================== */
/* Static n = 54 entry points.  Replaces src/geneig_fixed.cu.               */
/* ======================================================================== */

/* Single-pair Cholesky B = L L^H, L lower with zeroed strict upper triangle.
 * info uses zpotrf numbering (j when minor j is not PD).
 * Replaces chol_n54_launch, src/geneig_fixed.cu:134-147. */
int chol_n54_launch(const hipDoubleComplex* d_S, hipDoubleComplex* d_L, int* d_info,
                    hipStream_t stream);
unsigned int chol_n54_block_dim_x(void);        /* geneig_fixed.cu:150 */
unsigned int chol_n54_shared_mem_bytes(void);   /* geneig_fixed.cu:151 */

/* Single-pair reduction M = L^{-1} H L^{-H}; M lower, strict upper zeroed.
 * Replaces geneig_reduce_n54_launch, src/geneig_fixed.cu:208-222. */
int geneig_reduce_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                             hipDoubleComplex* d_M, int* d_info, hipStream_t stream);
unsigned int geneig_reduce_n54_block_dim_x(void);       /* geneig_fixed.cu:225 */
unsigned int geneig_reduce_n54_shared_mem_bytes(void);  /* geneig_fixed.cu:226 */

/* Single-pair reduction + Hermitian eigendecomposition: W and the
 * eigenvectors V of M (not back-transformed).
 * Replaces geneig_reduce_eig_n54_launch, src/geneig_fixed.cu:302-317. */
int geneig_reduce_eig_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                                 double* d_W, hipDoubleComplex* d_V, int* d_info,
                                 hipStream_t stream);
unsigned int geneig_reduce_eig_n54_block_dim_x(void);       /* geneig_fixed.cu:320 */
unsigned int geneig_reduce_eig_n54_shared_mem_bytes(void);  /* geneig_fixed.cu:321 */
unsigned int geneig_reduce_eig_n54_chol_smem_bytes(void);   /* geneig_fixed.cu:322 */
unsigned int geneig_reduce_eig_n54_heev_smem_bytes(void);   /* geneig_fixed.cu:323 */
unsigned int geneig_reduce_eig_n54_heev_workspace(void);    /* geneig_fixed.cu:324 */

/* Single-pair full generalized solve.
 * Replaces geneig_full_n54_launch, src/geneig_fixed.cu:385-400. */
int geneig_full_n54_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                           double* d_W, hipDoubleComplex* d_U, int* d_info,
                           hipStream_t stream);
unsigned int geneig_full_n54_block_dim_x(void);       /* geneig_fixed.cu:403 */
unsigned int geneig_full_n54_shared_mem_bytes(void);  /* geneig_fixed.cu:404 */

/* THE DROP-IN: batched full generalized solve over strided pairs.
 * Replaces geneig_full_n54_batched_launch, src/geneig_fixed.cu:474-492. */
int geneig_full_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                                   double* d_W, hipDoubleComplex* d_U, int* d_info,
                                   int batch_size, hipStream_t stream);
unsigned int geneig_full_n54_batched_block_dim_x(void);       /* geneig_fixed.cu:495 */
unsigned int geneig_full_n54_batched_shared_mem_bytes(void);  /* geneig_fixed.cu:496 */

/* In-place batched solve with LAPACK zhegvd's storage, for callers of
 * Sunny's hegvd_batched! (the Julia FFI, RISKS.md:23-34): on return d_H holds
 * the B-orthonormal eigenvectors (what d_U holds after the launch above) and
 * the lower triangle of d_S the Cholesky factor L (S = L L^H); W and info as
 * above.  When B is not positive definite, info = n + j and the pair's
 * d_H and W are NaN.  No separate reference entry point: the reference's
 * batched launcher (geneig_fixed.cu:474-492) writes a separate U. */
int geneig_full_n54_batched_inplace_launch(hipDoubleComplex* d_H, hipDoubleComplex* d_S, double* d_W,
                                           int* d_info, int batch_size, hipStream_t stream);

/* Spin-wave pairs built inside the solve.  The pair of q = d_q[3b .. 3b+2] is
 * (eta, H(q) + d_diag[b] I) with eta = diag(I_27, -I_27) and
 *   H(q) = [[Aq, Bq], [conj Bq, conj Aq]]:  bond k joins site i to site
 *   t = (i + d_shift[k]) mod 27 with coupling J = d_J[27 k + i] and offset
 *   r = d_r[3k .. 3k+2]; with theta = 2 pi q.r, it adds J/2 e^{i theta} to
 *   Aq(i, t) and Bq(i, t), J/2 e^{-i theta} to Aq(t, i), J/2 e^{i theta} to
 *   Bq(t, i), and J to Aq(i, i)  (d_diag may be NULL: no shift).
 * The FP64 Cholesky stage builds each column of H(q) in registers from these
 * tables and K1 builds eta, so neither matrix of the pair is ever stored in
 * or read from device memory; outputs W, U, info as
 * geneig_full_n54_batched_launch.  The spin-wave consumer of the reference
 * (benchmarks/sw08_baseline.jl:163-174) builds H(q) on the host and passes it
 * whole; this entry has no reference counterpart. */
int geneig_spinwave_n54_launch(int nbonds, const int* d_shift, const int* d_r, const double* d_J,
                               const double* d_q, const double* d_diag, double* d_W, hipDoubleComplex* d_U,
                               int* d_info, int batch_size, hipStream_t stream);
/* The spin-wave pencil (eta, S), eta = diag(I_27, -I_27) (not passed), for S
 * of the block form of a spin-wave Hamiltonian, S = [[P, Q], [conj Q, conj P]]
 * (27 x 27 blocks, as H(q) is), solved in real FP64 arithmetic: with
 * T = [[I, iI], [I, -iI]] / sqrt 2, T^H S T is real symmetric and
 * T^H eta T = i [[0, I], [-I, 0]], so the Cholesky factor, the reduction, the
 * Householder tridiagonalization and the back-transform are real (about a
 * quarter of the complex pipeline's multiplications) and the FP64 QL stages
 * run on a real tridiagonal with zero diagonal.  The 27 positive eigenpairs
 * are taken from it and their partners written exactly: W[53 - i] = -W[i],
 * X(:, i) = Sigma conj X(:, 53 - i), Sigma swapping the halves.  Outputs W, U
 * as geneig_full_n54_batched_launch's on (eta, S) to FP64 rounding; only the
 * lower triangle of S is read.  info = 54 + k when the leading minor of
 * order k of T^H S T (not of S) is not positive definite, with NaN outputs
 * for that pair.  The block form is not checked: an S without it gets wrong
 * results.  The FP64 bars hold up to cond(S) = 1e8 (tested).  No reference
 * counterpart (the reference solves the complex pencil as given). */
int geneig_eta_paired_n54_batched_launch(const hipDoubleComplex* d_S, double* d_W, hipDoubleComplex* d_U,
                                         int* d_info, int batch_size, hipStream_t stream);
/* The H(q) + diag I the launch above builds, written to d_H (batch x 54 x 54,
 * column-major): for checking the construction, or for callers that need the
 * matrices themselves. */
int geneig_spinwave_n54_build(int nbonds, const int* d_shift, const int* d_r, const double* d_J,
                              const double* d_q, const double* d_diag, hipDoubleComplex* d_H, int batch_size,
                              hipStream_t stream);

/* Batched finish from a given factor L (lower, zero upper) and reduced M:
 * heev(M) then U = L^{-H} V.
 * Replaces geneig_finish_n54_launch, src/geneig_fixed.cu:569-587. */
int geneig_finish_n54_launch(const hipDoubleComplex* d_L, const hipDoubleComplex* d_M,
                             double* d_W, hipDoubleComplex* d_U, int* d_info,
                             int batch_size, hipStream_t stream);
unsigned int geneig_finish_n54_block_dim_x(void);       /* geneig_fixed.cu:590 */
unsigned int geneig_finish_n54_shared_mem_bytes(void);  /* geneig_fixed.cu:591 */

/* ======================================================================== */
/* Workspace control.  No counterpart in the reference (whose kernels used  */
/* shared memory only); callers that capture launches into graphs or must   */
/* bound device memory use these.                                           */
/* ======================================================================== */

/* Workspace bytes a batched n = 54 launch of batch_size pairs uses on its
 * stream: form 0 geneig_full_n54_batched_launch, form 1
 * geneig_mixed_n54_batched_launch (geneig_mixed.h). */
size_t geneig_n54_workspace_bytes(int batch_size, int form);
/* Allocates that workspace on `stream` (and the internal streams of the
 * concurrent parts) ahead of time.  A launch of at most batch_size pairs on
 * `stream` afterwards allocates nothing; growing the workspace while the
 * stream is being captured is refused (hipErrorStreamCaptureUnsupported). */
int geneig_n54_reserve(int batch_size, int form, hipStream_t stream);
/* Frees the workspace held for `stream` on the current device, after waiting
 * for the stream (also the buffers a captured stream outgrew, which the
 * library keeps while a graph may use them).  *freed_bytes (nullable)
 * receives the bytes freed.  The caller must not replay graphs captured on
 * `stream` afterwards.  Works for the workspaces of runtime solvers too. */
int geneig_workspace_release(hipStream_t stream, size_t* freed_bytes);

/* ======================================================================== */
/* Allocation-free batched solve.  No counterpart in the reference: for     */
/* callers that cannot allocate device memory (code under graph capture,    */
/* tight serving loops, harnesses that forbid hipMalloc).                   */
/* ======================================================================== */

/* geneig_full_n54_batched_launch with its stage workspace in a fixed device
 * array of the library (512 MiB per GPU, held from the library's first
 * launch on that GPU until the process ends) instead of the per-stream
 * workspace: no hipMalloc, hipFree, stream or event creation.  At most
 * geneig_full_n54_batched_noalloc_max_batch() pairs per call (4,485); a
 * larger batch returns hipErrorInvalidValue with the reason in
 * geneig_last_error().  Runs on `stream` alone (no concurrent parts).
 * Every call on one GPU shares the array: calls that overlap on one GPU
 * (other streams, other threads) race; calls ordered on one stream do not.
 * Outputs equal geneig_full_n54_batched_launch's bit for bit. */
int geneig_full_n54_batched_launch_noalloc(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S,
                                           double* d_W, hipDoubleComplex* d_U, int* d_info,
                                           int batch_size, hipStream_t stream);
int geneig_full_n54_batched_noalloc_max_batch(void);

#ifdef __cplusplus
}
#endif
#endif /* GENEIG_H */
