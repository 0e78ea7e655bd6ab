/* geneig_learn.h -- learned warm start for the n = 54 spin-wave family
 * (gfx950).
 *
 * No counterpart in the reference.  A multilayer perceptron (8 features ->
 * 256 -> 256 -> 256 tanh -> 2*54*54 outputs, FP32, row-major weights
 * W_l[out][in] and biases b_l[out]) maps the parameters of a pair (wavevector
 * q, coupling scale s, lambda_min target t) to a complex 54 x 54 matrix Z whose
 * Cholesky-QR Q is an approximate eigenbasis of the reduced matrix
 * C = L^-1 A L^-H.  Training and inference run on these entry points; device
 * buffers are caller-owned, weights are passed as an array of 8 device
 * pointers {W1, b1, W2, b2, W3, b3, W4, b4}.  All return a hipError_t value.
 */
#ifndef GENEIG_LEARN_H
#define GENEIG_LEARN_H

#include <hip/hip_complex.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

int geneig_learn_num_features(void);   /* 8 */
int geneig_learn_hidden(void);         /* 256 */
int geneig_learn_num_outputs(void);    /* 5832 = 2 * 54 * 54 */

/* Features from per-pair parameters (q0, q1, q2, s, t) (geneig_spinwave_params_launch):
 * cos/sin(2 pi q_k), 5 (s - 1), log10 t + 2. */
int geneig_learn_features(const float* d_params, float* d_X, int batch, hipStream_t stream);

/* C = op(A) op(B) (+ bias[col]) (tanh when act == 1), row-major FP32;
 * op(A) is M x K (A stored K x M when trans_a), op(B) is K x N (B stored N x K
 * when trans_b). */
int geneig_learn_gemm(int trans_a, int trans_b, int M, int N, int K, const float* A, int lda, const float* B,
                      int ldb, float* C, int ldc, const float* bias, int act, hipStream_t stream);
/* dY <- dY * (1 - Y^2) over n elements. */
int geneig_learn_tanh_backward(const float* d_Y, float* d_dY, long long n, hipStream_t stream);
/* db[c] = sum_r dY[r][c]. */
int geneig_learn_colsum(const float* d_dY, int rows, int cols, float* d_db, hipStream_t stream);
/* Adam step `step` (1-based) on n parameters. */
int geneig_learn_adam(float* d_p, const float* d_g, float* d_m, float* d_v, long long n, float lr, float beta1,
                      float beta2, float eps, int step, hipStream_t stream);
/* Network forward: Z (batch x 5832) from X (batch x 8); d_H receives the
 * three hidden activations (3 x batch x 256). */
int geneig_learn_mlp_forward(const float* const* d_weights, const float* d_X, int batch, float* d_H, float* d_Z,
                             hipStream_t stream);
/* Per sample b: Q = Z_b R^-1 (R^H R = Z_b^H Z_b), loss[b] = ||off(Q^H C Q)||_F^2 /
 * ||C||_F^2 with C = d_C[cidx ? cidx[b] : b] (complex64, 54 x 54 column-major,
 * full Hermitian), and dZ_b = scale * dloss[b]/dZ_b.  Z(i, j) is
 * (Z_b[2 (i + 54 j)], Z_b[2 (i + 54 j) + 1]).  A rank-deficient Z gives NaN loss
 * and zero gradient. */
int geneig_learn_loss(const float* d_Z, const void* d_C, const int* d_cidx, int batch, float scale, float* d_loss,
                      float* d_dZ, hipStream_t stream);
/* The solver's reduced matrices: FP64 Cholesky of S (L into d_L, zhegv info),
 * then C = L^-1 H L^-H in complex64 into d_C (batch x 54 x 54, full Hermitian). */
int geneig_learn_reduce_n54(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, hipDoubleComplex* d_L,
                            int* d_info, void* d_C, int batch, hipStream_t stream);

/* Learned solve of n = 54 pairs, same outputs as geneig_full_n54_batched_launch:
 * FP64 Cholesky, complex64 C, network forward on d_features (batch x 8; NULL:
 * start Jacobi from Q = I), Cholesky-QR, at most max_sweeps FP32 cyclic Jacobi
 * sweeps on Q^H C Q (until off-norm <= jacobi_tol ||C||_F), back-transform, FP64
 * refinement (max_iter, tol as geneig_mixed.h) and FP64 re-solve of the pairs it
 * hands back.  d_sweeps / d_iters (nullable, device) receive the Jacobi sweeps and
 * refinement steps per pair; *h_fallback the re-solved count.  Synchronizes
 * `stream` once. */
int geneig_learned_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                                      hipDoubleComplex* d_U, int* d_info, int batch, const float* d_features,
                                      const float* const* d_weights, int max_sweeps, float jacobi_tol,
                                      int max_iter, double tol, int* d_sweeps, int* d_iters, int* h_fallback,
                                      hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GENEIG_LEARN_H */
