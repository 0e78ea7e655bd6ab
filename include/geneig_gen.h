/*
 * geneig_gen.h -- on-device synthetic pair generators (benchmark and test
 * inputs).  No counterpart in the reference, whose generators are host code
 * (tests/lapack_reference.cpp:86-150); these build the configured families in
 * HBM so a multi-GB batch never crosses PCIe.  Layout as in geneig.h.
 * Each returns a hipError_t value.
 */
#ifndef GENEIG_GEN_H
#define GENEIG_GEN_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A = (M + M^H) / (2 sqrt n), B = G G^H / n + I, M and G complex Gaussian
 * (stream 0 and 1 of the counter RNG, pair index first_pair + b); n <= 100. */
int geneig_gen_generic_launch(int n, int batch, uint64_t seed, uint64_t first_pair, void* d_A,
                              void* d_B, hipStream_t stream);

/* A = eta = diag(I_nsub, -I_nsub), B = H0(q) for q on a g0 x g1 x g2 grid,
 * q index first_q + b.  H0 = [[Aq, Bq], [conj Bq, conj Aq]]: bond b joins
 * site i to site (i + shift_b) mod nsub through the integer offset r_b with
 * coupling J_{b,i} ~ U(0.5, 1.5) * jscale (jscale < 0: * s per q, s = 0.8 + 0.4 U); Aq gets J/2 e^{+-i theta} (Hermitian)
 * plus J on the diagonal, Bq gets J/2 e^{i theta} on (i,j) and (j,i).
 * Finish with geneig_spinwave_shift_launch once lambda_min(H0) is known. */
int geneig_gen_spinwave_h0_launch(int nsub, int nbonds, uint64_t seed, int g0, int g1, int g2,
                                  uint64_t first_q, int batch, double jscale, void* d_A, void* d_B,
                                  hipStream_t stream);

/* B_ii += t_b - W[b*n] with t_b = 10^U(-3,-1) (stream 7 of the RNG), where
 * W[b*n] is lambda_min(H0) of pair b: sets lambda_min(B) = t_b. */
int geneig_spinwave_shift_launch(int n, int batch, uint64_t seed, uint64_t first_q, const double* d_W,
                                 void* d_B, hipStream_t stream);

/* B = I for every pair (used to get lambda(H0) from the (H0, I) problem). */
/* Per-pair parameters (q0, q1, q2, s, t) of the spin-wave family as float32
 * rows of 5: the grid point, the coupling scale (jscale, or per q when
 * jscale < 0) and the lambda_min(H) target. */
int geneig_spinwave_params_launch(uint64_t seed, int g0, int g1, int g2, uint64_t first_q, int batch,
                                  double jscale, float* d_params, hipStream_t stream);
int geneig_fill_identity_launch(int n, int batch, void* d_B, hipStream_t stream);

/* The spin-wave family's model as the tables geneig_spinwave_n54_launch
 * (geneig.h) takes: d_shift[b] and d_r[3b .. 3b+2] of each of the nbonds
 * bonds, and d_J[b nsub + i] = jscale (0.5 + U) the coupling of bond b at
 * site i (jscale >= 0; the draws of geneig_gen_spinwave_h0_launch). */
int geneig_spinwave_model_launch(int nsub, int nbonds, uint64_t seed, double jscale, int* d_shift, int* d_r,
                                 double* d_J, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* GENEIG_GEN_H */
