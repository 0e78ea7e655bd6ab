/*
 * geneig_oracle.h -- CPU oracle for the batched complex Hermitian-definite
 * generalized eigensolve  A x = lambda B x  (B Hermitian positive definite).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ (to check results), by
 * __graft_entry__.smoke() (to check its result) and by bench.py's cpu_baseline leg.
 *
 * What it restates: the reference pipeline of
 *   src/nvrtc/kernel_source.hpp:240-248 (and src/geneig_fixed.cu:455-464):
 *     potrf(B) -> L^{-1} A -> (.) L^{-H} -> heev -> L^{-H} V
 * with the heev step written out as Householder tridiagonalization (lower)
 * followed by implicit-shift QL on the real tridiagonal and an ascending sort,
 * i.e. LAPACK zhegv(itype=1, 'V', 'L') semantics, which is what the
 * reference's tests hold it to (tests/test_full_batched_n54.cpp:253-260).
 *
 * Matrices are column-major, leading dimension n, complex128 (C99 double
 * complex, layout-identical to hipDoubleComplex / std::complex<double>).
 *
 * Parity pinning of each function (the checks live in tests/):
 *   go_zpotrf_lower   vs LAPACK zpotrf (tests/test_chol_n54.cpp:141-165 rule),
 *                     L L^H = B, B = I -> L = I.
 *   go_reduce         vs LAPACK zpotrf+zhegst lower triangle
 *                     (tests/test_reduce_n54.cpp:127-160 rule), eig(C)=eig(A,B).
 *   go_heev           vs LAPACK zheev eigenvalues/phases
 *                     (tests/test_reduce_eig_n54.cpp:144-202 rule), diagonal
 *                     input -> sorted diagonal, 2x2 closed form.
 *   go_zhegv          vs the reference's GEIG v1 fixtures (zhegvd outputs,
 *                     tests/data/n*_b8_s42_{random,nearid}.bin) and LAPACK zhegv; invariants
 *                     A = alpha B, scaling, unitary congruence, +- pairing of
 *                     spin-wave pairs.
 *   go_compare / go_residuals: brute-force definitions, pinned by planted
 *                     errors in tests/test_oracle.py.
 *   generators:       counter-based RNG pinned by SplitMix64 known-answer
 *                     values and by statistical moments; spin-wave pairs
 *                     pinned by Hermiticity, lambda_min(H) = target shift.
 */
#ifndef GENEIG_ORACLE_H
#define GENEIG_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex go_cplx;

/* ---- counter-based RNG (SplitMix64 finalizer + Box-Muller) -------------- */
uint64_t go_splitmix64(uint64_t x);
uint64_t go_rng_key(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx);
double   go_u01(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx);
void     go_cnormal(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx,
                    double* re, double* im);

/* ---- synthetic pair families -------------------------------------------- */
/* Generic pair: A = (M + M^H) / (2 sqrt n), B = G G^H / n + I. */
void go_gen_generic(int n, int batch, uint64_t seed, uint64_t first_pair,
                    go_cplx* A, go_cplx* B);
/* Generic pair with a prescribed cond(B): B = Q diag(logspace(0, log10 cond)) Q^H,
 * Q unitary from a QR (Gram-Schmidt) of a complex Gaussian matrix. */
void go_gen_generic_cond(int n, int batch, uint64_t seed, uint64_t first_pair,
                         double cond, go_cplx* A, go_cplx* B);

typedef struct { int shift; int r[3]; double J; } go_bond;
void go_spinwave_bonds(int nsub, int nbonds, uint64_t seed, go_bond* out);
void go_spinwave_q(int g0, int g1, int g2, uint64_t qidx, double q[3]);
/* H0(q) = [[Aq, Bq], [conj Bq, conj Aq]] scaled by jscale (n = 2 nsub); go_gen_spinwave
 * draws a scale per q, s = 0.8 + 0.4 U (stream 8), when jscale < 0. */
void go_spinwave_H0(int nsub, int nbonds, const go_bond* bonds, uint64_t seed,
                    double jscale, const double q[3], go_cplx* H);
double go_spinwave_J(uint64_t seed, int b, int i);
/* Shift target t = 10^U(-3,-1) for grid point qidx. */
double go_spinwave_target(uint64_t seed, uint64_t qidx);
/* Full pair: A = eta = diag(I, -I), B = H0 + (t - lambda_min(H0)) I. */
int go_gen_spinwave(int nsub, int nbonds, uint64_t seed, int g0, int g1, int g2,
                    uint64_t first_q, int batch, double jscale,
                    go_cplx* A, go_cplx* B);

/* ---- solver restatement -------------------------------------------------- */
int  go_zpotrf_lower(int n, const go_cplx* B, go_cplx* L);
void go_reduce(int n, const go_cplx* A, const go_cplx* L, go_cplx* C);
int  go_heev(int n, const go_cplx* C, double* W, go_cplx* V);
int  go_heevals(int n, const go_cplx* C, double* W);
void go_backtransform(int n, const go_cplx* L, const go_cplx* V, go_cplx* X);
int  go_zhegv(int n, const go_cplx* A, const go_cplx* B, double* W, go_cplx* X);
/* Batched driver, OpenMP over pairs; returns number of pairs with info != 0. */
int  go_zhegv_batch(int n, int batch, const go_cplx* A, const go_cplx* B,
                    double* W, go_cplx* X, int* info, int nthreads);

/* ---- LAPACK loaded at run time (MKL or scipy's OpenBLAS) ---------------- */
int         go_lapack_load(const char* path);   /* 0 on success */
const char* go_lapack_name(void);
int  go_lapack_zhegv(int n, const go_cplx* A, const go_cplx* B, double* W, go_cplx* X);
int  go_lapack_zhegv_batch(int n, int batch, const go_cplx* A, const go_cplx* B,
                           double* W, go_cplx* X, int* info, int nthreads);
int  go_lapack_zpotrf(int n, go_cplx* a);
int  go_lapack_zhegst(int n, go_cplx* a, const go_cplx* l);
int  go_lapack_zheev(int n, go_cplx* a, double* w);

/* ---- metrics ------------------------------------------------------------ */
typedef struct {
    double eig_rel_scaled;  /* max |W-Wr| / max(|Wr|,1)   (test_full_n54.cpp:147-148) */
    double eig_rel;         /* max |W-Wr| / |Wr|                                     */
    double phase_err;       /* max 1-|x^H B xr| over gap >= degen_gap (:160-174)     */
    double phase_err_degen; /* same over the excluded near-degenerate columns        */
    int    degen_excluded;
    double resid;           /* max_j ||A x-w B x|| / ((||A||+|w| ||B||) ||x||)       */
    double orth;            /* max_ij |x_i^H B x_j - d_ij| / (||B|| ||x_i|| ||x_j||)  */
    double resid_raw;       /* ||A X - B X diag(W)||_1 / (||A||_1 ||X||_1)           */
    double orth_raw;        /* ||X^H B X - I||_1                                     */
} go_metrics;
void go_residuals(int n, const go_cplx* A, const go_cplx* B, const double* W,
                  const go_cplx* X, double* resid, double* orth);
void go_residuals4(int n, const go_cplx* A, const go_cplx* B, const double* W,
                   const go_cplx* X, double out[4]);
void go_compare(int n, const go_cplx* A, const go_cplx* B,
                const double* W, const go_cplx* X,
                const double* Wr, const go_cplx* Xr, double degen_gap,
                go_metrics* out);
/* Batched compare, max over pairs; Xr may be NULL (then no phase metric). */
void go_compare_batch(int n, int batch, const go_cplx* A, const go_cplx* B,
                      const double* W, const go_cplx* X,
                      const double* Wr, const go_cplx* Xr, double degen_gap,
                      int nthreads, go_metrics* out);

/* ---- GEIG v1 fixture format (tests/lapack_reference.cpp:256-283) --------- */
int go_geig_read_header(const char* path, int32_t hdr[6]);
int go_geig_read(const char* path, int n, int batch, go_cplx* H, go_cplx* S,
                 double* W, go_cplx* U);
int go_geig_write(const char* path, int n, int batch, int seed,
                  const go_cplx* H, const go_cplx* S, const double* W,
                  const go_cplx* U);

/* FNV-1a 64 (known-answer anchor for the module-cache key hashing). */
uint64_t go_fnv1a64(const void* data, uint64_t len, uint64_t seed);

/* Long-double Rayleigh quotients of every column of X (batch x n out): the
 * batched form of oracle.rayleigh_ld. */
void go_rayleigh_ld_batch(int n, int batch, const go_cplx* A, const go_cplx* B, const go_cplx* X, double* out,
                          int nthreads);

#ifdef __cplusplus
}
#endif
#endif
