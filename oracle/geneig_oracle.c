/*
 * geneig_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see the header for
 * what may use it and how each function is pinned).
 *
 * Restates the reference pipeline src/nvrtc/kernel_source.hpp:240-248:
 *   Cholesky (:240) -> TrsmLeft L^{-1} H (:242) -> TrsmRight (.) L^{-H} (:244)
 *   -> Heev (:246) -> TrsmLeftConj L^{-H} V (:248),
 * with the eigen-decomposition written out from the textbook method
 * (Householder reduction to real tridiagonal form + implicit-shift QL with
 * Wilkinson shifts, Golub & Van Loan, Matrix Computations, 4th ed., 8.3)
 * because cuSolverDx's heev is closed.
 */
#include "geneig_oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ------------------------------------------------------------------------ */
/* RNG                                                                       */
/* ------------------------------------------------------------------------ */

uint64_t go_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Key layout: stream in bits 56..63, pair in bits 24..55, element in 0..23,
 * xor-ed with a hashed seed.  Distinct (stream, pair < 2^32, idx < 2^24)
 * give distinct counters. */
uint64_t go_rng_key(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx) {
    uint64_t s = go_splitmix64(seed ^ 0x5851F42D4C957F2DULL);
    return s ^ ((uint64_t)(stream & 0xFFu) << 56) ^ ((pair & 0xFFFFFFFFULL) << 24)
             ^ (uint64_t)(idx & 0xFFFFFFu);
}

double go_u01(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx) {
    uint64_t r = go_splitmix64(go_rng_key(seed, stream, pair, idx));
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* Box-Muller on the uniform pair (2 idx, 2 idx + 1). */
void go_cnormal(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx,
                double* re, double* im) {
    double u1 = go_u01(seed, stream, pair, 2u * idx);
    double u2 = go_u01(seed, stream, pair, 2u * idx + 1u);
    double rad = sqrt(-2.0 * log(u1));
    double ang = 6.283185307179586476925286766559 * u2;
    *re = rad * cos(ang);
    *im = rad * sin(ang);
}

/* ------------------------------------------------------------------------ */
/* Generators                                                                */
/* ------------------------------------------------------------------------ */

void go_gen_generic(int n, int batch, uint64_t seed, uint64_t first_pair,
                    go_cplx* A, go_cplx* B) {
    const size_t nn = (size_t)n * n;
    const double sa = 1.0 / (2.0 * sqrt((double)n));
    const double sb = 1.0 / (double)n;
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < batch; ++b) {
        uint64_t p = first_pair + (uint64_t)b;
        go_cplx* Ab = A + nn * b;
        go_cplx* Bb = B + nn * b;
        go_cplx* M = (go_cplx*)malloc(nn * sizeof(go_cplx));
        go_cplx* G = (go_cplx*)malloc(nn * sizeof(go_cplx));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double re, im;
                go_cnormal(seed, 0, p, (uint32_t)IDX(i, j, n), &re, &im);
                M[IDX(i, j, n)] = re + I * im;
                go_cnormal(seed, 1, p, (uint32_t)IDX(i, j, n), &re, &im);
                G[IDX(i, j, n)] = re + I * im;
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                Ab[IDX(i, j, n)] = (M[IDX(i, j, n)] + conj(M[IDX(j, i, n)])) * sa;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                go_cplx acc = 0;
                for (int k = 0; k < n; ++k) acc += G[IDX(i, k, n)] * conj(G[IDX(j, k, n)]);
                Bb[IDX(i, j, n)] = acc * sb + (i == j ? 1.0 : 0.0);
            }
        /* exact Hermitian storage: diagonal real, upper = conj(lower) */
        for (int j = 0; j < n; ++j) {
            Ab[IDX(j, j, n)] = creal(Ab[IDX(j, j, n)]);
            Bb[IDX(j, j, n)] = creal(Bb[IDX(j, j, n)]);
            for (int i = j + 1; i < n; ++i) {
                Ab[IDX(j, i, n)] = conj(Ab[IDX(i, j, n)]);
                Bb[IDX(j, i, n)] = conj(Bb[IDX(i, j, n)]);
            }
        }
        free(M);
        free(G);
    }
}

void go_gen_generic_cond(int n, int batch, uint64_t seed, uint64_t first_pair,
                         double cond, go_cplx* A, go_cplx* B) {
    const size_t nn = (size_t)n * n;
    go_gen_generic(n, batch, seed, first_pair, A, B);
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < batch; ++b) {
        uint64_t p = first_pair + (uint64_t)b;
        go_cplx* Bb = B + nn * b;
        go_cplx* Q = (go_cplx*)malloc(nn * sizeof(go_cplx));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double re, im;
                go_cnormal(seed, 2, p, (uint32_t)IDX(i, j, n), &re, &im);
                Q[IDX(i, j, n)] = re + I * im;
            }
        /* modified Gram-Schmidt, twice for orthogonality to working accuracy */
        for (int pass = 0; pass < 2; ++pass)
            for (int j = 0; j < n; ++j) {
                for (int k = 0; k < j; ++k) {
                    go_cplx d = 0;
                    for (int i = 0; i < n; ++i) d += conj(Q[IDX(i, k, n)]) * Q[IDX(i, j, n)];
                    for (int i = 0; i < n; ++i) Q[IDX(i, j, n)] -= d * Q[IDX(i, k, n)];
                }
                double nr = 0;
                for (int i = 0; i < n; ++i) nr += creal(Q[IDX(i, j, n)] * conj(Q[IDX(i, j, n)]));
                nr = 1.0 / sqrt(nr);
                for (int i = 0; i < n; ++i) Q[IDX(i, j, n)] *= nr;
            }
        for (int j = 0; j < n; ++j)
            for (int i = j; i < n; ++i) {
                go_cplx acc = 0;
                for (int k = 0; k < n; ++k) {
                    double ev = (n > 1) ? pow(cond, (double)k / (double)(n - 1)) : 1.0;
                    acc += Q[IDX(i, k, n)] * ev * conj(Q[IDX(j, k, n)]);
                }
                if (i == j) acc = creal(acc);
                Bb[IDX(i, j, n)] = acc;
                Bb[IDX(j, i, n)] = conj(acc);
            }
        free(Q);
    }
}

void go_spinwave_bonds(int nsub, int nbonds, uint64_t seed, go_bond* out) {
    for (int b = 0; b < nbonds; ++b) {
        double u = go_u01(seed, 5, 0, (uint32_t)(8 * b + 0));
        int s = 1 + (int)(u * (double)(nsub - 1));
        if (s > nsub - 1) s = nsub - 1;
        out[b].shift = s;
        for (int c = 0; c < 3; ++c) {
            double v = go_u01(seed, 5, 0, (uint32_t)(8 * b + 1 + c));
            int r = (int)(v * 3.0) - 1;
            if (r > 1) r = 1;
            out[b].r[c] = r;
        }
        out[b].J = 1.0;   /* per-site couplings come from go_spinwave_J */
    }
}

/* Coupling of bond b at site i: J ~ U(0.5, 1.5), drawn per (bond, site) so the
 * sublattice ring has no translation symmetry (no systematic degeneracy). */
double go_spinwave_J(uint64_t seed, int b, int i) {
    return 0.5 + go_u01(seed, 6, (uint64_t)b, (uint32_t)i);
}

void go_spinwave_q(int g0, int g1, int g2, uint64_t qidx, double q[3]) {
    uint64_t k = qidx % (uint64_t)g2;
    uint64_t j = (qidx / (uint64_t)g2) % (uint64_t)g1;
    uint64_t i = (qidx / ((uint64_t)g2 * (uint64_t)g1)) % (uint64_t)g0;
    q[0] = (double)i / (double)g0;
    q[1] = (double)j / (double)g1;
    q[2] = (double)k / (double)g2;
}

void go_spinwave_H0(int nsub, int nbonds, const go_bond* bonds, uint64_t seed,
                    double jscale, const double q[3], go_cplx* H) {
    const int N = nsub, n = 2 * nsub;
    go_cplx* Aq = (go_cplx*)calloc((size_t)N * N, sizeof(go_cplx));
    go_cplx* Bq = (go_cplx*)calloc((size_t)N * N, sizeof(go_cplx));
    for (int b = 0; b < nbonds; ++b) {
        const double th = 6.283185307179586476925286766559 *
            (q[0] * bonds[b].r[0] + q[1] * bonds[b].r[1] + q[2] * bonds[b].r[2]);
        const go_cplx ph = cos(th) + I * sin(th);
        for (int i = 0; i < N; ++i) {
            const double J = bonds[b].J * jscale * go_spinwave_J(seed, b, i);
            int j = (i + bonds[b].shift) % N;
            Aq[IDX(i, j, N)] += 0.5 * J * ph;
            Aq[IDX(j, i, N)] += 0.5 * J * conj(ph);
            Bq[IDX(i, j, N)] += 0.5 * J * ph;
            Bq[IDX(j, i, N)] += 0.5 * J * ph;
            Aq[IDX(i, i, N)] += J;
        }
    }
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            H[IDX(i, j, n)] = Aq[IDX(i, j, N)];
            H[IDX(i, j + N, n)] = Bq[IDX(i, j, N)];
            H[IDX(i + N, j, n)] = conj(Bq[IDX(i, j, N)]);
            H[IDX(i + N, j + N, n)] = conj(Aq[IDX(i, j, N)]);
        }
    for (int i = 0; i < n; ++i) H[IDX(i, i, n)] = creal(H[IDX(i, i, n)]);
    free(Aq);
    free(Bq);
}

double go_spinwave_target(uint64_t seed, uint64_t qidx) {
    double u = go_u01(seed, 7, qidx, 0);
    return pow(10.0, -3.0 + 2.0 * u);
}

int go_gen_spinwave(int nsub, int nbonds, uint64_t seed, int g0, int g1, int g2,
                    uint64_t first_q, int batch, double jscale,
                    go_cplx* A, go_cplx* B) {
    const int n = 2 * nsub;
    const size_t nn = (size_t)n * n;
    go_bond* bonds = (go_bond*)malloc(sizeof(go_bond) * (size_t)nbonds);
    go_spinwave_bonds(nsub, nbonds, seed, bonds);
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
    for (int b = 0; b < batch; ++b) {
        uint64_t qi = first_q + (uint64_t)b;
        double q[3];
        go_spinwave_q(g0, g1, g2, qi, q);
        go_cplx* Hb = B + nn * b;
        go_cplx* Ab = A + nn * b;
        /* jscale < 0: a coupling scale per q, s = 0.8 + 0.4 U (stream 8) */
        const double js = jscale < 0.0 ? 0.8 + 0.4 * go_u01(seed, 8, qi, 0) : jscale;
        go_spinwave_H0(nsub, nbonds, bonds, seed, js, q, Hb);
        double* w = (double*)malloc(sizeof(double) * (size_t)n);
        if (go_heevals(n, Hb, w) != 0) bad++;
        double shift = go_spinwave_target(seed, qi) - w[0];
        for (int i = 0; i < n; ++i) Hb[IDX(i, i, n)] += shift;
        free(w);
        memset(Ab, 0, nn * sizeof(go_cplx));
        for (int i = 0; i < n; ++i) Ab[IDX(i, i, n)] = (i < nsub) ? 1.0 : -1.0;
    }
    free(bonds);
    return bad;
}

/* ------------------------------------------------------------------------ */
/* Stage 1: Cholesky, lower (reference Cholesky potrf, kernel_source.hpp:240) */
/* Reads the lower triangle of B only (potrf lower).  L has zero upper part.  */
/* Returns 0, or j+1 when the leading minor of order j+1 is not PD (zpotrf).  */
/* ------------------------------------------------------------------------ */
int go_zpotrf_lower(int n, const go_cplx* B, go_cplx* L) {
    memset(L, 0, sizeof(go_cplx) * (size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) L[IDX(i, j, n)] = B[IDX(i, j, n)];
    for (int j = 0; j < n; ++j) {
        double ajj = creal(L[IDX(j, j, n)]);
        for (int k = 0; k < j; ++k) {
            go_cplx l = L[IDX(j, k, n)];
            ajj -= creal(l) * creal(l) + cimag(l) * cimag(l);
        }
        if (!(ajj > 0.0)) return j + 1;
        double ljj = sqrt(ajj);
        L[IDX(j, j, n)] = ljj;
        for (int i = j + 1; i < n; ++i) {
            go_cplx acc = L[IDX(i, j, n)];
            for (int k = 0; k < j; ++k) acc -= L[IDX(i, k, n)] * conj(L[IDX(j, k, n)]);
            L[IDX(i, j, n)] = acc / ljj;
        }
    }
    return 0;
}

/* Forward substitution on every column of X: X <- L^{-1} X.               */
static void trsm_lower_left(int n, const go_cplx* L, go_cplx* X) {
    for (int c = 0; c < n; ++c) {
        go_cplx* x = X + (size_t)c * n;
        for (int k = 0; k < n; ++k) {
            x[k] /= creal(L[IDX(k, k, n)]);
            for (int i = k + 1; i < n; ++i) x[i] -= L[IDX(i, k, n)] * x[k];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Stage 2: two-sided reduction C = L^{-1} A L^{-H}                          */
/* (TrsmLeft kernel_source.hpp:242 then TrsmRight :244).  The right solve is */
/* done as C = L^{-1} (L^{-1} A)^H, valid because C is Hermitian.  Output is  */
/* stored as an exactly Hermitian full matrix built from its lower triangle. */
/* ------------------------------------------------------------------------ */
void go_reduce(int n, const go_cplx* A, const go_cplx* L, go_cplx* C) {
    const size_t nn = (size_t)n * n;
    go_cplx* Wm = (go_cplx*)malloc(nn * sizeof(go_cplx));
    memcpy(Wm, A, nn * sizeof(go_cplx));
    trsm_lower_left(n, L, Wm);                   /* W = L^{-1} A */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) C[IDX(i, j, n)] = conj(Wm[IDX(j, i, n)]);
    trsm_lower_left(n, L, C);                    /* C = L^{-1} W^H */
    for (int j = 0; j < n; ++j) {
        C[IDX(j, j, n)] = creal(C[IDX(j, j, n)]);
        for (int i = j + 1; i < n; ++i) C[IDX(j, i, n)] = conj(C[IDX(i, j, n)]);
    }
    free(Wm);
}

/* ------------------------------------------------------------------------ */
/* Stage 3 (Heev, kernel_source.hpp:246): Householder tridiagonalization of  */
/* the lower triangle.  Reflector k: H_k = I - tau_k v_k v_k^H with          */
/* H_k^H (alpha; x) = (beta; 0), beta real (LAPACK zlarfg convention), so    */
/* the tridiagonal T = Q^H C Q is real with Q = H_0 H_1 ... H_{n-2}.         */
/* On return C holds v_k below the subdiagonal of column k (v_k[k+1] = 1 is  */
/* implicit), d the diagonal, e the subdiagonal (e[n-1] = 0).                */
/* ------------------------------------------------------------------------ */
static void hetrd_lower(int n, go_cplx* C, double* d, double* e, go_cplx* tau) {
    go_cplx* p = (go_cplx*)malloc(sizeof(go_cplx) * (size_t)n);
    for (int k = 0; k + 1 < n; ++k) {
        const int m = n - k - 1;            /* length of the column below the diagonal */
        go_cplx* x = C + IDX(k + 1, k, n);  /* x[0] = alpha */
        go_cplx alpha = x[0];
        double xnorm2 = 0.0;
        for (int i = 1; i < m; ++i) xnorm2 += creal(x[i] * conj(x[i]));
        go_cplx t;
        double beta;
        if (xnorm2 == 0.0 && cimag(alpha) == 0.0) {
            t = 0.0;
            beta = creal(alpha);
        } else {
            double nrm = sqrt(creal(alpha) * creal(alpha) + cimag(alpha) * cimag(alpha) + xnorm2);
            beta = (creal(alpha) >= 0.0) ? -nrm : nrm;
            t = (beta - creal(alpha)) / beta - I * (cimag(alpha) / beta);
            go_cplx sc = 1.0 / (alpha - beta);
            for (int i = 1; i < m; ++i) x[i] *= sc;
        }
        x[0] = 1.0;
        e[k] = beta;
        tau[k] = t;
        if (t != 0.0) {
            go_cplx* C22 = C + IDX(k + 1, k + 1, n);
            /* p = tau * C22 v (C22 Hermitian, full storage kept in sync) */
            for (int i = 0; i < m; ++i) {
                go_cplx acc = 0;
                for (int j = 0; j < m; ++j) acc += C22[IDX(i, j, n)] * x[j];
                p[i] = t * acc;
            }
            /* w = p - (tau/2) (p^H v) v */
            go_cplx pv = 0;
            for (int i = 0; i < m; ++i) pv += conj(p[i]) * x[i];
            go_cplx a2 = -0.5 * t * pv;
            for (int i = 0; i < m; ++i) p[i] += a2 * x[i];
            /* C22 -= v w^H + w v^H */
            for (int j = 0; j < m; ++j)
                for (int i = 0; i < m; ++i)
                    C22[IDX(i, j, n)] -= x[i] * conj(p[j]) + p[i] * conj(x[j]);
        }
        x[0] = beta;
    }
    for (int k = 0; k < n; ++k) d[k] = creal(C[IDX(k, k, n)]);
    e[n - 1] = 0.0;
    free(p);
}

/* Implicit-shift QL on the symmetric tridiagonal (d, e), e[i] couples i and */
/* i+1.  Rotations are accumulated into the columns of Z (n x n, ld n) when  */
/* Z != NULL.  Returns 0 or the number of unconverged off-diagonals.         */
static int tridiag_ql(int n, double* d, double* e, double* Z) {
    const int max_iter = 30 * n;
    int total = 0;
    for (int l = 0; l < n; ++l) {
        for (;;) {
            int m = l;
            while (m < n - 1) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= DBL_EPSILON * dd) break;
                ++m;
            }
            if (m == l) break;
            if (++total > max_iter) {
                int bad = 0;
                for (int i = 0; i < n - 1; ++i) bad += (e[i] != 0.0);
                return bad > 0 ? bad : 1;
            }
            /* Wilkinson shift from the leading 2x2 block at l */
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            int early = 0;
            for (i = m - 1; i >= l; --i) {
                double f = s * e[i];
                double b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {          /* underflow: split and restart */
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (Z) {
                    double* zi = Z + (size_t)i * n;
                    double* zj = Z + (size_t)(i + 1) * n;
                    for (int k = 0; k < n; ++k) {
                        double t = zj[k];
                        zj[k] = s * zi[k] + c * t;
                        zi[k] = c * zi[k] - s * t;
                    }
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

/* Ascending selection sort of d with matching column swaps of Z.            */
static void sort_eig(int n, double* d, double* Z) {
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            double t = d[i]; d[i] = d[k]; d[k] = t;
            if (Z)
                for (int r = 0; r < n; ++r) {
                    double z = Z[IDX(r, i, n)];
                    Z[IDX(r, i, n)] = Z[IDX(r, k, n)];
                    Z[IDX(r, k, n)] = z;
                }
        }
    }
}

int go_heev(int n, const go_cplx* Cin, double* W, go_cplx* V) {
    const size_t nn = (size_t)n * n;
    go_cplx* C = (go_cplx*)malloc(nn * sizeof(go_cplx));
    /* Hermitian input: use the lower triangle, mirror it (FillMode lower). */
    for (int j = 0; j < n; ++j) {
        C[IDX(j, j, n)] = creal(Cin[IDX(j, j, n)]);
        for (int i = j + 1; i < n; ++i) {
            C[IDX(i, j, n)] = Cin[IDX(i, j, n)];
            C[IDX(j, i, n)] = conj(Cin[IDX(i, j, n)]);
        }
    }
    double* e = (double*)malloc(sizeof(double) * (size_t)n);
    go_cplx* tau = (go_cplx*)malloc(sizeof(go_cplx) * (size_t)n);
    double* Z = (double*)calloc(nn, sizeof(double));
    hetrd_lower(n, C, W, e, tau);
    for (int i = 0; i < n; ++i) Z[IDX(i, i, n)] = 1.0;
    int info = tridiag_ql(n, W, e, Z);
    sort_eig(n, W, Z);
    /* V = Q Z = H_0 (H_1 (... H_{n-2} Z)) */
    for (size_t t = 0; t < nn; ++t) V[t] = Z[t];
    for (int k = n - 2; k >= 0; --k) {
        if (tau[k] == 0.0) continue;
        const int m = n - k - 1;
        go_cplx* v = C + IDX(k + 1, k, n);
        go_cplx v0 = v[0];
        v[0] = 1.0;
        for (int j = 0; j < n; ++j) {
            go_cplx* y = V + IDX(k + 1, j, n);
            go_cplx w = 0;
            for (int i = 0; i < m; ++i) w += conj(v[i]) * y[i];
            w *= tau[k];
            for (int i = 0; i < m; ++i) y[i] -= v[i] * w;
        }
        v[0] = v0;
    }
    free(C); free(e); free(tau); free(Z);
    return info;
}

int go_heevals(int n, const go_cplx* Cin, double* W) {
    const size_t nn = (size_t)n * n;
    go_cplx* C = (go_cplx*)malloc(nn * sizeof(go_cplx));
    for (int j = 0; j < n; ++j) {
        C[IDX(j, j, n)] = creal(Cin[IDX(j, j, n)]);
        for (int i = j + 1; i < n; ++i) {
            C[IDX(i, j, n)] = Cin[IDX(i, j, n)];
            C[IDX(j, i, n)] = conj(Cin[IDX(i, j, n)]);
        }
    }
    double* e = (double*)malloc(sizeof(double) * (size_t)n);
    go_cplx* tau = (go_cplx*)malloc(sizeof(go_cplx) * (size_t)n);
    hetrd_lower(n, C, W, e, tau);
    int info = tridiag_ql(n, W, e, NULL);
    sort_eig(n, W, NULL);
    free(C); free(e); free(tau);
    return info;
}

/* ------------------------------------------------------------------------ */
/* Stage 4 (TrsmLeftConj, kernel_source.hpp:248): X = L^{-H} V, backward     */
/* substitution with L^H, written in dot-product form over column i of L.    */
/* ------------------------------------------------------------------------ */
void go_backtransform(int n, const go_cplx* L, const go_cplx* V, go_cplx* X) {
    for (int c = 0; c < n; ++c) {
        const go_cplx* y = V + (size_t)c * n;
        go_cplx* x = X + (size_t)c * n;
        for (int i = n - 1; i >= 0; --i) {
            go_cplx acc = y[i];
            for (int k = i + 1; k < n; ++k) acc -= conj(L[IDX(k, i, n)]) * x[k];
            x[i] = acc / creal(L[IDX(i, i, n)]);
        }
    }
}

/* Full pipeline with LAPACK zhegv's info convention: n+i when the leading   */
/* minor of order i of B is not PD, i > 0 (count) when the QL fails.         */
int go_zhegv(int n, const go_cplx* A, const go_cplx* B, double* W, go_cplx* X) {
    const size_t nn = (size_t)n * n;
    go_cplx* L = (go_cplx*)malloc(nn * sizeof(go_cplx));
    go_cplx* C = (go_cplx*)malloc(nn * sizeof(go_cplx));
    go_cplx* V = (go_cplx*)malloc(nn * sizeof(go_cplx));
    int info = go_zpotrf_lower(n, B, L);
    if (info != 0) {
        for (int i = 0; i < n; ++i) W[i] = NAN;
        for (size_t t = 0; t < nn; ++t) X[t] = NAN;
        free(L); free(C); free(V);
        return n + info;
    }
    go_reduce(n, A, L, C);
    info = go_heev(n, C, W, V);
    go_backtransform(n, L, V, X);
    free(L); free(C); free(V);
    return info;
}

int go_zhegv_batch(int n, int batch, const go_cplx* A, const go_cplx* B,
                   double* W, go_cplx* X, int* info, int nthreads) {
    const size_t nn = (size_t)n * n;
    int bad = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(+ : bad)
#endif
    for (int b = 0; b < batch; ++b) {
        int r = go_zhegv(n, A + nn * b, B + nn * b, W + (size_t)n * b, X + nn * b);
        if (info) info[b] = r;
        bad += (r != 0);
    }
    (void)nthreads;
    return bad;
}

/* ------------------------------------------------------------------------ */
/* LAPACK at run time                                                        */
/* ------------------------------------------------------------------------ */
typedef void (*zhegv_fn)(const int*, const char*, const char*, const int*, go_cplx*,
                         const int*, go_cplx*, const int*, double*, go_cplx*, const int*,
                         double*, int*);
typedef void (*zpotrf_fn)(const char*, const int*, go_cplx*, const int*, int*);
typedef void (*zhegst_fn)(const int*, const char*, const int*, go_cplx*, const int*,
                          const go_cplx*, const int*, int*);
typedef void (*zheev_fn)(const char*, const char*, const int*, go_cplx*, const int*,
                         double*, go_cplx*, const int*, double*, int*);
typedef void (*setthreads_fn)(int);

static zhegv_fn p_zhegv;
static zpotrf_fn p_zpotrf;
static zhegst_fn p_zhegst;
static zheev_fn p_zheev;
static char g_lapack_name[512];

static void* sym2(void* h, const char* a, const char* b) {
    void* p = dlsym(h, a);
    return p ? p : dlsym(h, b);
}

int go_lapack_load(const char* path) {
    if (p_zhegv) return 0;
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    p_zhegv = (zhegv_fn)sym2(h, "zhegv_", "scipy_zhegv_");
    p_zpotrf = (zpotrf_fn)sym2(h, "zpotrf_", "scipy_zpotrf_");
    p_zhegst = (zhegst_fn)sym2(h, "zhegst_", "scipy_zhegst_");
    p_zheev = (zheev_fn)sym2(h, "zheev_", "scipy_zheev_");
    if (!p_zhegv || !p_zpotrf || !p_zhegst || !p_zheev) {
        p_zhegv = NULL;
        return 2;
    }
    /* one BLAS thread per call; batching is done with OpenMP over pairs */
    setthreads_fn st = (setthreads_fn)dlsym(h, "MKL_Set_Num_Threads");
    if (st) st(1);
    setthreads_fn st2 = (setthreads_fn)sym2(h, "openblas_set_num_threads", "scipy_openblas_set_num_threads64_");
    if (st2) st2(1);
    snprintf(g_lapack_name, sizeof g_lapack_name, "%s", path);
    return 0;
}

const char* go_lapack_name(void) { return g_lapack_name; }

int go_lapack_zhegv(int n, const go_cplx* A, const go_cplx* B, double* W, go_cplx* X) {
    if (!p_zhegv) return -100;
    const size_t nn = (size_t)n * n;
    go_cplx* Bw = (go_cplx*)malloc(nn * sizeof(go_cplx));
    memcpy(X, A, nn * sizeof(go_cplx));
    memcpy(Bw, B, nn * sizeof(go_cplx));
    int itype = 1, info = 0, lwork = 64 * n;
    go_cplx* work = (go_cplx*)malloc(sizeof(go_cplx) * (size_t)lwork);
    double* rwork = (double*)malloc(sizeof(double) * (size_t)(3 * n));
    p_zhegv(&itype, "V", "L", &n, X, &n, Bw, &n, W, work, &lwork, rwork, &info);
    free(Bw); free(work); free(rwork);
    return info;
}

int go_lapack_zhegv_batch(int n, int batch, const go_cplx* A, const go_cplx* B,
                          double* W, go_cplx* X, int* info, int nthreads) {
    const size_t nn = (size_t)n * n;
    int bad = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(+ : bad)
#endif
    for (int b = 0; b < batch; ++b) {
        int r = go_lapack_zhegv(n, A + nn * b, B + nn * b, W + (size_t)n * b, X + nn * b);
        if (info) info[b] = r;
        bad += (r != 0);
    }
    (void)nthreads;
    return bad;
}

int go_lapack_zpotrf(int n, go_cplx* a) {
    if (!p_zpotrf) return -100;
    int info = 0;
    p_zpotrf("L", &n, a, &n, &info);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) a[IDX(i, j, n)] = 0;
    return info;
}

int go_lapack_zhegst(int n, go_cplx* a, const go_cplx* l) {
    if (!p_zhegst) return -100;
    int itype = 1, info = 0;
    p_zhegst(&itype, "L", &n, a, &n, l, &n, &info);
    return info;
}

int go_lapack_zheev(int n, go_cplx* a, double* w) {
    if (!p_zheev) return -100;
    int info = 0, lwork = 64 * n;
    go_cplx* work = (go_cplx*)malloc(sizeof(go_cplx) * (size_t)lwork);
    double* rwork = (double*)malloc(sizeof(double) * (size_t)(3 * n));
    p_zheev("V", "L", &n, a, &n, w, work, &lwork, rwork, &info);
    free(work); free(rwork);
    return info;
}

/* ------------------------------------------------------------------------ */
/* Metrics                                                                   */
/* ------------------------------------------------------------------------ */
static double norm1(int n, const go_cplx* M) {
    double best = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += cabs(M[IDX(i, j, n)]);
        if (s > best) best = s;
    }
    return best;
}

static void matmul(int n, const go_cplx* A, const go_cplx* X, go_cplx* Y) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Y[IDX(i, j, n)] = 0;
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            go_cplx xkj = X[IDX(k, j, n)];
            for (int i = 0; i < n; ++i) Y[IDX(i, j, n)] += A[IDX(i, k, n)] * xkj;
        }
}

/* Scale-invariant backward-error style measures (the per-eigenpair form of
 * tests/lapack_reference.cpp:207-254):
 *   resid     = max_j ||A x_j - w_j B x_j||_2 / ((||A||_F + |w_j| ||B||_F) ||x_j||_2)
 *   orth      = max_ij |x_i^H B x_j - delta_ij| / (||B||_F ||x_i||_2 ||x_j||_2)
 * and the literal forms
 *   resid_raw = ||A X - B X diag(W)||_1 / (||A||_1 ||X||_1)
 *   orth_raw  = ||X^H B X - I||_1. */
static double fro(int n, const go_cplx* M) {
    double s = 0.0;
    for (size_t t = 0; t < (size_t)n * n; ++t) s += creal(M[t]) * creal(M[t]) + cimag(M[t]) * cimag(M[t]);
    return sqrt(s);
}

void go_residuals4(int n, const go_cplx* A, const go_cplx* B, const double* W,
                   const go_cplx* X, double out[4]) {
    const size_t nn = (size_t)n * n;
    go_cplx* AX = (go_cplx*)malloc(nn * sizeof(go_cplx));
    go_cplx* BX = (go_cplx*)malloc(nn * sizeof(go_cplx));
    go_cplx* R = (go_cplx*)malloc(nn * sizeof(go_cplx));
    double* xn = (double*)malloc(sizeof(double) * (size_t)n);
    matmul(n, A, X, AX);
    matmul(n, B, X, BX);
    const double fa = fro(n, A), fb = fro(n, B);
    double resid = 0.0;
    for (int j = 0; j < n; ++j) {
        double r2 = 0.0, x2 = 0.0;
        for (int i = 0; i < n; ++i) {
            go_cplx r = AX[IDX(i, j, n)] - BX[IDX(i, j, n)] * W[j];
            R[IDX(i, j, n)] = r;
            r2 += creal(r) * creal(r) + cimag(r) * cimag(r);
            go_cplx x = X[IDX(i, j, n)];
            x2 += creal(x) * creal(x) + cimag(x) * cimag(x);
        }
        xn[j] = sqrt(x2);
        double v = sqrt(r2) / ((fa + fabs(W[j]) * fb) * xn[j]);
        if (!(v == v)) v = INFINITY;
        if (v > resid) resid = v;
    }
    out[0] = resid;
    out[2] = norm1(n, R) / (norm1(n, A) * norm1(n, X));
    double orth = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            go_cplx acc = 0;
            for (int k = 0; k < n; ++k) acc += conj(X[IDX(k, i, n)]) * BX[IDX(k, j, n)];
            acc -= (i == j ? 1.0 : 0.0);
            R[IDX(i, j, n)] = acc;
            double v = cabs(acc) / (fb * xn[i] * xn[j]);
            if (!(v == v)) v = INFINITY;
            if (v > orth) orth = v;
        }
    out[1] = orth;
    out[3] = norm1(n, R);
    free(AX); free(BX); free(R); free(xn);
}

void go_residuals(int n, const go_cplx* A, const go_cplx* B, const double* W,
                  const go_cplx* X, double* resid, double* orth) {
    double o[4];
    go_residuals4(n, A, B, W, X, o);
    *resid = o[0];
    *orth = o[1];
}

void go_compare(int n, const go_cplx* A, const go_cplx* B,
                const double* W, const go_cplx* X,
                const double* Wr, const go_cplx* Xr, double degen_gap,
                go_metrics* out) {
    memset(out, 0, sizeof *out);
    for (int i = 0; i < n; ++i) {
        double diff = fabs(W[i] - Wr[i]);
        if (!(diff == diff)) diff = INFINITY;
        double a = fabs(Wr[i]);
        double s = diff / (a > 1.0 ? a : 1.0);
        double r = a > 0.0 ? diff / a : diff;
        if (s > out->eig_rel_scaled) out->eig_rel_scaled = s;
        if (r > out->eig_rel) out->eig_rel = r;
    }
    if (Xr) {
        go_cplx* Bx = (go_cplx*)malloc(sizeof(go_cplx) * (size_t)n);
        for (int c = 0; c < n; ++c) {
            const go_cplx* xr = Xr + (size_t)c * n;
            const go_cplx* x = X + (size_t)c * n;
            for (int i = 0; i < n; ++i) {
                go_cplx acc = 0;
                for (int k = 0; k < n; ++k) acc += B[IDX(i, k, n)] * xr[k];
                Bx[i] = acc;
            }
            go_cplx inner = 0;
            for (int i = 0; i < n; ++i) inner += conj(x[i]) * Bx[i];
            double pe = 1.0 - cabs(inner);
            if (!(pe == pe)) pe = INFINITY;
            double gl = c > 0 ? fabs(Wr[c] - Wr[c - 1]) : INFINITY;
            double gr = c < n - 1 ? fabs(Wr[c + 1] - Wr[c]) : INFINITY;
            double gap = gl < gr ? gl : gr;
            if (gap < degen_gap) {
                out->degen_excluded++;
                if (pe > out->phase_err_degen) out->phase_err_degen = pe;
            } else if (pe > out->phase_err) {
                out->phase_err = pe;
            }
        }
        free(Bx);
    }
    double o[4];
    go_residuals4(n, A, B, W, X, o);
    out->resid = o[0];
    out->orth = o[1];
    out->resid_raw = o[2] == o[2] ? o[2] : INFINITY;
    out->orth_raw = o[3] == o[3] ? o[3] : INFINITY;
}

void go_compare_batch(int n, int batch, const go_cplx* A, const go_cplx* B,
                      const double* W, const go_cplx* X,
                      const double* Wr, const go_cplx* Xr, double degen_gap,
                      int nthreads, go_metrics* out) {
    const size_t nn = (size_t)n * n;
    memset(out, 0, sizeof *out);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int b = 0; b < batch; ++b) {
        go_metrics m;
        go_compare(n, A + nn * b, B + nn * b, W + (size_t)n * b, X + nn * b,
                   Wr + (size_t)n * b, Xr ? Xr + nn * b : NULL, degen_gap, &m);
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            if (m.eig_rel_scaled > out->eig_rel_scaled) out->eig_rel_scaled = m.eig_rel_scaled;
            if (m.eig_rel > out->eig_rel) out->eig_rel = m.eig_rel;
            if (m.phase_err > out->phase_err) out->phase_err = m.phase_err;
            if (m.phase_err_degen > out->phase_err_degen) out->phase_err_degen = m.phase_err_degen;
            out->degen_excluded += m.degen_excluded;
            if (m.resid > out->resid) out->resid = m.resid;
            if (m.orth > out->orth) out->orth = m.orth;
            if (m.resid_raw > out->resid_raw) out->resid_raw = m.resid_raw;
            if (m.orth_raw > out->orth_raw) out->orth_raw = m.orth_raw;
        }
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* GEIG v1: <6 x int32> magic 0x47454947, version 1, N, B, seed, 0; then    */
/* H, S (B N^2 c128 each), W (B N f64), U (B N^2 c128).                      */
/* ------------------------------------------------------------------------ */
int go_geig_read_header(const char* path, int32_t hdr[6]) {
    FILE* f = fopen(path, "rb");
    if (!f) return 1;
    size_t r = fread(hdr, sizeof(int32_t), 6, f);
    fclose(f);
    if (r != 6) return 2;
    if (hdr[0] != 0x47454947) return 3;
    if (hdr[1] != 1) return 4;
    return 0;
}

int go_geig_read(const char* path, int n, int batch, go_cplx* H, go_cplx* S,
                 double* W, go_cplx* U) {
    int32_t hdr[6];
    int rc = go_geig_read_header(path, hdr);
    if (rc) return rc;
    if (hdr[2] != n || hdr[3] < batch) return 5;
    FILE* f = fopen(path, "rb");
    if (!f) return 1;
    const size_t nn = (size_t)n * n, B = (size_t)hdr[3];
    int ok = fseek(f, 24, SEEK_SET) == 0;
    ok = ok && fread(H, sizeof(go_cplx), nn * batch, f) == nn * batch;
    ok = ok && fseek(f, 24 + (long)(sizeof(go_cplx) * nn * B), SEEK_SET) == 0;
    ok = ok && fread(S, sizeof(go_cplx), nn * batch, f) == nn * batch;
    ok = ok && fseek(f, 24 + (long)(2 * sizeof(go_cplx) * nn * B), SEEK_SET) == 0;
    ok = ok && fread(W, sizeof(double), (size_t)n * batch, f) == (size_t)n * batch;
    ok = ok && fseek(f, 24 + (long)(2 * sizeof(go_cplx) * nn * B + sizeof(double) * n * B), SEEK_SET) == 0;
    ok = ok && fread(U, sizeof(go_cplx), nn * batch, f) == nn * batch;
    fclose(f);
    return ok ? 0 : 6;
}

int go_geig_write(const char* path, int n, int batch, int seed,
                  const go_cplx* H, const go_cplx* S, const double* W,
                  const go_cplx* U) {
    FILE* f = fopen(path, "wb");
    if (!f) return 1;
    int32_t hdr[6] = {0x47454947, 1, n, batch, seed, 0};
    const size_t nn = (size_t)n * n;
    int ok = fwrite(hdr, sizeof(int32_t), 6, f) == 6;
    ok = ok && fwrite(H, sizeof(go_cplx), nn * batch, f) == nn * batch;
    ok = ok && fwrite(S, sizeof(go_cplx), nn * batch, f) == nn * batch;
    ok = ok && fwrite(W, sizeof(double), (size_t)n * batch, f) == (size_t)n * batch;
    ok = ok && fwrite(U, sizeof(go_cplx), nn * batch, f) == nn * batch;
    fclose(f);
    return ok ? 0 : 2;
}

uint64_t go_fnv1a64(const void* data, uint64_t len, uint64_t seed) {
    const unsigned char* p = (const unsigned char*)data;
    uint64_t h = seed ? seed : 0xcbf29ce484222325ULL;
    for (uint64_t i = 0; i < len; ++i) {
        h ^= (uint64_t)p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

/* Rayleigh quotients x_j^H A x_j / x_j^H B x_j of every column of X, in long
 * double (64-bit significand), for a batch stored column-major like the
 * solver's data: the batched form of oracle.rayleigh_ld (same sums, in C, one
 * pair per OpenMP thread) for checking many pairs at once.  A checker with no
 * reference counterpart. */
void go_rayleigh_ld_batch(int n, int batch, const go_cplx* A, const go_cplx* B, const go_cplx* X, double* out,
                          int nthreads) {
    const size_t nn = (size_t)n * n;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int p = 0; p < batch; ++p) {
        const go_cplx* Ap = A + nn * p;
        const go_cplx* Bp = B + nn * p;
        const go_cplx* Xp = X + nn * p;
        for (int j = 0; j < n; ++j) {
            const go_cplx* x = Xp + (size_t)j * n;
            long double num = 0.0L, den = 0.0L;
            for (int i = 0; i < n; ++i) {
                long double ar = 0.0L, ai = 0.0L, br = 0.0L, bi = 0.0L;
                for (int k = 0; k < n; ++k) {
                    const long double xr = creal(x[k]), xi = cimag(x[k]);
                    const go_cplx a = Ap[i + (size_t)k * n], b = Bp[i + (size_t)k * n];
                    ar += (long double)creal(a) * xr - (long double)cimag(a) * xi;
                    ai += (long double)creal(a) * xi + (long double)cimag(a) * xr;
                    br += (long double)creal(b) * xr - (long double)cimag(b) * xi;
                    bi += (long double)creal(b) * xi + (long double)cimag(b) * xr;
                }
                const long double cr = creal(x[i]), ci = cimag(x[i]);   /* conj(x_i) (y_i) */
                num += cr * ar + ci * ai;
                den += cr * br + ci * bi;
            }
            out[(size_t)p * n + j] = (double)(num / den);
        }
    }
}
