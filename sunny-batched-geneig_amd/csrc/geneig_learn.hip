// geneig_learn.hip -- learned warm start for the n = 54 spin-wave family
// (hipcc, gfx950).
//
// A small MLP maps the parameters of a pair -- wavevector q, coupling scale s
// and the lambda_min(H) target t -- to an approximate eigenbasis of the
// reduced matrix C = L^{-1} A L^{-H}.  Everything it needs runs here as
// hand-written kernels:
//   * training: FP32 GEMM with fused bias / tanh (forward and both backward
//     products), tanh derivative, bias gradient, Adam, and the loss kernel:
//     Cholesky-QR of the network output Z (Q = Z R^{-1}, R^H R = Z^H Z) and the
//     gauge-invariant loss off(Q^H C Q)^2 / ||C||_F^2 with its gradient dL/dZ
//     (QR backward, see cholqr_loss_kernel);
//   * inference: the same network forward, Cholesky-QR, C' = Q^H C Q and a few
//     FP32 cyclic Jacobi sweeps on C' that also rotate Q, giving eigenvectors
//     Y of C, then the precision form's back-transform, FP64 refinement and
//     FP64 fallback (geneig_learned_n54_batched_launch).
// Without features the Jacobi kernel starts from Q = I (plain Jacobi), the
// comparison that measures what the network saves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <set>

#include "geneig_device.inc"
#include "geneig_refine.inc"
#include "geneig_ws.hpp"
#include "../../include/geneig_learn.h"

namespace {

constexpr int kN = 54;
constexpr int kLD = 55;                    // odd LDS stride (float2 elements)
constexpr int kNN = kN * kLD;              // one LDS matrix, in float2
constexpr int kOut = 2 * kN * kN;          // network outputs per pair
constexpr int kThreads = 256;
constexpr int kF = 8;                      // network inputs per pair
constexpr int kHidden = 256;               // width of the three hidden layers
using geneig::cplx;
using geneig::cplxf;
using CD = geneig::Cfg<kN, double>;
using CF = geneig::Cfg<kN, float>;
using WD = geneig::Ws<kN, double>;
using WF = geneig::Ws<kN, float>;
using RC = geneig::RCfg<kN>;
constexpr unsigned kK3LdsF = geneig::k3_lds_bytes<kN, float>();
constexpr unsigned kK3LdsD = geneig::k3_lds_bytes<kN, double>();
constexpr unsigned kK1LdsF = geneig::k1_lds_bytes<kN, float>();
constexpr unsigned kK1LdsD = geneig::k1_lds_bytes<kN, double>();

// ---- FP32 GEMM: C = op(A) op(B) (+ bias per column) (tanh) ---------------------
// Row-major operands.  op(A) is M x K (A stored M x K, or K x M when TA),
// op(B) is K x N (B stored K x N, or N x K when TB).  64 x 64 tiles, k-steps
// of 16, 256 threads computing 4 x 4 outputs each.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) sgemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                     const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                     int ldc, const float* __restrict__ bias, int act) {
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid + 256 * e;
            int m, k;
            if (TA) { k = idx >> 6; m = idx & 63; } else { m = idx >> 4; k = idx & 15; }
            const int gm = m0 + m, gk = k0 + k;
            float v = 0.0f;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[k][m] = v;
            int n, kb;
            if (TB) { n = idx >> 4; kb = idx & 15; } else { kb = idx >> 6; n = idx & 63; }
            const int gn = n0 + n, gkb = k0 + kb;
            float w = 0.0f;
            if (gn < N && gkb < K) w = TB ? B[(size_t)gn * ldb + gkb] : B[(size_t)gkb * ldb + gn];
            Bs[kb][n] = w;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[k][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty * 4 + i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + tx * 4 + j;
            if (n >= N) continue;
            float v = acc[i][j] + (bias ? bias[n] : 0.0f);
            if (act == 1) v = tanhf(v);
            C[(size_t)m * ldc + n] = v;
        }
    }
}

// dY <- dY * (1 - Y^2)   (backward through tanh)
__global__ void tanh_back_kernel(const float* __restrict__ Y, float* __restrict__ dY, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dY[i] *= 1.0f - Y[i] * Y[i];
}

// db[c] = sum over rows of dY[:, c]
__global__ void colsum_kernel(const float* __restrict__ dY, int rows, int cols, float* __restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += dY[(size_t)r * cols + c];
    db[c] = s;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Features of (q0, q1, q2, s, t): cos/sin of 2 pi q (periodic in q), the scale
// centred and normalized, log10 t centred.
__global__ void features_kernel(const float* __restrict__ params, float* __restrict__ X, int batch) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const float* p = params + 5 * (size_t)b;
    float* x = X + (size_t)kF * b;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float sn, cs;
        sincospif(2.0f * p[c], &sn, &cs);
        x[2 * c] = cs;
        x[2 * c + 1] = sn;
    }
    x[6] = (p[3] - 1.0f) * 5.0f;
    x[7] = log10f(p[4]) + 2.0f;
}

// ---- small complex helpers on LDS matrices (column-major, stride kLD) ----------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {   // conj(a) * b
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconjf(float2 a) { return make_float2(a.x, -a.y); }

// D = scale * op(A) op(B) (+ Cinit).  op(A) = A^H when HA; op(B) has its
// diagonal dropped when OFFB.  D must not alias A or B (it may alias Cinit).
template <bool HA, bool OFFB>
__device__ void lds_mm(float2* D, const float2* A, const float2* B, const float2* Cinit, float scale) {
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
        const int i = t % kN, j = t / kN;
        float2 acc = make_float2(0.0f, 0.0f);
        for (int k = 0; k < kN; ++k) {
            if (OFFB && k == j) continue;
            const float2 b = B[k + j * kLD];
            acc = HA ? cadd(acc, cmulc(A[k + i * kLD], b)) : cadd(acc, cmul(A[i + k * kLD], b));
        }
        acc.x *= scale;
        acc.y *= scale;
        D[i + j * kLD] = Cinit ? cadd(acc, Cinit[i + j * kLD]) : acc;
    }
}

__device__ float block_sum(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float s = 0.0f;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w];
    __syncthreads();
    return s;
}

// Cholesky-QR: G = Z^H Z, R upper with R^H R = G, Q = Z R^{-1}.  Z in Zs, R out
// in Rs, Q out in Qs.  Returns false (on every thread) if G is not positive
// definite to working precision.
__device__ bool cholqr(const float2* Zs, float2* Rs, float2* Qs, int* flag) {
    lds_mm<true, false>(Rs, Zs, Zs, nullptr, 1.0f);
    if (threadIdx.x == 0) *flag = 0;
    __syncthreads();
    for (int k = 0; k < kN; ++k) {
        const float gkk = Rs[k + k * kLD].x;
        const bool ok = gkk > 0.0f;
        if (!ok) {
            if (threadIdx.x == 0) *flag = 1;
            break;
        }
        const float inv = 1.0f / sqrtf(gkk);
        for (int j = k + 1 + threadIdx.x; j < kN; j += kThreads) {
            const float2 g = Rs[k + j * kLD];
            Rs[k + j * kLD] = make_float2(g.x * inv, g.y * inv);
        }
        __syncthreads();
        if (threadIdx.x == 0) Rs[k + k * kLD] = make_float2(gkk * inv, 0.0f);
        // trailing update of the upper triangle: G(i,j) -= conj(R(k,i)) R(k,j), k < i <= j
        const int m = kN - 1 - k;
        for (int t = threadIdx.x; t < m * m; t += kThreads) {
            const int i = k + 1 + t % m, j = k + 1 + t / m;
            if (i <= j) Rs[i + j * kLD] = csub(Rs[i + j * kLD], cmulc(Rs[k + i * kLD], Rs[k + j * kLD]));
        }
        __syncthreads();
    }
    __syncthreads();
    if (*flag) return false;
    // Q(i, :) R = Z(i, :), forward over columns, one row per thread
    for (int i = threadIdx.x; i < kN; i += kThreads) {
        for (int k = 0; k < kN; ++k) {
            float2 acc = Zs[i + k * kLD];
            for (int m = 0; m < k; ++m) acc = csub(acc, cmul(Qs[i + m * kLD], Rs[m + k * kLD]));
            const float inv = 1.0f / Rs[k + k * kLD].x;
            Qs[i + k * kLD] = make_float2(acc.x * inv, acc.y * inv);
        }
    }
    __syncthreads();
    return true;
}

// Loss and gradient for one sample per workgroup.
//   Q = Z R^{-1},  M = Q^H C Q,  loss = sum_{i != j} |M_ij|^2 / ||C||_F^2
//   Gbar = dloss/dQ = (4 / ||C||^2) C Q off(M)                    (Wirtinger form)
//   A = Q^H Gbar,  K = tril(A - A^H, -1) + diag(A - A^H) / 2
//   dloss/dZ = (Gbar + Q (K - A)) R^{-H}                          (QR backward)
// The gradient is scaled by `scale` (1 / batch for a mean loss).  Z is a row of
// the network output: Z(i, j) = (out[2 (i + 54 j)], out[2 (i + 54 j) + 1]).
__global__ void __launch_bounds__(kThreads) cholqr_loss_kernel(const float* __restrict__ Zf, const float2* __restrict__ Cset,
                                                               const int* __restrict__ cidx, int batch, float scale,
                                                               float* __restrict__ loss, float* __restrict__ dZf) {
    extern __shared__ __attribute__((aligned(16))) float2 lsm[];
    __shared__ float red[kThreads / 64];
    __shared__ int flag;
    const int b = blockIdx.x;
    if (b >= batch) return;
    float2 *Zs = lsm, *Rs = lsm + kNN, *Qs = lsm + 2 * kNN, *Cs = lsm + 3 * kNN, *Bs4 = lsm + 4 * kNN,
           *Bs5 = lsm + 5 * kNN;
    const float* z = Zf + (size_t)kOut * b;
    const float2* c = Cset + (size_t)kN * kN * (cidx ? cidx[b] : b);
    float cn = 0.0f;
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
        const int i = t % kN, j = t / kN;
        Zs[i + j * kLD] = make_float2(z[2 * t], z[2 * t + 1]);
        const float2 cv = c[t];
        Cs[i + j * kLD] = cv;
        cn += cv.x * cv.x + cv.y * cv.y;
    }
    __syncthreads();
    cn = block_sum(cn, red);
    float* dz = dZf + (size_t)kOut * b;
    if (!cholqr(Zs, Rs, Qs, &flag) || !(cn > 0.0f)) {
        for (int t = threadIdx.x; t < kOut; t += kThreads) dz[t] = 0.0f;
        if (threadIdx.x == 0) loss[b] = __builtin_nanf("");
        return;
    }
    lds_mm<false, false>(Bs4, Cs, Qs, nullptr, 1.0f);            // CQ
    __syncthreads();
    lds_mm<true, false>(Bs5, Qs, Bs4, nullptr, 1.0f);            // M = Q^H C Q
    __syncthreads();
    float off = 0.0f;
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
        const int i = t % kN, j = t / kN;
        if (i != j) {
            const float2 m = Bs5[i + j * kLD];
            off += m.x * m.x + m.y * m.y;
        }
    }
    off = block_sum(off, red);
    if (threadIdx.x == 0) loss[b] = off / cn;
    const float g = 4.0f * scale / cn;
    lds_mm<false, true>(Zs, Bs4, Bs5, nullptr, g);                // Gbar = g C Q off(M)   (Z no longer needed)
    __syncthreads();
    lds_mm<true, false>(Bs4, Qs, Zs, nullptr, 1.0f);             // A = Q^H Gbar          (CQ no longer needed)
    __syncthreads();
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {       // K - A                 (M no longer needed)
        const int i = t % kN, j = t / kN;
        const float2 a = Bs4[i + j * kLD], at = cconjf(Bs4[j + i * kLD]);
        float2 k = make_float2(0.0f, 0.0f);
        if (i > j) k = csub(a, at);
        else if (i == j) k = make_float2(0.0f, a.y);              // (A - A^H)_ii / 2 = i Im A_ii
        Bs5[i + j * kLD] = csub(k, a);
    }
    __syncthreads();
    lds_mm<false, false>(Bs4, Qs, Bs5, Zs, 1.0f);                // T = Gbar + Q (K - A)
    __syncthreads();
    // dZ = T R^{-H}: dZ(i, k) = (T(i, k) - sum_{m > k} dZ(i, m) conj(R(k, m))) / R(k, k)
    for (int i = threadIdx.x; i < kN; i += kThreads) {
        for (int k = kN - 1; k >= 0; --k) {
            float2 acc = Bs4[i + k * kLD];
            for (int m = k + 1; m < kN; ++m) acc = csub(acc, cmul(Bs4[i + m * kLD], cconjf(Rs[k + m * kLD])));
            const float inv = 1.0f / Rs[k + k * kLD].x;
            Bs4[i + k * kLD] = make_float2(acc.x * inv, acc.y * inv);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
        const float2 v = Bs4[(t % kN) + (t / kN) * kLD];
        dz[2 * t] = v.x;
        dz[2 * t + 1] = v.y;
    }
}

// Cyclic Jacobi on C' = Q^H C Q (FP32), round-robin ordering: 27 disjoint
// (p, q) pairs per round, 53 rounds per sweep.  Each rotation U = diag(1,
// e^{-i phi}) [[c, s], [-s, c]] (b = |b| e^{i phi}, (c, s) from the symmetric
// Schur 2 x 2 of Golub & Van Loan) is applied as C' <- U^H C' U and Q <- Q U.
// Seeds with the Cholesky-QR of the network output, or Q = I without it.
// Writes the diagonal (eigenvalues) and Q (eigenvectors of C) to the
// workspace, and the sweeps taken.  Pairs with info != 0 are skipped.
__global__ void __launch_bounds__(kThreads) jacobi_kernel(const float* __restrict__ Zf, unsigned char* ws, int chunk,
                                                          const int* __restrict__ info, int max_sweeps, float tol,
                                                          int* __restrict__ sweeps_out) {
    extern __shared__ __attribute__((aligned(16))) float2 lsm[];
    __shared__ float red[kThreads / 64];
    __shared__ int flag;
    __shared__ float rc[kN / 2], rs[kN / 2];
    __shared__ float2 rph[kN / 2];
    __shared__ int rp[kN / 2], rq[kN / 2];
    const int b = blockIdx.x;
    if (b >= chunk) return;
    if (info[b] != 0) {
        if (threadIdx.x == 0 && sweeps_out) sweeps_out[b] = 0;
        return;
    }
    geneig::Ws<kN, float> w(ws, b, chunk);
    float2 *Zs = lsm, *Rs = lsm + kNN, *Qs = lsm + 2 * kNN, *Cs = lsm + 3 * kNN, *CQ = lsm + 4 * kNN,
           *Ms = lsm + 5 * kNN;
    const float2* cw = reinterpret_cast<const float2*>(w.Q);
    float cn = 0.0f;
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
        const int i = t % kN, j = t / kN;
        const float2 cv = cw[t];
        Cs[i + j * kLD] = cv;
        cn += cv.x * cv.x + cv.y * cv.y;
        if (Zf) Zs[i + j * kLD] = make_float2(Zf[(size_t)kOut * b + 2 * t], Zf[(size_t)kOut * b + 2 * t + 1]);
        else Qs[i + j * kLD] = make_float2(i == j ? 1.0f : 0.0f, 0.0f);
    }
    __syncthreads();
    cn = block_sum(cn, red);
    bool seeded = Zf != nullptr;
    if (seeded && !cholqr(Zs, Rs, Qs, &flag)) {
        // unusable network output: start from the identity
        for (int t = threadIdx.x; t < kN * kN; t += kThreads)
            Qs[(t % kN) + (t / kN) * kLD] = make_float2((t % kN) == (t / kN) ? 1.0f : 0.0f, 0.0f);
        seeded = false;
        __syncthreads();
    }
    if (seeded) {
        lds_mm<false, false>(CQ, Cs, Qs, nullptr, 1.0f);
        __syncthreads();
        lds_mm<true, false>(Ms, Qs, CQ, nullptr, 1.0f);
    } else {
        for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
            const int i = t % kN, j = t / kN;
            Ms[i + j * kLD] = Cs[i + j * kLD];
        }
    }
    __syncthreads();
    // exactly Hermitian from the lower triangle
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
        const int i = t % kN, j = t / kN;
        if (i < j) Ms[i + j * kLD] = cconjf(Ms[j + i * kLD]);
        else if (i == j) Ms[i + i * kLD].y = 0.0f;
    }
    __syncthreads();
    const float stop = tol * tol * cn;
    int sweep = 0;
    for (;;) {
        float off = 0.0f;
        for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
            const int i = t % kN, j = t / kN;
            if (i != j) {
                const float2 m = Ms[i + j * kLD];
                off += m.x * m.x + m.y * m.y;
            }
        }
        off = block_sum(off, red);
        if (!(off > stop) || sweep >= max_sweeps) break;
        ++sweep;
        for (int r = 0; r < kN - 1; ++r) {
            if (threadIdx.x < kN / 2) {
                const int k = threadIdx.x;
                auto pos = [&](int i) { return i == 0 ? 0 : ((i - 1 + r) % (kN - 1)) + 1; };
                int p = pos(k), q = pos(kN - 1 - k);
                if (p > q) { const int x = p; p = q; q = x; }
                const float a = Ms[p + p * kLD].x, d = Ms[q + q * kLD].x;
                const float2 bb = Ms[p + q * kLD];
                const float ab = sqrtf(bb.x * bb.x + bb.y * bb.y);
                float c = 1.0f, s = 0.0f;
                float2 ph = make_float2(1.0f, 0.0f);
                if (ab > 1e-30f && ab > 1e-7f * sqrtf(fabsf(a * d))) {
                    const float tau = (d - a) / (2.0f * ab);
                    const float tt = (tau >= 0.0f ? 1.0f : -1.0f) / (fabsf(tau) + sqrtf(1.0f + tau * tau));
                    c = 1.0f / sqrtf(1.0f + tt * tt);
                    s = tt * c;
                    ph = make_float2(bb.x / ab, bb.y / ab);   // e^{i phi}
                }
                rc[k] = c;
                rs[k] = s;
                rph[k] = ph;
                rp[k] = p;
                rq[k] = q;
            }
            __syncthreads();
            // rows: new row p = c row_p - s e^{i phi} row_q, new row q = s row_p + c e^{i phi} row_q
            for (int t = threadIdx.x; t < (kN / 2) * kN; t += kThreads) {
                const int k = t / kN, j = t % kN;
                const int p = rp[k], q = rq[k];
                const float c = rc[k], s = rs[k];
                const float2 ph = rph[k];
                const float2 x = Ms[p + j * kLD], y = cmul(ph, Ms[q + j * kLD]);
                Ms[p + j * kLD] = make_float2(c * x.x - s * y.x, c * x.y - s * y.y);
                Ms[q + j * kLD] = make_float2(s * x.x + c * y.x, s * x.y + c * y.y);
            }
            __syncthreads();
            // columns of C' and Q: new col p = c col_p - s e^{-i phi} col_q, new col q = s col_p + c e^{-i phi} col_q
            for (int t = threadIdx.x; t < (kN / 2) * kN * 2; t += kThreads) {
                const int half = t / ((kN / 2) * kN), u = t % ((kN / 2) * kN);
                const int k = u / kN, i = u % kN;
                float2* X = half ? Qs : Ms;
                const int p = rp[k], q = rq[k];
                const float c = rc[k], s = rs[k];
                const float2 ph = cconjf(rph[k]);
                const float2 x = X[i + p * kLD], y = cmul(ph, X[i + q * kLD]);
                X[i + p * kLD] = make_float2(c * x.x - s * y.x, c * x.y - s * y.y);
                X[i + q * kLD] = make_float2(s * x.x + c * y.x, s * x.y + c * y.y);
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < kN * kN; t += kThreads) {
        const int i = t % kN, j = t / kN;
        w.Q[t] = geneig::mk<float>(Qs[i + j * kLD].x, Qs[i + j * kLD].y);
    }
    for (int i = threadIdx.x; i < kN; i += kThreads) w.d[i] = Ms[i + i * kLD].x;
    if (threadIdx.x == 0 && sweeps_out) sweeps_out[b] = sweep;
}

constexpr unsigned kLossLds = 6u * kNN * sizeof(float2);   // 142,560 B

// pipeline kernels of the learned form (one pair per workgroup)
__global__ void __launch_bounds__(64, GENEIG_CHOL64_WAVES_PER_EU) chol64_kernel(const cplx* __restrict__ B, double* __restrict__ W,
                                                     cplx* __restrict__ U, int* __restrict__ info, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplx vecs[CD::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_CHOL64, double>(B, B, W, U, info, nullptr, batch, batch, smem, vecs);
}
__global__ void __launch_bounds__(64) k1c_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                  double* __restrict__ W, cplx* __restrict__ U,
                                                  int* __restrict__ info, unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplxf vecs[CF::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_MIXED_C, float>(A, B, W, U, info, ws, batch, batch, smem, vecs);
}
__global__ void __launch_bounds__(64, GENEIG_K3F_WAVES_PER_EU) k3j_kernel(const cplx* __restrict__ B, double* __restrict__ W,
                                                  cplx* __restrict__ U, const int* __restrict__ info,
                                                  unsigned char* ws, int batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplxf vecs[CF::VEC];
    geneig::k3_batched<kN, 1, geneig::MODE_JACOBI, float>(B, W, U, info, ws, batch, batch, smem, vecs);
}
__global__ void __launch_bounds__(64 * RC::NT, geneig::refine_waves_per_eu<kN>()) refine_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                                 double* __restrict__ W, cplx* __restrict__ U,
                                                                 const int* __restrict__ info, int* __restrict__ fb_list,
                                                                 int* __restrict__ fb_count, int* __restrict__ iters,
                                                                 int batch, int maxit, double tol) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    geneig::refine_batched<kN>(A, B, W, U, info, fb_list, fb_count, iters, batch, maxit, tol, smem);
}
__global__ void __launch_bounds__(64) k1l_kernel(const cplx* __restrict__ A, const cplx* __restrict__ B,
                                                  double* __restrict__ W, cplx* __restrict__ U,
                                                  int* __restrict__ info, unsigned char* ws, int chunk,
                                                  const int* __restrict__ list, const int* __restrict__ dcount,
                                                  int off) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ cplx vecs[CD::VEC];
    geneig::k1_batched<kN, 1, geneig::MODE_FULL, double, true>(
        A, B, W, U, info, ws,
        geneig::dispatch_slots(0, dcount, off, chunk), chunk, smem,
                                                         vecs, list + off);
}
__global__ void __launch_bounds__(64 * CD::R, GENEIG_K2_WAVES_PER_EU) k2l_kernel(int* __restrict__ info,
                                                                               unsigned char* ws, int chunk,
                                                                               const int* __restrict__ list,
                                                                               const int* __restrict__ dcount,
                                                                               int off) {
    __shared__ cplx rotlog[CD::R * CD::K2_VEC];
    geneig::k2_batched<kN, double, true>(info, ws, geneig::dispatch_slots(0, dcount, off, chunk), chunk, rotlog,
                                   list + off);
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

hipError_t gemm(int ta, int tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc, const float* bias, int act, hipStream_t s) {
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
    if (!ta && !tb) hipLaunchKernelGGL((sgemm_kernel<false, false>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, act);
    else if (!ta && tb) hipLaunchKernelGGL((sgemm_kernel<false, true>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, act);
    else if (ta && !tb) hipLaunchKernelGGL((sgemm_kernel<true, false>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, act);
    else hipLaunchKernelGGL((sgemm_kernel<true, true>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, act);
    return hipGetLastError();
}

// Network forward for P rows of features into Z (P x kOut); H holds the three
// hidden activations (3 x P x hidden).
hipError_t mlp_forward(const float* const* wts, const float* X, int P, float* H, float* Z, hipStream_t s) {
    const int F = kF, Hd = kHidden;
    float* h1 = H;
    float* h2 = H + (size_t)P * Hd;
    float* h3 = H + 2 * (size_t)P * Hd;
    RETURN_IF(gemm(0, 1, P, Hd, F, X, F, wts[0], F, h1, Hd, wts[1], 1, s));
    RETURN_IF(gemm(0, 1, P, Hd, Hd, h1, Hd, wts[2], Hd, h2, Hd, wts[3], 1, s));
    RETURN_IF(gemm(0, 1, P, Hd, Hd, h2, Hd, wts[4], Hd, h3, Hd, wts[5], 1, s));
    return gemm(0, 1, P, kOut, Hd, h3, Hd, wts[6], Hd, Z, kOut, wts[7], 0, s);
}

}  // namespace

extern "C" {

int geneig_learn_num_features(void) { return kF; }
int geneig_learn_hidden(void) { return kHidden; }
int geneig_learn_num_outputs(void) { return kOut; }

int geneig_learn_features(const float* d_params, float* d_X, int batch, hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, stream, d_params, d_X,
                       batch);
    return hipGetLastError();
}

int geneig_learn_gemm(int trans_a, int trans_b, int M, int N, int K, const float* A, int lda, const float* B,
                      int ldb, float* C, int ldc, const float* bias, int act, hipStream_t stream) {
    return gemm(trans_a, trans_b, M, N, K, A, lda, B, ldb, C, ldc, bias, act, stream);
}

int geneig_learn_tanh_backward(const float* d_Y, float* d_dY, long long n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tanh_back_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, d_Y, d_dY, n);
    return hipGetLastError();
}

int geneig_learn_colsum(const float* d_dY, int rows, int cols, float* d_db, hipStream_t stream) {
    if (cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, stream, d_dY, rows, cols,
                       d_db);
    return hipGetLastError();
}

int geneig_learn_adam(float* d_p, const float* d_g, float* d_m, float* d_v, long long n, float lr, float beta1,
                      float beta2, float eps, int step, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, d_p, d_g, d_m, d_v, n,
                       lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

int geneig_learn_loss(const float* d_Z, const void* d_C, const int* d_cidx, int batch, float scale, float* d_loss,
                      float* d_dZ, hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(cholqr_loss_kernel), kLossLds));
    hipLaunchKernelGGL(cholqr_loss_kernel, dim3((unsigned)batch), dim3(kThreads), kLossLds, stream, d_Z,
                       static_cast<const float2*>(d_C), d_cidx, batch, scale, d_loss, d_dZ);
    return hipGetLastError();
}

int geneig_learn_mlp_forward(const float* const* d_weights, const float* d_X, int batch, float* d_H, float* d_Z,
                             hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    return mlp_forward(d_weights, d_X, batch, d_H, d_Z, stream);
}

int geneig_learn_reduce_n54(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, hipDoubleComplex* d_L,
                            int* d_info, void* d_C, int batch, hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(chol64_kernel), geneig::chol64_lds_bytes<kN>()));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1c_kernel), kK1LdsF));
    const int chunk = geneig_host::chunk_pairs(batch, WF::PAIR_BYTES);
    void* ws = nullptr;
    RETURN_IF(geneig_host::workspace(stream, (size_t)chunk * WF::PAIR_BYTES, &ws));
    const auto* A = reinterpret_cast<const cplx*>(d_H);
    const auto* B = reinterpret_cast<const cplx*>(d_S);
    auto* L = reinterpret_cast<cplx*>(d_L);
    hipLaunchKernelGGL(chol64_kernel, dim3((unsigned)batch), dim3(64), geneig::chol64_lds_bytes<kN>(), stream, B, nullptr, L, d_info,
                       batch);
    RETURN_IF(hipGetLastError());
    const size_t mat = (size_t)kN * kN;
    for (int off = 0; off < batch; off += chunk) {
        const int P = std::min(chunk, batch - off);
        hipLaunchKernelGGL(k1c_kernel, dim3((unsigned)P), dim3(64), kK1LdsF, stream, A + mat * off, B + mat * off,
                           nullptr, L + mat * off, d_info + off, static_cast<unsigned char*>(ws), P);
        RETURN_IF(hipGetLastError());
        // the chunk's C matrices are contiguous in the workspace's Q slot
        const unsigned char* q = static_cast<unsigned char*>(ws) + 2ull * sizeof(float) * kN * P;
        RETURN_IF(hipMemcpyAsync(static_cast<char*>(d_C) + sizeof(cplxf) * mat * off, q, sizeof(cplxf) * mat * P,
                                 hipMemcpyDeviceToDevice, stream));
    }
    return hipSuccess;
}

int geneig_learned_n54_batched_launch(const hipDoubleComplex* d_H, const hipDoubleComplex* d_S, double* d_W,
                                      hipDoubleComplex* d_U, int* d_info, int batch, const float* d_features,
                                      const float* const* d_weights, int max_sweeps, float jacobi_tol,
                                      int max_iter, double tol, int* d_sweeps, int* d_iters, int* h_fallback,
                                      hipStream_t stream) {
    if (h_fallback) *h_fallback = 0;
    if (batch <= 0) return hipSuccess;   // the library's rule: no work, no error
    if (d_features && !d_weights) return hipErrorInvalidValue;
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(chol64_kernel), geneig::chol64_lds_bytes<kN>()));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1c_kernel), kK1LdsF));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(jacobi_kernel), kLossLds));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3j_kernel), kK3LdsF));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(refine_kernel), RC::LDS));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k1l_kernel), kK1LdsD));
    RETURN_IF(set_lds_once(reinterpret_cast<const void*>(k3l_kernel), kK3LdsD));
    if (max_iter <= 0) max_iter = 4;
    if (!(tol > 0.0)) tol = 1e-7;
    if (max_sweeps < 0) max_sweeps = 0;
    if (!(jacobi_tol > 0.0f)) jacobi_tol = 1e-6f;
    const int F = kF, Hd = kHidden;
    // per pair in a chunk: the FP32 workspace, plus network activations and output
    const unsigned long long mlp_pair = d_features ? sizeof(float) * (3ull * Hd + kOut) : 0ull;
    const geneig_host::MixedLayout L = geneig_host::mixed_layout(batch, WF::PAIR_BYTES, WD::PAIR_BYTES, 1, mlp_pair);
    const int chunk_f = L.chunk_f;
    void* ws = nullptr;
    RETURN_IF(geneig_host::workspace(stream, L.bytes, &ws));
    int* fb_count = static_cast<int*>(ws);
    int* fb_list = fb_count + 1;
    unsigned char* solver = static_cast<unsigned char*>(ws) + L.list_bytes;
    const auto* A = reinterpret_cast<const cplx*>(d_H);
    const auto* B = reinterpret_cast<const cplx*>(d_S);
    auto* U = reinterpret_cast<cplx*>(d_U);
    const size_t mat = (size_t)kN * kN;

    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, fb_count, 1);
    RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(chol64_kernel, dim3((unsigned)batch), dim3(64), geneig::chol64_lds_bytes<kN>(), stream, B, d_W, U, d_info, batch);
    RETURN_IF(hipGetLastError());
    for (int off = 0; off < batch; off += chunk_f) {
        const int P = std::min(chunk_f, batch - off);
        float* H = reinterpret_cast<float*>(solver + (size_t)P * WF::PAIR_BYTES);
        float* Z = H + 3ull * Hd * P;
        hipLaunchKernelGGL(k1c_kernel, dim3((unsigned)P), dim3(64), kK1LdsF, stream, A + mat * off, B + mat * off,
                           d_W + (size_t)kN * off, U + mat * off, d_info + off, solver, P);
        RETURN_IF(hipGetLastError());
        if (d_features) RETURN_IF(mlp_forward(d_weights, d_features + (size_t)F * off, P, H, Z, stream));
        hipLaunchKernelGGL(jacobi_kernel, dim3((unsigned)P), dim3(kThreads), kLossLds, stream,
                           d_features ? (const float*)Z : nullptr, solver, P, (const int*)(d_info + off), max_sweeps,
                           jacobi_tol, d_sweeps ? d_sweeps + off : nullptr);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k3j_kernel, dim3((unsigned)P), dim3(64), kK3LdsF, stream, B + mat * off,
                           d_W + (size_t)kN * off, U + mat * off, (const int*)(d_info + off), solver, P);
        RETURN_IF(hipGetLastError());
    }
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)batch), dim3(64 * RC::NT), RC::LDS, stream, A, B, d_W, U,
                       (const int*)d_info, fb_list, fb_count, d_iters, batch, max_iter, tol);
    RETURN_IF(hipGetLastError());
    // the FP64 re-solve of the listed pairs, its length read on the device
    const int cd = L.chunk_d;
    const unsigned g1 = (unsigned)std::min(cd, geneig::kListGridK1);
    const unsigned g2 = (unsigned)std::min(cd, geneig::kListGridK2x);
    for (int r = 0; r < L.rounds; ++r) {
        const int off = r * cd;
        hipLaunchKernelGGL(k1l_kernel, dim3(g1), dim3(64), kK1LdsD, stream, A, B, d_W, U, d_info, solver, cd,
                           (const int*)fb_list, (const int*)fb_count, off);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k2l_kernel, dim3(g2), dim3(64 * CD::R), 0, stream, d_info, solver, cd,
                           (const int*)fb_list, (const int*)fb_count, off);
        RETURN_IF(hipGetLastError());
        hipLaunchKernelGGL(k3l_kernel, dim3(g1), dim3(64), kK3LdsD, stream, B, d_W, U, (const int*)d_info, solver, cd,
                           (const int*)fb_list, (const int*)fb_count, off);
        RETURN_IF(hipGetLastError());
    }
    if (h_fallback) {
        int count = 0;
        RETURN_IF(hipMemcpyAsync(&count, fb_count, sizeof(int), hipMemcpyDeviceToHost, stream));
        RETURN_IF(hipStreamSynchronize(stream));
        if (count < 0 || count > batch) return hipErrorUnknown;
        *h_fallback = count;
    }
    return hipSuccess;
}

}  // extern "C"
