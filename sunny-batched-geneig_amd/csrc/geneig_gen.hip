// geneig_gen.hip -- on-device synthetic pair generators.
//
// Builds the benchmark and test families directly in HBM (a 65,536-pair
// n = 54 batch is 3 GB per matrix, too much to stream from host generators):
//   * generic pairs  A = (M + M^H) / (2 sqrt n),  B = G G^H / n + I
//   * spin-wave pairs A = eta = diag(I, -I),  B = H0(q) + (t - lambda_min(H0)) I
//     with H0(q) = [[Aq, Bq], [conj Bq, conj Aq]] summed over bonds.
// The RNG is the counter-based SplitMix64 + Box-Muller scheme shared with the
// host-side test generators, so both sides draw the same variates (the device
// transcendental functions may differ from the host's in the last bit, which
// is why parity tests always copy the actual device inputs back).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/geneig_gen.h"

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rng_key(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx) {
    const uint64_t s = splitmix64(seed ^ 0x5851F42D4C957F2DULL);
    return s ^ ((uint64_t)(stream & 0xFFu) << 56) ^ ((pair & 0xFFFFFFFFULL) << 24) ^
           (uint64_t)(idx & 0xFFFFFFu);
}
__device__ __forceinline__ double u01(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx) {
    const uint64_t r = splitmix64(rng_key(seed, stream, pair, idx));
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double2 cnormal(uint64_t seed, uint32_t stream, uint64_t pair, uint32_t idx) {
    const double u1 = u01(seed, stream, pair, 2u * idx);
    const double u2 = u01(seed, stream, pair, 2u * idx + 1u);
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincos(6.283185307179586476925286766559 * u2, &s, &c);
    return make_double2(rad * c, rad * s);
}

// One workgroup per pair; G is staged in LDS (n <= 100 fits 160 KB).
__global__ void __launch_bounds__(256) gen_generic_kernel(int n, uint64_t seed, uint64_t first_pair,
                                                          double2* __restrict__ A, double2* __restrict__ B,
                                                          int batch) {
    extern __shared__ double2 G[];
    const int b = blockIdx.x;
    if (b >= batch) return;
    const uint64_t p = first_pair + (uint64_t)b;
    const size_t nn = (size_t)n * n;
    double2* Ab = A + nn * b;
    double2* Bb = B + nn * b;
    const double sa = 1.0 / (2.0 * sqrt((double)n)), sb = 1.0 / (double)n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) G[t] = cnormal(seed, 1, p, (uint32_t)t);
    __syncthreads();
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int i = t % n, j = t / n;
        if (i < j) continue;
        const double2 mij = cnormal(seed, 0, p, (uint32_t)(i + j * n));
        const double2 mji = cnormal(seed, 0, p, (uint32_t)(j + i * n));
        double2 a = make_double2((mij.x + mji.x) * sa, (mij.y - mji.y) * sa);
        double re = 0.0, im = 0.0;
        for (int k = 0; k < n; ++k) {
            const double2 g1 = G[i + k * n], g2 = G[j + k * n];   // g1 * conj(g2)
            re += g1.x * g2.x + g1.y * g2.y;
            im += g1.y * g2.x - g1.x * g2.y;
        }
        double2 bv = make_double2(re * sb + (i == j ? 1.0 : 0.0), im * sb);
        if (i == j) {
            a.y = 0.0;
            bv.y = 0.0;
        }
        Ab[i + (size_t)j * n] = a;
        Bb[i + (size_t)j * n] = bv;
        if (i != j) {
            Ab[j + (size_t)i * n] = make_double2(a.x, -a.y);
            Bb[j + (size_t)i * n] = make_double2(bv.x, -bv.y);
        }
    }
}

struct Bond {
    int shift, r0, r1, r2;
};

__device__ __forceinline__ Bond bond(uint64_t seed, int nsub, int b) {
    Bond o;
    int s = 1 + (int)(u01(seed, 5, 0, (uint32_t)(8 * b + 0)) * (double)(nsub - 1));
    o.shift = s > nsub - 1 ? nsub - 1 : s;
    int r[3];
    for (int c = 0; c < 3; ++c) {
        int v = (int)(u01(seed, 5, 0, (uint32_t)(8 * b + 1 + c)) * 3.0) - 1;
        r[c] = v > 1 ? 1 : v;
    }
    o.r0 = r[0]; o.r1 = r[1]; o.r2 = r[2];
    return o;
}

// H0(q) for one grid point per workgroup; the bond loop is tiny (nbonds x nsub).
__global__ void __launch_bounds__(256) gen_spinwave_kernel(int nsub, int nbonds, uint64_t seed, int g0,
                                                           int g1, int g2, uint64_t first_q, double jscale,
                                                           double2* __restrict__ A, double2* __restrict__ B,
                                                           int batch) {
    extern __shared__ double2 H[];   // n x n
    const int b = blockIdx.x;
    if (b >= batch) return;
    const int N = nsub, n = 2 * nsub;
    const uint64_t qi = first_q + (uint64_t)b;
    const double q0 = (double)((qi / ((uint64_t)g2 * g1)) % (uint64_t)g0) / g0;
    const double q1 = (double)((qi / (uint64_t)g2) % (uint64_t)g1) / g1;
    const double q2 = (double)(qi % (uint64_t)g2) / g2;
    // jscale < 0: a coupling scale per q, s = 0.8 + 0.4 U (stream 8)
    const double js = jscale < 0.0 ? 0.8 + 0.4 * u01(seed, 8, qi, 0) : jscale;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) H[t] = make_double2(0.0, 0.0);
    __syncthreads();
    // Each (bond, site) term touches fixed entries; serialize per bond so no two
    // threads update the same entry concurrently (a site i hits (i,j), (j,i), (i,i)).
    for (int bd = 0; bd < nbonds; ++bd) {
        const Bond bo = bond(seed, nsub, bd);
        const double th = 6.283185307179586476925286766559 * (q0 * bo.r0 + q1 * bo.r1 + q2 * bo.r2);
        double sn, cs;
        sincos(th, &sn, &cs);
        if (threadIdx.x == 0) {
            for (int i = 0; i < N; ++i) {
                const double J = js * (0.5 + u01(seed, 6, (uint64_t)bd, (uint32_t)i));
                const int j = (i + bo.shift) % N;
                const double h = 0.5 * J;
                // Aq (top-left) and conj(Aq) (bottom-right)
                H[i + j * n].x += h * cs;  H[i + j * n].y += h * sn;
                H[j + i * n].x += h * cs;  H[j + i * n].y -= h * sn;
                H[i + i * n].x += J;
                // Bq (top-right): symmetric, phase e^{i th} on both (i,j) and (j,i)
                H[i + (j + N) * n].x += h * cs;  H[i + (j + N) * n].y += h * sn;
                H[j + (i + N) * n].x += h * cs;  H[j + (i + N) * n].y += h * sn;
            }
        }
        __syncthreads();
    }
    // fill the lower blocks from the upper ones: [[Aq, Bq], [conj Bq, conj Aq]]
    for (int t = threadIdx.x; t < N * N; t += blockDim.x) {
        const int i = t % N, j = t / N;
        const double2 a = H[i + j * n], bq = H[i + (j + N) * n];
        H[(i + N) + (j + N) * n] = make_double2(a.x, -a.y);
        H[(i + N) + j * n] = make_double2(bq.x, -bq.y);
    }
    __syncthreads();
    const size_t nn = (size_t)n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double2 h = H[t];
        if (i == j) h.y = 0.0;
        B[nn * b + t] = h;
        A[nn * b + t] = make_double2(i == j ? (i < N ? 1.0 : -1.0) : 0.0, 0.0);
    }
}

__global__ void shift_kernel(int n, uint64_t seed, uint64_t first_q, const double* __restrict__ W,
                             double2* __restrict__ B, int batch) {
    const int b = blockIdx.x * blockDim.y + threadIdx.y;
    if (b >= batch) return;
    const double t = pow(10.0, -3.0 + 2.0 * u01(seed, 7, first_q + (uint64_t)b, 0));
    const double sh = t - W[(size_t)b * n];
    for (int i = threadIdx.x; i < n; i += blockDim.x) B[(size_t)b * n * n + i + (size_t)i * n].x += sh;
}

// Per-pair parameters of the spin-wave family: q (3), the coupling scale s
// and the lambda_min target t, as float32 rows of 5.
__global__ void spinwave_params_kernel(uint64_t seed, int g0, int g1, int g2, uint64_t first_q, double jscale,
                                       float* __restrict__ out, int batch) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const uint64_t qi = first_q + (uint64_t)b;
    out[5 * b + 0] = (float)((double)((qi / ((uint64_t)g2 * g1)) % (uint64_t)g0) / g0);
    out[5 * b + 1] = (float)((double)((qi / (uint64_t)g2) % (uint64_t)g1) / g1);
    out[5 * b + 2] = (float)((double)(qi % (uint64_t)g2) / g2);
    out[5 * b + 3] = (float)(jscale < 0.0 ? 0.8 + 0.4 * u01(seed, 8, qi, 0) : jscale);
    out[5 * b + 4] = (float)pow(10.0, -3.0 + 2.0 * u01(seed, 7, qi, 0));
}

// The family's model as tables: bond b's site shift and offset r_b, and its
// coupling at site i, J[b nsub + i] = jscale (0.5 + U) (the draws of
// gen_spinwave_kernel).
__global__ void spinwave_model_kernel(int nsub, int nbonds, uint64_t seed, double jscale, int* __restrict__ shift,
                                      int* __restrict__ r, double* __restrict__ J) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nbonds) {
        const Bond bo = bond(seed, nsub, t);
        shift[t] = bo.shift;
        r[3 * t] = bo.r0;
        r[3 * t + 1] = bo.r1;
        r[3 * t + 2] = bo.r2;
    }
    if (t < nbonds * nsub) J[t] = jscale * (0.5 + u01(seed, 6, (uint64_t)(t / nsub), (uint32_t)(t % nsub)));
}

__global__ void identity_kernel(int n, double2* __restrict__ B, long long total) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const long long e = t % ((long long)n * n);
    const int i = (int)(e % n), j = (int)(e / n);
    B[t] = make_double2(i == j ? 1.0 : 0.0, 0.0);
}

}  // namespace

extern "C" {

int geneig_gen_generic_launch(int n, int batch, uint64_t seed, uint64_t first_pair, void* d_A, void* d_B,
                              hipStream_t stream) {
    if (n < 1 || n > 100) return (int)hipErrorInvalidValue;
    if (batch <= 0) return 0;
    const unsigned lds = (unsigned)(16 * n * n);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(gen_generic_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(gen_generic_kernel, dim3(batch), dim3(256), lds, stream, n, seed, first_pair,
                       static_cast<double2*>(d_A), static_cast<double2*>(d_B), batch);
    return (int)hipGetLastError();
}

int geneig_gen_spinwave_h0_launch(int nsub, int nbonds, uint64_t seed, int g0, int g1, int g2,
                                  uint64_t first_q, int batch, double jscale, void* d_A, void* d_B,
                                  hipStream_t stream) {
    if (nsub < 2 || 2 * nsub > 100 || nbonds < 1 || g0 < 1 || g1 < 1 || g2 < 1) return (int)hipErrorInvalidValue;
    if (batch <= 0) return 0;
    const int n = 2 * nsub;
    const unsigned lds = (unsigned)(16 * n * n);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(gen_spinwave_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(gen_spinwave_kernel, dim3(batch), dim3(256), lds, stream, nsub, nbonds, seed, g0, g1, g2,
                       first_q, jscale, static_cast<double2*>(d_A), static_cast<double2*>(d_B), batch);
    return (int)hipGetLastError();
}

int geneig_spinwave_shift_launch(int n, int batch, uint64_t seed, uint64_t first_q, const double* d_W,
                                 void* d_B, hipStream_t stream) {
    if (n < 1) return (int)hipErrorInvalidValue;
    if (batch <= 0) return 0;
    hipLaunchKernelGGL(shift_kernel, dim3((batch + 3) / 4), dim3(64, 4), 0, stream, n, seed, first_q, d_W,
                       static_cast<double2*>(d_B), batch);
    return (int)hipGetLastError();
}

int geneig_spinwave_params_launch(uint64_t seed, int g0, int g1, int g2, uint64_t first_q, int batch,
                                  double jscale, float* d_params, hipStream_t stream) {
    if (batch <= 0) return hipSuccess;
    if (g0 <= 0 || g1 <= 0 || g2 <= 0 || !d_params) return hipErrorInvalidValue;
    hipLaunchKernelGGL(spinwave_params_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, stream, seed, g0,
                       g1, g2, first_q, jscale, d_params, batch);
    return hipGetLastError();
}

int geneig_spinwave_model_launch(int nsub, int nbonds, uint64_t seed, double jscale, int* d_shift, int* d_r,
                                 double* d_J, hipStream_t stream) {
    if (nsub < 2 || nbonds < 1 || !(jscale >= 0.0) || !d_shift || !d_r || !d_J) return (int)hipErrorInvalidValue;
    const int total = nbonds * nsub;
    hipLaunchKernelGGL(spinwave_model_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, nsub,
                       nbonds, seed, jscale, d_shift, d_r, d_J);
    return (int)hipGetLastError();
}

int geneig_fill_identity_launch(int n, int batch, void* d_B, hipStream_t stream) {
    if (n < 1) return (int)hipErrorInvalidValue;
    if (batch <= 0) return 0;
    const long long total = (long long)n * n * batch;
    hipLaunchKernelGGL(identity_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, n,
                       static_cast<double2*>(d_B), total);
    return (int)hipGetLastError();
}

}  // extern "C"
