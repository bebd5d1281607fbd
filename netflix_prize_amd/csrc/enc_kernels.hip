// Fold-in encoder kernels for gfx950 (declarations and contracts in
// include/tsvd_enc.h): set pooling, an MFMA GEMM with fused epilogues for the
// MLP's forward and backward passes, the loss gradient, column sums, Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "tsvd_enc.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int launched(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(-3, "%s: %s", what, hipGetErrorString(e));
    return 0;
}

typedef uint16_t bf16_t;  // bfloat16 bits
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float to_f32(bf16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ float to_f32(float x) { return x; }

// FP32 -> BF16, round to nearest even; NaN stays NaN.
__device__ __forceinline__ bf16_t bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return static_cast<bf16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<bf16_t>(u >> 16);
}

template <typename T>
__device__ __forceinline__ T from_f32(float x);
template <>
__device__ __forceinline__ bf16_t from_f32<bf16_t>(float x) { return bf16_rne(x); }
template <>
__device__ __forceinline__ float from_f32<float>(float x) { return x; }

// gelu(z) = z Phi(z) and its derivative Phi(z) + z phi(z), erf form.
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// ---------------------------------------------------------------------------
// GEMM: 256 threads = 4 waves per 64 x 64 tile of C; each wave owns a 32 x 32
// quarter as 2 x 2 MFMA 16 x 16 blocks.  K advances 32 at a time through LDS
// (A tile 64 x 32, B tile 32 x 64), loaded so that consecutive threads read
// consecutive addresses of whichever layout A and B are stored in.
//
// MFMA 16 x 16 operand layout (wave64, lane l, element e):
//   BF16 (16x16x16): A[l%16][4*(l/16)+e], B[4*(l/16)+e][l%16]   (e < 4)
//   FP32 (16x16x4):  A[l%16][l/16],       B[l/16][l%16]
//   result:          D[4*(l/16)+e][l%16]                       (e < 4)
// ---------------------------------------------------------------------------
constexpr int BM = 64, BN = 64, BK = 32;

template <typename T>
struct EpiArgs {
    float* cf;
    T* ct;
    const float* bias;
    const float* zin;
    int ldc;
};

template <typename T, bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const T* __restrict__ A, const T* __restrict__ B, int M, int N, int K,
                                              int lda, int ldb, EpiArgs<T> ep) {
    __shared__ T As[BM][BK + 2];
    __shared__ T Bs[BK][BN + 2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int r = 0; r < (BM * BK) / 256; ++r) {
            const int e = tid + 256 * r;
            if constexpr (TA) {  // A stored K x M: walk m fastest
                const int m = e % BM, k = e / BM;
                As[m][k] = A[static_cast<size_t>(k0 + k) * lda + m0 + m];
            } else {
                const int k = e % BK, m = e / BK;
                As[m][k] = A[static_cast<size_t>(m0 + m) * lda + k0 + k];
            }
        }
#pragma unroll
        for (int r = 0; r < (BK * BN) / 256; ++r) {
            const int e = tid + 256 * r;
            if constexpr (TB) {  // B stored N x K: walk k fastest
                const int k = e % BK, n = e / BK;
                Bs[k][n] = B[static_cast<size_t>(n0 + n) * ldb + k0 + k];
            } else {
                const int n = e % BN, k = e / BN;
                Bs[k][n] = B[static_cast<size_t>(k0 + k) * ldb + n0 + n];
            }
        }
        __syncthreads();
        if constexpr (sizeof(T) == 2) {
#pragma unroll
            for (int ks = 0; ks < BK; ks += 16) {
                v4s a[2], b[2];
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        a[t][e] = static_cast<short>(As[wm + t * 16 + (lane & 15)][ks + 4 * (lane >> 4) + e]);
                        b[t][e] = static_cast<short>(Bs[ks + 4 * (lane >> 4) + e][wn + t * 16 + (lane & 15)]);
                    }
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                    for (int tj = 0; tj < 2; ++tj)
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < BK; ks += 4) {
                float a[2], b[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    a[t] = to_f32(As[wm + t * 16 + (lane & 15)][ks + (lane >> 4)]);
                    b[t] = to_f32(Bs[ks + (lane >> 4)][wn + t * 16 + (lane & 15)]);
                }
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                    for (int tj = 0; tj < 2; ++tj)
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = m0 + wm + ti * 16 + 4 * (lane >> 4) + e;
                const int c = n0 + wn + tj * 16 + (lane & 15);
                const size_t o = static_cast<size_t>(r) * ep.ldc + c;
                const float v = acc[ti][tj][e];
                if constexpr (EPI == ENC_EPI_STORE) {
                    ep.cf[o] = v;
                } else if constexpr (EPI == ENC_EPI_BIAS) {
                    ep.cf[o] = v + ep.bias[c];
                } else if constexpr (EPI == ENC_EPI_BIAS_GELU) {
                    const float z = v + ep.bias[c];
                    ep.cf[o] = z;
                    ep.ct[o] = from_f32<T>(gelu(z));
                } else {
                    const float g = v * gelu_grad(ep.zin[o]);
                    ep.cf[o] = g;
                    ep.ct[o] = from_f32<T>(g);
                }
            }
}

template <typename T>
__global__ __launch_bounds__(256) void k_loss_grad(const float* Y, const float* Tg, const float* colw, float* dyf,
                                                   T* dyt, int M, int N, int ld, double* loss) {
    __shared__ double part[256];
    double s = 0.0;
    const int64_t n = static_cast<int64_t>(M) * N;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int r = static_cast<int>(e / N), c = static_cast<int>(e % N);
        const size_t o = static_cast<size_t>(r) * ld + c;
        const float d = Y[o] - Tg[o];
        const float g = 2.0f * colw[c] * d / static_cast<float>(M);
        dyf[o] = g;
        dyt[o] = from_f32<T>(g);
        s += static_cast<double>(colw[c]) * d * d;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] / M);
}

__global__ __launch_bounds__(256) void k_colsum(const float* X, int M, int N, int ld, float* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float s = 0.0f;
    for (int r = 0; r < M; ++r) s += X[static_cast<size_t>(r) * ld + c];
    out[c] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_adam(float* w, const float* g, float* m, float* v, T* wc, int64_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.0f - b1) * gi;
        const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[e] = wi;
        wc[e] = from_f32<T>(wi);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_cast(const float* x, T* y, int64_t n) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        y[e] = from_f32<T>(x[e]);
}

// One wave per user: row chunks of 64 consecutive elements per load.
__global__ __launch_bounds__(256) void k_pool(const float* q, int fs, const uint32_t* off, const uint32_t* items,
                                              const float* w, int U, int max_items, float* out, int ldo) {
    const int lane = threadIdx.x & 63;
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= U) return;
    const uint32_t s0 = off[u], cnt = off[u + 1] - s0;
    const uint32_t n = cnt < static_cast<uint32_t>(max_items) ? cnt : static_cast<uint32_t>(max_items);
    const float inv = n ? 1.0f / static_cast<float>(n) : 0.0f;
    for (int k0 = 0; k0 < fs; k0 += 64) {
        const int k = k0 + lane;
        float a = 0.0f, b = 0.0f;
        for (uint32_t t = 0; t < n; ++t) {
            const uint32_t pos = s0 + static_cast<uint32_t>((static_cast<uint64_t>(t) * cnt) / n);
            const float x = k < fs ? q[static_cast<size_t>(items[pos]) * fs + k] : 0.0f;
            a += x;
            b += w[pos] * x;
        }
        if (k < fs) {
            out[static_cast<size_t>(u) * ldo + k] = a * inv;
            out[static_cast<size_t>(u) * ldo + fs + k] = b * inv;
        }
    }
}

unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return static_cast<unsigned>(b);
}

template <typename T, bool TA, bool TB>
int gemm_t(int epi, const void* A, const void* B, int M, int N, int K, int lda, int ldb, EpiArgs<T> ep,
           hipStream_t st) {
    dim3 grid(N / BN, M / BM);
    const T* a = static_cast<const T*>(A);
    const T* b = static_cast<const T*>(B);
    (void)hipGetLastError();
    switch (epi) {
    case ENC_EPI_STORE: hipLaunchKernelGGL((k_gemm<T, TA, TB, ENC_EPI_STORE>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep); break;
    case ENC_EPI_BIAS: hipLaunchKernelGGL((k_gemm<T, TA, TB, ENC_EPI_BIAS>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep); break;
    case ENC_EPI_BIAS_GELU: hipLaunchKernelGGL((k_gemm<T, TA, TB, ENC_EPI_BIAS_GELU>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep); break;
    case ENC_EPI_GELU_BWD: hipLaunchKernelGGL((k_gemm<T, TA, TB, ENC_EPI_GELU_BWD>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep); break;
    default: return fail(-1, "enc_gemm: unknown epilogue %d", epi);
    }
    return launched("enc_gemm");
}

template <typename T>
int gemm_d(int ta, int tb, int epi, const void* A, const void* B, int M, int N, int K, int lda, int ldb, float* cf,
           void* ct, const float* bias, const float* zin, int ldc, hipStream_t st) {
    EpiArgs<T> ep{cf, static_cast<T*>(ct), bias, zin, ldc};
    if (ta) return tb ? gemm_t<T, true, true>(epi, A, B, M, N, K, lda, ldb, ep, st)
                      : gemm_t<T, true, false>(epi, A, B, M, N, K, lda, ldb, ep, st);
    return tb ? gemm_t<T, false, true>(epi, A, B, M, N, K, lda, ldb, ep, st)
              : gemm_t<T, false, false>(epi, A, B, M, N, K, lda, ldb, ep, st);
}

}  // namespace

extern "C" {

int enc_gemm(int dtype, int ta, int tb, int epi, const void* A, const void* B, int M, int N, int K, int lda, int ldb,
             float* cf, void* ct, const float* bias, const float* zin, int ldc, void* stream) {
    g_err.clear();
    if (M <= 0 || N <= 0 || K <= 0 || M % BM || N % BN || K % BK)
        return fail(-1, "enc_gemm: M, N must be positive multiples of 64 and K of 32 (got %d, %d, %d)", M, N, K);
    if (!A || !B || !cf) return fail(-1, "enc_gemm: NULL operand");
    if ((epi == ENC_EPI_BIAS || epi == ENC_EPI_BIAS_GELU) && !bias) return fail(-1, "enc_gemm: epilogue needs bias");
    if ((epi == ENC_EPI_BIAS_GELU || epi == ENC_EPI_GELU_BWD) && !ct) return fail(-1, "enc_gemm: epilogue needs ct");
    if (epi == ENC_EPI_GELU_BWD && !zin) return fail(-1, "enc_gemm: GELU backward needs zin");
    if (lda < (ta ? M : K) || ldb < (tb ? K : N) || ldc < N) return fail(-1, "enc_gemm: leading dimension too small");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dtype == ENC_BF16) return gemm_d<bf16_t>(ta, tb, epi, A, B, M, N, K, lda, ldb, cf, ct, bias, zin, ldc, st);
    if (dtype == ENC_F32) return gemm_d<float>(ta, tb, epi, A, B, M, N, K, lda, ldb, cf, ct, bias, zin, ldc, st);
    return fail(-1, "enc_gemm: unknown dtype %d", dtype);
}

int enc_loss_grad(int dtype, const float* Y, const float* T, const float* colw, float* dyf, void* dyt, int M, int N,
                  int ld, double* loss, void* stream) {
    g_err.clear();
    if (M <= 0 || N <= 0 || ld < N) return fail(-1, "enc_loss_grad: bad shape");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const unsigned g = grid_for(static_cast<int64_t>(M) * N) > 1024 ? 1024 : grid_for(static_cast<int64_t>(M) * N);
    (void)hipGetLastError();
    if (dtype == ENC_BF16)
        hipLaunchKernelGGL(k_loss_grad<bf16_t>, dim3(g), dim3(256), 0, st, Y, T, colw, dyf, static_cast<bf16_t*>(dyt),
                           M, N, ld, loss);
    else if (dtype == ENC_F32)
        hipLaunchKernelGGL(k_loss_grad<float>, dim3(g), dim3(256), 0, st, Y, T, colw, dyf, static_cast<float*>(dyt), M,
                           N, ld, loss);
    else
        return fail(-1, "enc_loss_grad: unknown dtype %d", dtype);
    return launched("enc_loss_grad");
}

int enc_colsum(const float* X, int M, int N, int ld, float* out, void* stream) {
    g_err.clear();
    if (M <= 0 || N <= 0 || ld < N) return fail(-1, "enc_colsum: bad shape");
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), X, M, N, ld,
                       out);
    return launched("enc_colsum");
}

int enc_adam(int dtype, float* w, const float* g, float* m, float* v, void* wcopy, int64_t n, float lr, float beta1,
             float beta2, float eps, int32_t t, void* stream) {
    g_err.clear();
    if (n <= 0 || t < 1) return fail(-1, "enc_adam: bad size or step");
    const float c1 = 1.0f - std::pow(beta1, static_cast<float>(t));
    const float c2 = 1.0f - std::pow(beta2, static_cast<float>(t));
    hipStream_t st = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();
    if (dtype == ENC_BF16)
        hipLaunchKernelGGL(k_adam<bf16_t>, dim3(grid_for(n)), dim3(256), 0, st, w, g, m, v,
                           static_cast<bf16_t*>(wcopy), n, lr, beta1, beta2, eps, c1, c2);
    else if (dtype == ENC_F32)
        hipLaunchKernelGGL(k_adam<float>, dim3(grid_for(n)), dim3(256), 0, st, w, g, m, v, static_cast<float*>(wcopy),
                           n, lr, beta1, beta2, eps, c1, c2);
    else
        return fail(-1, "enc_adam: unknown dtype %d", dtype);
    return launched("enc_adam");
}

int enc_cast(int dtype, const float* x, void* y, int64_t n, void* stream) {
    g_err.clear();
    if (n <= 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();
    if (dtype == ENC_BF16)
        hipLaunchKernelGGL(k_cast<bf16_t>, dim3(grid_for(n)), dim3(256), 0, st, x, static_cast<bf16_t*>(y), n);
    else if (dtype == ENC_F32)
        hipLaunchKernelGGL(k_cast<float>, dim3(grid_for(n)), dim3(256), 0, st, x, static_cast<float*>(y), n);
    else
        return fail(-1, "enc_cast: unknown dtype %d", dtype);
    return launched("enc_cast");
}

int enc_pool(const float* q, int fs, const uint32_t* off, const uint32_t* items, const float* w, int U, int max_items,
             float* out, int ldo, void* stream) {
    g_err.clear();
    if (U <= 0) return 0;
    if (fs <= 0 || max_items <= 0 || ldo < 2 * fs) return fail(-1, "enc_pool: bad shape");
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_pool, dim3((U + 3) / 4), dim3(256), 0, static_cast<hipStream_t>(stream), q, fs, off, items, w,
                       U, max_items, out, ldo);
    return launched("enc_pool");
}

const char* enc_last_error(void) { return g_err.c_str(); }

}  // extern "C"
