// Fold-in encoder kernels for gfx950 (declarations and contracts in
// include/tsvd_enc.h): per-user token features with set pooling, column
// statistics and row normalisation, an MFMA GEMM with fused epilogues for the
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

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------

constexpr float kDays = 2243.0f;  // day scale of the Netflix date range

// Per-item rating count, day sum and squared-day sum as exact 64-bit integers.
__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, int64_t n) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[e] = 0ull;
}

__global__ __launch_bounds__(256) void k_item_day_sums(const uint32_t* items, const int16_t* days, int64_t n,
                                                       int n_items, unsigned long long* work) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint32_t j = items[e];
        if (j >= static_cast<uint32_t>(n_items)) continue;  // checked on the host before the launch
        const long long d = days[e];
        atomicAdd(work + j, 1ull);
        atomicAdd(work + n_items + j, static_cast<unsigned long long>(d));  // two's complement wraps exactly
        atomicAdd(work + 2 * static_cast<int64_t>(n_items) + j, static_cast<unsigned long long>(d * d));
    }
}

__global__ __launch_bounds__(256) void k_item_day_final(const unsigned long long* work, int n_items, float* mean,
                                                        float* inv_std) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_items) return;
    const unsigned long long c = work[j];
    if (c == 0) {
        mean[j] = 0.0f;
        inv_std[j] = 0.0f;
        return;
    }
    const double n = static_cast<double>(c);
    const double m = static_cast<double>(static_cast<long long>(work[n_items + j])) / n;
    const double v = static_cast<double>(work[2 * static_cast<int64_t>(n_items) + j]) / n - m * m;
    const double sd = v > 0.0 ? sqrt(v) : 0.0;
    mean[j] = static_cast<float>(m);
    inv_std[j] = static_cast<float>(1.0 / (sd > 1.0 ? sd : 1.0));
}

// VEC consecutive BF16 values of a row, widened to FP32.
template <int VEC>
__device__ __forceinline__ void load_bf16_row(const bf16_t* p, float* x) {
    if constexpr (VEC == 1) {
        x[0] = to_f32(*p);
    } else {
        uint32_t w[VEC / 2];
        if constexpr (VEC == 2) {
            w[0] = *reinterpret_cast<const uint32_t*>(p);
        } else if constexpr (VEC == 4) {
            const uint2 v = *reinterpret_cast<const uint2*>(p);
            w[0] = v.x;
            w[1] = v.y;
        } else {
#pragma unroll
            for (int h = 0; h < VEC / 8; ++h) {
                const uint4 v = reinterpret_cast<const uint4*>(p)[h];
                w[4 * h] = v.x;
                w[4 * h + 1] = v.y;
                w[4 * h + 2] = v.z;
                w[4 * h + 3] = v.w;
            }
        }
#pragma unroll
        for (int e = 0; e < VEC / 2; ++e) {
            x[2 * e] = __uint_as_float(w[e] << 16);
            x[2 * e + 1] = __uint_as_float(w[e] & 0xffff0000u);
        }
    }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float bcast(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// One wave per user; lane l owns row elements [l*VEC, l*VEC + VEC).  Each
// rating's token is (q_j in BF16, one-hot of the rating, normalised day d_j,
// residual w_j); the set pooling is the mean over the sampled ratings of the
// one-hot (x) q_j blocks, d_j q_j and w_j q_j (the contract in tsvd_enc.h).
// The sample's records are loaded one per lane, 64 at a time, and broadcast
// with readlane; lanes past the sample hold item 0 with a zero one-hot,
// d and w, so padded steps of 8 add exact zeros.
template <int VEC>
__global__ __launch_bounds__(256) void k_user_features(const bf16_t* __restrict__ q, int f, int ldq,
                                                       const float* __restrict__ ibias, float mu,
                                                       const float* __restrict__ day_mean,
                                                       const float* __restrict__ day_inv, const uint32_t* off,
                                                       const uint32_t* items, const int8_t* ratings,
                                                       const float* targets, const int16_t* days, int U,
                                                       int max_items, float* out, int64_t ldo) {
    const int lane = threadIdx.x & 63;
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= U) return;
    const uint32_t s0 = off[u], cnt = off[u + 1] - s0;

    // scalars over every training rating of the user
    float c5[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float sw = 0.0f, sw2 = 0.0f, sd = 0.0f, sd2 = 0.0f, sday = 0.0f;
    float dmin = 3.0e38f, dmax = -3.0e38f;
    for (uint32_t t = lane; t < cnt; t += 64) {
        const uint32_t pos = s0 + t;
        const uint32_t j = items[pos];
        const int r = ratings[pos];
        const float w = targets[pos] - mu - ibias[j];
        const float day = static_cast<float>(days[pos]);
        const float d = (day - day_mean[j]) * day_inv[j];
#pragma unroll
        for (int b = 0; b < 5; ++b) c5[b] += r == b + 1 ? 1.0f : 0.0f;
        sw += w;
        sw2 += w * w;
        sd += d;
        sd2 += d * d;
        sday += day;
        dmin = fminf(dmin, day);
        dmax = fmaxf(dmax, day);
    }
#pragma unroll
    for (int b = 0; b < 5; ++b) c5[b] = wave_sum(c5[b]);
    sw = wave_sum(sw);
    sw2 = wave_sum(sw2);
    sd = wave_sum(sd);
    sd2 = wave_sum(sd2);
    sday = wave_sum(sday);
    dmin = wave_min(dmin);
    dmax = wave_max(dmax);

    // pooled blocks over up to max_items ratings at evenly spaced positions
    const uint32_t ns = cnt < static_cast<uint32_t>(max_items) ? cnt : static_cast<uint32_t>(max_items);
    float acc[7][VEC];
#pragma unroll
    for (int b = 0; b < 7; ++b)
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[b][e] = 0.0f;
    const bf16_t* qrow = q + lane * VEC;
    for (uint32_t t0 = 0; t0 < ns; t0 += 64) {
        const uint32_t t = t0 + lane;
        uint32_t jl = 0;
        float rl = 0.0f, wl = 0.0f, dl = 0.0f;
        if (t < ns) {
            const uint32_t pos = s0 + static_cast<uint32_t>((static_cast<uint64_t>(t) * cnt) / ns);
            jl = items[pos];
            rl = static_cast<float>(ratings[pos]);
            wl = targets[pos] - mu - ibias[jl];
            dl = (static_cast<float>(days[pos]) - day_mean[jl]) * day_inv[jl];
        }
        const uint32_t m = ns - t0 < 64u ? ns - t0 : 64u;
        const int m8 = static_cast<int>((m + 7u) & ~7u);
        for (int k0 = 0; k0 < m8; k0 += 8) {
            float x[8][VEC];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t j = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(jl), k0 + k));
                load_bf16_row<VEC>(qrow + static_cast<size_t>(j) * ldq, x[k]);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float r = bcast(rl, k0 + k), w = bcast(wl, k0 + k), d = bcast(dl, k0 + k);
#pragma unroll
                for (int b = 0; b < 5; ++b) {
                    const float hot = r == static_cast<float>(b + 1) ? 1.0f : 0.0f;
#pragma unroll
                    for (int e = 0; e < VEC; ++e) acc[b][e] += hot * x[k][e];
                }
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    acc[5][e] += d * x[k][e];
                    acc[6][e] += w * x[k][e];
                }
            }
        }
    }
    float* o = out + static_cast<int64_t>(u) * ldo;
    const float inv = ns ? 1.0f / static_cast<float>(ns) : 0.0f;
#pragma unroll
    for (int b = 0; b < 7; ++b)
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            const int col = lane * VEC + e;
            if (col < f) o[static_cast<int64_t>(b) * f + col] = acc[b][e] * inv;
        }
    const float n = static_cast<float>(cnt), in = cnt ? 1.0f / n : 0.0f;
    const float mw = sw * in, md = sd * in;
    float v = 0.0f;
    switch (lane) {
    case 0: v = log1pf(n) * 0.125f; break;
    case 1: case 2: case 3: case 4: case 5: v = c5[lane - 1] * in; break;
    case 6: v = mw; break;
    case 7: v = sqrtf(fmaxf(sw2 * in - mw * mw, 0.0f)); break;
    case 8: v = md; break;
    case 9: v = sqrtf(fmaxf(sd2 * in - md * md, 0.0f)); break;
    case 10: v = sday * in / kDays; break;
    case 11: v = cnt ? (dmax - dmin) / kDays : 0.0f; break;
    case 12: v = cnt ? dmin / kDays : 0.0f; break;
    case 13: v = cnt ? dmax / kDays : 0.0f; break;
    case 14: v = 1.0f; break;
    default: break;
    }
    o[7 * static_cast<int64_t>(f) + lane] = v;
}

// out[r, c] = (x[rows[r], c] - shift[c]) * scale[c] for r < n, c < d; zero in
// the padding (n <= r < n_out or d <= c < ldo).
template <typename T>
__global__ __launch_bounds__(256) void k_affine_rows(const float* x, int64_t ldx, const int64_t* rows, int n, int d,
                                                     const float* shift, const float* scale, T* out, int64_t ldo,
                                                     int n_out) {
    const int64_t total = static_cast<int64_t>(n_out) * ldo;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t r = e / ldo;
        const int c = static_cast<int>(e - r * ldo);
        float v = 0.0f;
        if (r < n && c < d) {
            const int64_t src = rows ? rows[r] : r;
            v = x[src * ldx + c];
            if (shift) v -= shift[c];
            if (scale) v *= scale[c];
        }
        out[e] = from_f32<T>(v);
    }
}

constexpr int kStatGroups = 256;

// FP64 partial sums of x and x^2 per column over the rows g, g + 256, ...
__global__ __launch_bounds__(256) void k_colstats_part(const float* x, int M, int N, int64_t ld, double* work) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (c >= N) return;
    double s = 0.0, s2 = 0.0;
    for (int r = g; r < M; r += kStatGroups) {
        const double v = x[static_cast<int64_t>(r) * ld + c];
        s += v;
        s2 += v * v;
    }
    work[static_cast<int64_t>(g) * N + c] = s;
    work[static_cast<int64_t>(kStatGroups + g) * N + c] = s2;
}

__global__ __launch_bounds__(256) void k_colstats_final(const double* work, int M, int N, float* mean, float* inv_std) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    double s = 0.0, s2 = 0.0;
    for (int g = 0; g < kStatGroups; ++g) {
        s += work[static_cast<int64_t>(g) * N + c];
        s2 += work[static_cast<int64_t>(kStatGroups + g) * N + c];
    }
    const double m = s / M, v = s2 / M - m * m;
    const double sd = v > 0.0 ? sqrt(v) : 0.0;
    mean[c] = static_cast<float>(m);
    inv_std[c] = static_cast<float>(sd > 1e-6 ? 1.0 / sd : 1.0);
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

int enc_item_day_stats(const uint32_t* items, const int16_t* days, int64_t n, int n_items, uint64_t* work,
                       float* day_mean, float* day_inv_std, void* stream) {
    g_err.clear();
    if (n < 0 || n_items <= 0) return fail(-1, "enc_item_day_stats: bad size");
    if (!work || !day_mean || !day_inv_std || (n > 0 && (!items || !days)))
        return fail(-1, "enc_item_day_stats: NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned long long* w = reinterpret_cast<unsigned long long*>(work);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(3 * static_cast<int64_t>(n_items))), dim3(256), 0, st, w,
                       3 * static_cast<int64_t>(n_items));
    if (n > 0)
        hipLaunchKernelGGL(k_item_day_sums, dim3(grid_for(n) > 8192 ? 8192 : grid_for(n)), dim3(256), 0, st, items,
                           days, n, n_items, w);
    hipLaunchKernelGGL(k_item_day_final, dim3((n_items + 255) / 256), dim3(256), 0, st, w, n_items, day_mean,
                       day_inv_std);
    return launched("enc_item_day_stats");
}

int enc_user_features(const uint16_t* q, int f, int ldq, const float* ibias, float mu, const float* day_mean,
                      const float* day_inv_std, const uint32_t* off, const uint32_t* items, const int8_t* ratings,
                      const float* targets, const int16_t* days, int U, int max_items, float* out, int64_t ldo,
                      void* stream) {
    g_err.clear();
    if (U <= 0) return 0;
    if (f <= 0 || f > 1024 || max_items <= 0) return fail(-1, "enc_user_features: need 0 < f <= 1024, max_items > 0");
    int vec = 1;
    while (64 * vec < f) vec *= 2;
    if (ldq < 64 * vec || ldq % vec) return fail(-1, "enc_user_features: ldq must be a multiple of %d, >= %d", vec,
                                                 64 * vec);
    if (ldo < 7 * static_cast<int64_t>(f) + ENC_SCALAR_COLS)
        return fail(-1, "enc_user_features: ldo < 7 f + %d", ENC_SCALAR_COLS);
    if (!q || !ibias || !day_mean || !day_inv_std || !off || !items || !ratings || !targets || !days || !out)
        return fail(-1, "enc_user_features: NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const dim3 grid((U + 3) / 4);
    const bf16_t* qb = reinterpret_cast<const bf16_t*>(q);
    (void)hipGetLastError();
#define ENC_FEAT(V)                                                                                               \
    hipLaunchKernelGGL(k_user_features<V>, grid, dim3(256), 0, st, qb, f, ldq, ibias, mu, day_mean, day_inv_std, \
                       off, items, ratings, targets, days, U, max_items, out, ldo)
    switch (vec) {
    case 1: ENC_FEAT(1); break;
    case 2: ENC_FEAT(2); break;
    case 4: ENC_FEAT(4); break;
    case 8: ENC_FEAT(8); break;
    default: ENC_FEAT(16); break;
    }
#undef ENC_FEAT
    return launched("enc_user_features");
}

int enc_affine_rows(int dtype, const float* x, int64_t ldx, const int64_t* rows, int n, int d, const float* shift,
                    const float* scale, void* out, int64_t ldo, int n_out, void* stream) {
    g_err.clear();
    if (n < 0 || d < 0 || n_out < n || ldo < d || (n > 0 && ldx < d)) return fail(-1, "enc_affine_rows: bad shape");
    if (n_out == 0 || ldo == 0) return 0;
    if (!out || (n > 0 && d > 0 && !x)) return fail(-1, "enc_affine_rows: NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t total = static_cast<int64_t>(n_out) * ldo;
    (void)hipGetLastError();
    if (dtype == ENC_BF16)
        hipLaunchKernelGGL(k_affine_rows<bf16_t>, dim3(grid_for(total)), dim3(256), 0, st, x, ldx, rows, n, d, shift,
                           scale, static_cast<bf16_t*>(out), ldo, n_out);
    else if (dtype == ENC_F32)
        hipLaunchKernelGGL(k_affine_rows<float>, dim3(grid_for(total)), dim3(256), 0, st, x, ldx, rows, n, d, shift,
                           scale, static_cast<float*>(out), ldo, n_out);
    else
        return fail(-1, "enc_affine_rows: unknown dtype %d", dtype);
    return launched("enc_affine_rows");
}

int enc_colstats(const float* x, int M, int N, int64_t ld, double* work, float* mean, float* inv_std, void* stream) {
    g_err.clear();
    if (M <= 0 || N <= 0 || ld < N) return fail(-1, "enc_colstats: bad shape");
    if (!x || !work || !mean || !inv_std) return fail(-1, "enc_colstats: NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_colstats_part, dim3((N + 255) / 256, kStatGroups), dim3(256), 0, st, x, M, N, ld, work);
    hipLaunchKernelGGL(k_colstats_final, dim3((N + 255) / 256), dim3(256), 0, st, work, M, N, mean, inv_std);
    return launched("enc_colstats");
}

const char* enc_last_error(void) { return g_err.c_str(); }

}  // extern "C"
