// Fold-in encoder kernels for gfx950 (declarations and contracts in
// include/tsvd_enc.h): per-user token features with set pooling, column
// statistics and row normalisation, an MFMA GEMM with fused epilogues for the
// MLP's forward and backward passes, the loss gradient, column sums, Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <mutex>
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

// ---------------------------------------------------------------------------
// BF16 GEMM on gfx950's 16 x 16 x 32 MFMA (K a multiple of 64, operands and
// leading dimensions 16-byte aligned).  256 threads = 4 waves in 2 x 2 per
// BM x BN tile of C (BM, BN in {64, 128}); a wave owns a (BM/2) x (BN/2)
// quarter as 16 x 16 blocks.  K advances 64 per stage through two LDS
// buffers, both operands stored [row][k] (row = m for A, n for B; rows padded
// to 72 elements = 144 bytes): a lane's 8 consecutive k of one row are one
// 16-byte LDS read.  Global loads are 16 bytes (8 elements) along whichever
// dimension is contiguous in memory; a k-contiguous chunk is one LDS write,
// an m/n-contiguous chunk is scattered over 8 rows.  The next stage's loads
// are issued before the current stage's MFMAs, written to the other buffer
// after them, and one barrier per stage separates the two.
//
// Operands of v_mfma_f32_16x16x32_bf16 (wave64, lane l, element e < 8):
//   A[l%16][8*(l/16)+e], B[8*(l/16)+e][l%16]; result D[4*(l/16)+e][l%16] (e < 4).
// A and B take the same lane/element -> k assignment, so the product is the
// sum over k whatever order the unit adds the 32 products in.
// ---------------------------------------------------------------------------
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
// staging registers: a native vector type (HIP's uint4 is a union wrapper that
// the compiler keeps in memory when arrays of it are passed around)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Workgroup barrier that waits for this wave's LDS operations only.
// __syncthreads() is a workgroup-scope release/acquire around s_barrier, and
// the release also waits for every global load in flight, which would drain
// the next stages' prefetch at each barrier; the stages share nothing but LDS.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
constexpr int GK = 64;       // K per stage
// tiles a product should have before a larger tile is taken: two workgroups
// per CU, so that one CU holds two tiles' loads in flight (a 4,096 x 512
// output: 128 x 64 tiles 41.0 us, 64 x 64 tiles 36.2 us)
constexpr int kMinTiles = 512;
constexpr int GKP = GK + 8;  // LDS row pitch in elements

template <int R, bool KCONT>
__device__ __forceinline__ void stage_load(const bf16_t* __restrict__ X, int ld, int r0, int k0, int tid,
                                           u32x4 (&v)[R * GK / 8 / 256]) {
#pragma unroll
    for (int q = 0; q < R * GK / 8 / 256; ++q) {
        const int c = tid + 256 * q;
        if constexpr (KCONT) {  // X stored [row][k]
            const int row = c >> 3, kc = c & 7;
            v[q] = *reinterpret_cast<const u32x4*>(X + static_cast<size_t>(r0 + row) * ld + k0 + kc * 8);
        } else {  // X stored [k][row]: rows g*8 .. g*8+7 of k rows kq*Q .. kq*Q+Q-1
            constexpr int Q = R * GK / 8 / 256;
            const int g = tid % (R / 8), kq = tid / (R / 8);
            v[q] = *reinterpret_cast<const u32x4*>(X + static_cast<size_t>(k0 + kq * Q + q) * ld + r0 + g * 8);
        }
    }
}

template <int R, bool KCONT>
__device__ __forceinline__ void stage_store(bf16_t* __restrict__ S, int tid, const u32x4 (&v)[R * GK / 8 / 256]) {
#pragma unroll
    for (int q = 0; q < R * GK / 8 / 256; ++q) {
        const int c = tid + 256 * q;
        if constexpr (KCONT) {
            const int row = c >> 3, kc = c & 7;
            *reinterpret_cast<u32x4*>(S + row * GKP + kc * 8) = v[q];
        }
    }
    if constexpr (!KCONT) {
        // the thread's Q chunks are Q consecutive k of the same 8 rows: row
        // g*8+e takes element e of every chunk, Q consecutive k in one write
        constexpr int Q = R * GK / 8 / 256;
        static_assert(Q == 2 || Q == 4, "two or four k per row and thread");
        const int g = tid % (R / 8), kq = tid / (R / 8);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int w = e >> 1, sh = 16 * (e & 1);
            uint32_t h[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) h[q] = (v[q][w] >> sh) & 0xffffu;
            bf16_t* dst = S + (g * 8 + e) * GKP + kq * Q;
            if constexpr (Q == 2) {
                *reinterpret_cast<uint32_t*>(dst) = h[0] | (h[1] << 16);
            } else {
                typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
                *reinterpret_cast<u32x2*>(dst) = u32x2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
            }
        }
    }
}

template <int BM, int BN, bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm_mfma(const bf16_t* __restrict__ A, const bf16_t* __restrict__ B, int M,
                                                   int N, int K, int lda, int ldb, EpiArgs<bf16_t> ep, int kper,
                                                   float* __restrict__ part) {
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
    constexpr int QA = BM * GK / 8 / 256, QB = BN * GK / 8 / 256;
    __shared__ __attribute__((aligned(16))) bf16_t As[2][BM * GKP];
    __shared__ __attribute__((aligned(16))) bf16_t Bs[2][BN * GKP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    // split-K: workgroup z takes K range [z kper, (z + 1) kper)
    const int kb = static_cast<int>(blockIdx.z) * kper;
    const int nst = kper / GK;
    v4f acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    // two register sets: stage s + 2 loads while stage s computes and stage
    // s + 1 (loaded one stage earlier) goes to LDS
    u32x4 ra0[QA], rb0[QB], ra1[QA], rb1[QB];
    const int arow = lane & 15, koff = 8 * (lane >> 4);
    auto compute = [&](int buf) {
        const bf16_t* as = As[buf];
        const bf16_t* bs = Bs[buf];
#pragma unroll
        for (int ks = 0; ks < GK; ks += 32) {
            v8bf a[TM], b[TN];
#pragma unroll
            for (int t = 0; t < TM; ++t)
                a[t] = *reinterpret_cast<const v8bf*>(as + (wm + t * 16 + arow) * GKP + ks + koff);
#pragma unroll
            for (int t = 0; t < TN; ++t)
                b[t] = *reinterpret_cast<const v8bf*>(bs + (wn + t * 16 + arow) * GKP + ks + koff);
#pragma unroll
            for (int ti = 0; ti < TM; ++ti)
#pragma unroll
                for (int tj = 0; tj < TN; ++tj)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
        }
    };

    stage_load<BM, !TA>(A, lda, m0, kb, tid, ra0);
    stage_load<BN, TB>(B, ldb, n0, kb, tid, rb0);
    if (nst > 1) {
        stage_load<BM, !TA>(A, lda, m0, kb + GK, tid, ra1);
        stage_load<BN, TB>(B, ldb, n0, kb + GK, tid, rb1);
    }
    stage_store<BM, !TA>(As[0], tid, ra0);
    stage_store<BN, TB>(Bs[0], tid, rb0);
    lds_barrier();
    for (int s = 0; s < nst; s += 2) {
        // even stage s: buffer 0 computes, set 1 (stage s + 1) goes to buffer 1
        if (s + 2 < nst) {
            stage_load<BM, !TA>(A, lda, m0, kb + (s + 2) * GK, tid, ra0);
            stage_load<BN, TB>(B, ldb, n0, kb + (s + 2) * GK, tid, rb0);
        }
        compute(0);
        if (s + 1 < nst) {
            stage_store<BM, !TA>(As[1], tid, ra1);
            stage_store<BN, TB>(Bs[1], tid, rb1);
        }
        lds_barrier();
        if (s + 1 >= nst) break;
        // odd stage s + 1: buffer 1 computes, set 0 (stage s + 2) goes to buffer 0
        if (s + 3 < nst) {
            stage_load<BM, !TA>(A, lda, m0, kb + (s + 3) * GK, tid, ra1);
            stage_load<BN, TB>(B, ldb, n0, kb + (s + 3) * GK, tid, rb1);
        }
        compute(1);
        if (s + 2 < nst) {
            stage_store<BM, !TA>(As[0], tid, ra0);
            stage_store<BN, TB>(Bs[0], tid, rb0);
        }
        lds_barrier();
    }

    // the epilogue's operands (bias, Z) are loaded before its first store: the
    // compiler cannot rule out that a store aliases them, and would otherwise
    // issue each element's load only after the previous element's store
    float bv[TN], zv[TM][TN][4];
    if (!part) {
#pragma unroll
        for (int tj = 0; tj < TN; ++tj) {
            const int c = n0 + wn + tj * 16 + (lane & 15);
            if constexpr (EPI == ENC_EPI_BIAS || EPI == ENC_EPI_BIAS_GELU) bv[tj] = ep.bias[c];
#pragma unroll
            for (int ti = 0; ti < TM; ++ti)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = m0 + wm + ti * 16 + 4 * (lane >> 4) + e;
                    if constexpr (EPI != ENC_EPI_STORE && EPI != ENC_EPI_BIAS && EPI != ENC_EPI_BIAS_GELU)
                        zv[ti][tj][e] = ep.zin[static_cast<size_t>(r) * ep.ldc + c];
                }
        }
    }

#pragma unroll
    for (int ti = 0; ti < TM; ++ti)
#pragma unroll
        for (int tj = 0; tj < TN; ++tj)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = m0 + wm + ti * 16 + 4 * (lane >> 4) + e;
                const int c = n0 + wn + tj * 16 + (lane & 15);
                const float v = acc[ti][tj][e];
                if (part) {  // split-K partial (EPI_STORE only): dense M x N slab z
                    part[(static_cast<size_t>(blockIdx.z) * M + r) * N + c] = v;
                    continue;
                }
                const size_t o = static_cast<size_t>(r) * ep.ldc + c;
                if constexpr (EPI == ENC_EPI_STORE) {
                    ep.cf[o] = v;
                } else if constexpr (EPI == ENC_EPI_BIAS) {
                    ep.cf[o] = v + bv[tj];
                } else if constexpr (EPI == ENC_EPI_BIAS_GELU) {
                    const float z = v + bv[tj];
                    ep.cf[o] = z;
                    ep.ct[o] = bf16_rne(gelu(z));
                } else {
                    const float g = v * gelu_grad(zv[ti][tj][e]);
                    ep.cf[o] = g;
                    ep.ct[o] = bf16_rne(g);
                }
            }
}

// Split-K partial slabs of the MFMA GEMM (at most kSplitWs floats; calls that
// split K share it, so such calls must not run concurrently).
constexpr size_t kSplitWs = size_t(4) << 20;
__device__ __align__(256) float g_split_ws[kSplitWs];

// cf[r*ldc + c] = sum over z (in order) of slab z's (r, c).
__global__ __launch_bounds__(256) void k_splitk_reduce(const float* __restrict__ part, int S, int M, int N,
                                                       float* __restrict__ cf, int ldc) {
    const int64_t n = static_cast<int64_t>(M) * N;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        float t = 0.0f;
        for (int z = 0; z < S; ++z) t += part[z * n + e];
        const int r = static_cast<int>(e / N), c = static_cast<int>(e % N);
        cf[static_cast<size_t>(r) * ldc + c] = t;
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

// Column sums: one 1,024-thread block per 64 columns; thread (row phase p,
// column c) sums rows p, p + 16, ... in order, then phase 0 adds the 16
// partial sums in phase order (a fixed order: the result does not depend on
// scheduling).
__global__ __launch_bounds__(1024) void k_colsum(const float* X, int M, int N, int ld, float* out) {
    __shared__ float part[16][64];
    const int cx = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cx;
    float s = 0.0f;
    if (c < N) {
        // 8 rows per round in flight, added in row order
        int r = ph;
        for (; r + 7 * 16 < M; r += 8 * 16) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = X[static_cast<size_t>(r + 16 * j) * ld + c];
#pragma unroll
            for (int j = 0; j < 8; ++j) s += v[j];
        }
        for (; r < M; r += 16) s += X[static_cast<size_t>(r) * ld + c];
    }
    part[ph][cx] = s;
    __syncthreads();
    if (ph == 0 && c < N) {
        float t = 0.0f;
#pragma unroll
        for (int q = 0; q < 16; ++q) t += part[q][cx];
        out[c] = t;
    }
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

// Tile of the MFMA GEMM for an M x N x K product: the largest of 128 x 128,
// 128 x 64, 64 x 128 that divides M and N and still gives at least kMinTiles
// tiles, else 64 x 64; {0, 0} where K is not a multiple of 64 (the generic
// kernel then runs).
void mfma_tile(int M, int N, int K, int* bm, int* bn) {
    *bm = *bn = 0;
    if (K % GK) return;
    const int cand[3][2] = {{128, 128}, {128, 64}, {64, 128}};
    for (const auto& t : cand)
        if (M % t[0] == 0 && N % t[1] == 0 && static_cast<int64_t>(M / t[0]) * (N / t[1]) >= kMinTiles) {
            *bm = t[0];
            *bn = t[1];
            return;
        }
    *bm = *bn = 64;
}

// Device address of g_split_ws, looked up once per device.
float* split_ws() {
    static std::mutex mu;
    static float* addr[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_split_ws)) != hipSuccess) return nullptr;
        addr[dev] = static_cast<float*>(p);
    }
    return addr[dev];
}

// K splits of a plain-store product with fewer than kMinTiles tiles: the
// largest power of two <= 8 that keeps at least 4 stages per split, at most
// 1,024 workgroups and the slabs within the workspace.
int split_k(int epi, int tiles, int M, int N, int K) {
    if (epi != ENC_EPI_STORE || tiles >= kMinTiles) return 1;
    int s = 1;
    while (s < 8 && K % (GK * 2 * s) == 0 && K / (GK * 2 * s) >= 4 && tiles * 2 * s <= 1024 &&
           static_cast<size_t>(2 * s) * M * N <= kSplitWs)
        s *= 2;
    return s;
}

template <int BM, int BN, bool TA, bool TB>
int gemm_mfma(int epi, const bf16_t* a, const bf16_t* b, int M, int N, int K, int lda, int ldb, EpiArgs<bf16_t> ep,
              hipStream_t st) {
    const int tiles = (N / BN) * (M / BM);
    const int S = split_k(epi, tiles, M, N, K);
    float* part = nullptr;
    if (S > 1 && !(part = split_ws())) return fail(-3, "enc_gemm: split-K workspace address lookup failed");
    dim3 grid(N / BN, M / BM, S);
    const int kper = K / S;
    (void)hipGetLastError();
    switch (epi) {
    case ENC_EPI_STORE: hipLaunchKernelGGL((k_gemm_mfma<BM, BN, TA, TB, ENC_EPI_STORE>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep, kper, part); break;
    case ENC_EPI_BIAS: hipLaunchKernelGGL((k_gemm_mfma<BM, BN, TA, TB, ENC_EPI_BIAS>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep, kper, part); break;
    case ENC_EPI_BIAS_GELU: hipLaunchKernelGGL((k_gemm_mfma<BM, BN, TA, TB, ENC_EPI_BIAS_GELU>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep, kper, part); break;
    case ENC_EPI_GELU_BWD: hipLaunchKernelGGL((k_gemm_mfma<BM, BN, TA, TB, ENC_EPI_GELU_BWD>), grid, dim3(256), 0, st, a, b, M, N, K, lda, ldb, ep, kper, part); break;
    default: return fail(-1, "enc_gemm: unknown epilogue %d", epi);
    }
    if (S > 1) {
        const int64_t n = static_cast<int64_t>(M) * N;
        const unsigned g = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 2048));
        hipLaunchKernelGGL(k_splitk_reduce, dim3(g), dim3(256), 0, st, part, S, M, N, ep.cf, ep.ldc);
    }
    return launched("enc_gemm");
}

template <bool TA, bool TB>
int gemm_mfma_tile(int bm, int bn, int epi, const bf16_t* a, const bf16_t* b, int M, int N, int K, int lda, int ldb,
                   EpiArgs<bf16_t> ep, hipStream_t st) {
    if (bm == 128 && bn == 128) return gemm_mfma<128, 128, TA, TB>(epi, a, b, M, N, K, lda, ldb, ep, st);
    if (bm == 128) return gemm_mfma<128, 64, TA, TB>(epi, a, b, M, N, K, lda, ldb, ep, st);
    if (bn == 128) return gemm_mfma<64, 128, TA, TB>(epi, a, b, M, N, K, lda, ldb, ep, st);
    return gemm_mfma<64, 64, TA, TB>(epi, a, b, M, N, K, lda, ldb, ep, st);
}

template <typename T, bool TA, bool TB>
int gemm_t(int epi, const void* A, const void* B, int M, int N, int K, int lda, int ldb, EpiArgs<T> ep,
           hipStream_t st) {
    if constexpr (sizeof(T) == 2) {
        int bm, bn;
        mfma_tile(M, N, K, &bm, &bn);
        const bool aligned = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0) &&
                             lda % 8 == 0 && ldb % 8 == 0;
        if (bm && aligned)
            return gemm_mfma_tile<TA, TB>(bm, bn, epi, static_cast<const bf16_t*>(A), static_cast<const bf16_t*>(B), M,
                                          N, K, lda, ldb, ep, st);
    }
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

int enc_gemm_tile(int dtype, int epi, int M, int N, int K, int* bm, int* bn, int* splits) {
    g_err.clear();
    if (!bm || !bn || !splits) return fail(-1, "enc_gemm_tile: NULL output");
    *bm = *bn = 0;
    *splits = 1;
    if (M <= 0 || N <= 0 || K <= 0 || M % BM || N % BN || K % BK)
        return fail(-1, "enc_gemm_tile: M, N must be positive multiples of 64 and K of 32 (got %d, %d, %d)", M, N, K);
    if (dtype == ENC_BF16) mfma_tile(M, N, K, bm, bn);
    if (*bm) *splits = split_k(epi, (M / *bm) * (N / *bn), M, N, K);
    return 0;
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
    hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64), dim3(1024), 0, static_cast<hipStream_t>(stream), X, M, N, ld,
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

int enc_train_step(const enc_mlp* n, const enc_mlp_work* w, const void* X, const float* T, const float* colw, int B,
                   float lr, float beta1, float beta2, float eps, int32_t t, double* loss, void* stream) {
    g_err.clear();
    if (!n || !w || !X || !T || !colw || !loss) return fail(-1, "enc_train_step: NULL argument");
    const int dt = n->dtype, di = n->d_in, hd = n->hid, dout = n->d_out;
    int rc;
    // forward: Z = X W1 + b1, H = gelu(Z); Y = H W2 + b2
    if ((rc = enc_gemm(dt, 0, 0, ENC_EPI_BIAS_GELU, X, n->W1c, B, hd, di, di, hd, w->Z, w->H, n->b1, nullptr, hd,
                       stream)))
        return rc;
    if ((rc = enc_gemm(dt, 0, 0, ENC_EPI_BIAS, w->H, n->W2c, B, dout, hd, hd, dout, w->Y, nullptr, n->b2, nullptr,
                       dout, stream)))
        return rc;
    if ((rc = enc_loss_grad(dt, w->Y, T, colw, w->dyf, w->dyt, B, dout, dout, loss, stream))) return rc;
    // dW2 = H^T dY, db2 = colsum dY; dZ = (dY W2^T) gelu'(Z); dW1 = X^T dZ, db1 = colsum dZ
    if ((rc = enc_gemm(dt, 1, 0, ENC_EPI_STORE, w->H, w->dyt, hd, dout, B, hd, dout, w->gW2, nullptr, nullptr, nullptr,
                       dout, stream)))
        return rc;
    if ((rc = enc_colsum(w->dyf, B, dout, dout, w->gb2, stream))) return rc;
    if ((rc = enc_gemm(dt, 0, 1, ENC_EPI_GELU_BWD, w->dyt, n->W2c, B, hd, dout, dout, dout, w->dzf, w->dzt, nullptr,
                       w->Z, hd, stream)))
        return rc;
    if ((rc = enc_gemm(dt, 1, 0, ENC_EPI_STORE, X, w->dzt, di, hd, B, di, hd, w->gW1, nullptr, nullptr, nullptr, hd,
                       stream)))
        return rc;
    if ((rc = enc_colsum(w->dzf, B, hd, hd, w->gb1, stream))) return rc;
    float* ws[4] = {n->W1, n->b1, n->W2, n->b2};
    float* gs[4] = {w->gW1, w->gb1, w->gW2, w->gb2};
    void* cs[4] = {n->W1c, n->b1, n->W2c, n->b2};
    const int dts[4] = {dt, ENC_F32, dt, ENC_F32};
    const int64_t sz[4] = {static_cast<int64_t>(di) * hd, hd, static_cast<int64_t>(hd) * dout, dout};
    for (int k = 0; k < 4; ++k)
        if ((rc = enc_adam(dts[k], ws[k], gs[k], n->m[k], n->v[k], cs[k], sz[k], lr, beta1, beta2, eps, t, stream)))
            return rc;
    return 0;
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
