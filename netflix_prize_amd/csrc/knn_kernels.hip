// KnnfModel, KnnsModel and KnnModel (`__knn-d` form) on the GPU (src/knnf.rs,
// src/knns.rs, src/knn.rs): the item-item cosine similarity matrix of a factor
// table by FP32 MFMA, the co-occurrence support and its shrunk lift, and the
// per-query temporal top-k weighted mean of the user's residuals (or of a
// per-pair value).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "knn_launch.hpp"

namespace knn {

typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// S = max(0, F F^T / (n_i n_j)) over the padded Ip x Ip matrix (rows of F are
// items, fp >= f columns zero-padded, Ip a multiple of 64).  256 threads =
// 4 waves per 64 x 64 tile of S, each wave a 32 x 32 quarter as 2 x 2 MFMA
// 16 x 16 x 4 blocks; K advances 32 at a time through LDS.  Operand layout of
// v_mfma_f32_16x16x4f32 (wave64, lane l): A[l%16][l/16], B[l/16][l%16],
// result D[4*(l/16)+e][l%16].  A zero norm gives similarity 0, as
// KnnfModel::sim (src/knnf.rs:46-52).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cosine(const float* __restrict__ F, int fp, const float* __restrict__ norm,
                                                float* __restrict__ S, int ip) {
    constexpr int BM = 64, BK = 32;
    __shared__ float As[BM][BK + 1];
    __shared__ float Bs[BM][BK + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BM;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < fp; k0 += BK) {
#pragma unroll
        for (int r = 0; r < (BM * BK) / 256; ++r) {
            const int e = tid + 256 * r;
            const int k = e % BK, m = e / BK;  // consecutive threads: consecutive columns of a row
            As[m][k] = F[static_cast<size_t>(m0 + m) * fp + k0 + k];
            Bs[m][k] = F[static_cast<size_t>(n0 + m) * fp + k0 + k];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            float a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = As[wm + t * 16 + (lane & 15)][ks + (lane >> 4)];
                b[t] = Bs[wn + t * 16 + (lane & 15)][ks + (lane >> 4)];
            }
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
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
                const float nr = norm[r], nc = norm[c];
                float v = 0.0f;
                if (nr != 0.0f && nc != 0.0f) {
                    v = acc[ti][tj][e] / (nr * nc);
                    v = v > 0.0f ? v : 0.0f;
                }
                S[static_cast<size_t>(r) * ip + c] = v;
            }
}

// Row norms sqrt(sum f^2), one wave per row.
__global__ __launch_bounds__(256) void k_norms(const float* __restrict__ F, int fp, int rows, float* __restrict__ norm) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= rows) return;
    float s = 0.0f;
    for (int k = lane; k < fp; k += 64) {
        const float x = F[static_cast<size_t>(wave) * fp + k];
        s += x * x;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) norm[wave] = __builtin_sqrtf(s);
}

// ---------------------------------------------------------------------------
// KnnfModel::predict (src/knnf.rs:101-141), one wave per query (u, i, day).
// Candidates are the user's training ratings t (item j, residual r, day d_j)
// with j != i and weight w = sim(i, j)^scaling / (1 + tau |day - d_j|) > 0;
// the k largest weights are kept and the prediction is sum w r / sum w (0
// without candidates).  The k-th largest weight is found exactly by a radix
// select over the weights' bit patterns (positive floats order as unsigned
// integers), 8 bits per pass through a 256-bin LDS histogram.  Equal weights
// at the k-th place follow the reference's select_nth_unstable_by: with k + 1
// candidates it drops the first smallest (core::slice::select, index ==
// len - 1: max_index keeps the first of equals), with up to 16 its insertion
// sort keeps equal weights in rating order, which is also taken beyond 16
// (there the reference's quickselect partitioning decides).
// ---------------------------------------------------------------------------
constexpr int kWCache = 32; // candidate weights per lane kept in registers

__device__ __forceinline__ float knn_weight(const KnnView& v, const float* srow, int32_t i, float day, uint32_t t) {
    const int32_t j = v.t_item[t];
    if (j == i) return 0.0f;
    float sim = srow[j];
    if (v.scaling != 1.0f) sim = __builtin_powf(sim, v.scaling);
    const float dt = __builtin_fabsf(day - v.t_day[t]);
    return sim / (1.0f + v.tau * dt);
}

// the value averaged for candidate t: its residual, or D[i][j] (KnnModel use_stat1)
__device__ __forceinline__ float knn_target(const KnnView& v, int32_t i, uint32_t t) {
    return v.D ? v.D[static_cast<size_t>(i) * v.ip + v.t_item[t]] : v.t_res[t];
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x += static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), o, 64));
    return x;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), o, 64));
        x = y < x ? y : x;
    }
    return x;
}

__device__ __forceinline__ float wave_sum_f32(float x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__global__ __launch_bounds__(256) void k_knnf_predict(KnnView v, int64_t n, const int32_t* __restrict__ us,
                                                      const int32_t* __restrict__ is, const int16_t* __restrict__ ds,
                                                      float* __restrict__ out) {
    __shared__ uint32_t hist_all[4][256];
    const int lane = threadIdx.x & 63, wslot = threadIdx.x >> 6;
    uint32_t* hist = hist_all[wslot];
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + wslot; q < n; q += nwaves) {
        const int32_t u = us[q], i = is[q];
        const float day = static_cast<float>(ds[q]);
        const uint32_t t0 = v.t_off[u], t1 = v.t_off[u + 1];
        const float* srow = v.S + static_cast<size_t>(i) * v.ip;
        // a list of up to 64 kWCache ratings keeps its weights and residuals in
        // registers (kWCache per lane) for all passes; a longer one recomputes
        // them per pass
        const bool cached = t1 - t0 <= 64u * kWCache;
        float wc[kWCache], rc[kWCache];
#pragma unroll
        for (int k = 0; k < kWCache; ++k) {
            const uint32_t t = t0 + lane + 64u * k;
            wc[k] = 0.0f;
            rc[k] = 0.0f;
            if (cached && t < t1) {
                wc[k] = knn_weight(v, srow, i, day, t);
                rc[k] = knn_target(v, i, t);
            }
        }
        // f(w) for every candidate weight of this lane (rating order not kept)
        auto visit = [&](auto&& f) {
            if (cached) {
#pragma unroll
                for (int k = 0; k < kWCache; ++k)
                    if (t0 + lane + 64u * k < t1) f(wc[k]);
            } else {
                for (uint32_t t = t0 + lane; t < t1; t += 64) f(knn_weight(v, srow, i, day, t));
            }
        };
        // positive candidates
        uint32_t m = 0;
        visit([&](float w) { m += w > 0.0f ? 1u : 0u; });
        m = wave_sum_u32(m);
        uint32_t thr = 0, take_eq = 0xffffffffu;  // keep w > thr, and the first take_eq weights == thr
        bool drop_first = false;                   // k + 1 candidates: keep all but the first smallest
        if (m == static_cast<uint32_t>(v.k) + 1u) {
            uint32_t mn = 0xffffffffu;
            visit([&](float w) {
                if (w > 0.0f && __float_as_uint(w) < mn) mn = __float_as_uint(w);
            });
            thr = wave_min_u32(mn);
            drop_first = true;
        } else if (m > static_cast<uint32_t>(v.k)) {
            uint32_t prefix = 0, mask = 0, krem = static_cast<uint32_t>(v.k);
            for (int shift = 24; shift >= 0; shift -= 8) {
                for (int b = lane; b < 256; b += 64) hist[b] = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                visit([&](float w) {
                    const uint32_t bits = __float_as_uint(w);
                    if (w > 0.0f && (bits & mask) == prefix) atomicAdd(&hist[(bits >> shift) & 255u], 1u);
                });
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // lane l holds bins 255-4l .. 252-4l (descending); inclusive scan in lane order
                uint32_t c[4], own = 0;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    c[e] = hist[255 - 4 * lane - e];
                    own += c[e];
                }
                uint32_t incl = own;
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t up = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), o, 64));
                    if (lane >= o) incl += up;
                }
                const uint64_t hit = __ballot(incl >= krem);
                const int cl = __builtin_ctzll(hit);  // first lane whose bins reach the k-th weight
                uint32_t digit = 0, above = 0;
                if (lane == cl) {
                    uint32_t cum = incl - own;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (cum + c[e] >= krem) {
                            digit = 255u - 4u * static_cast<uint32_t>(lane) - static_cast<uint32_t>(e);
                            above = cum;
                            break;
                        }
                        cum += c[e];
                    }
                }
                digit = static_cast<uint32_t>(__shfl(static_cast<int>(digit), cl, 64));
                above = static_cast<uint32_t>(__shfl(static_cast<int>(above), cl, 64));
                krem -= above;
                prefix |= digit << shift;
                mask |= 255u << shift;
            }
            thr = prefix;
            take_eq = krem;
        }
        // weighted mean over the kept candidates, batches of 64 in rating order
        float num = 0.0f, den = 0.0f;
        uint32_t eq_seen = 0;
        auto batch = [&](float w, float r) {
            const uint32_t bits = __float_as_uint(w);
            bool keep = w > 0.0f && (m <= static_cast<uint32_t>(v.k) || bits > thr);
            const bool eq = w > 0.0f && m > static_cast<uint32_t>(v.k) && bits == thr;
            const uint64_t eqm = __ballot(eq);
            const uint32_t rank = eq_seen + static_cast<uint32_t>(__builtin_popcountll(eqm & ((1ull << lane) - 1ull)));
            if (eq && (drop_first ? rank >= 1u : rank < take_eq)) keep = true;
            eq_seen += static_cast<uint32_t>(__builtin_popcountll(eqm));
            if (keep) {
                num += w * r;
                den += w;
            }
        };
        if (cached) {
#pragma unroll
            for (int k = 0; k < kWCache; ++k)
                if (t0 + 64u * k < t1) batch(wc[k], rc[k]);  // wc, rc are 0 past the list end
        } else {
            for (uint32_t tb = t0; tb < t1; tb += 64) {
                const uint32_t t = tb + lane;
                float w = 0.0f, r = 0.0f;
                if (t < t1) {
                    w = knn_weight(v, srow, i, day, t);
                    r = knn_target(v, i, t);
                }
                batch(w, r);
            }
        }
        num = wave_sum_f32(num);
        den = wave_sum_f32(den);
        if (lane == 0) out[q] = m == 0 ? 0.0f : num / den;
    }
}

// ---------------------------------------------------------------------------
// KnnsModel (src/knns.rs): the support matrix of build_bin_nbstats(.., "supp")
// (src/nbstats.rs:73-117): supp[i][j] += 1 for every ordered pair of items of
// each user's list (diagonal included).  The matrix is symmetric and each
// user's list is sorted by item, so the entry of item i at list position t
// contributes the pairs (t, s), s >= t, of its list tail to the upper
// triangle (an equal item at s > t stands for both orders and adds 2);
// k_mirror then copies the upper triangle down.
// Row-owned counting: the list entries are indexed by item (k_item_hist,
// k_scan_offsets, k_item_scatter: the position t and the list end of each
// entry), then one block per row i counts the tails of i's entries with LDS
// atomics over a kRowChunk-wide window of columns and writes the row once.
// Integer counts, converted to FP32 exactly below 2^24.
// ---------------------------------------------------------------------------
constexpr int kRowChunk = 20480;  // columns per LDS window (80 KB)
constexpr int kRowThreads = 1024;  // threads per row block

__global__ __launch_bounds__(256) void k_item_hist(const int32_t* __restrict__ items, int64_t n_rows,
                                                   uint32_t* __restrict__ icnt) {
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n_rows;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x)
        atomicAdd(&icnt[items[t] + 1], 1u);
}

// exclusive offsets in place over icnt[0..n] (icnt[0] = 0 on entry), copied to cursor
__global__ __launch_bounds__(1024) void k_scan_offsets(uint32_t* __restrict__ icnt, uint32_t* __restrict__ cursor,
                                                       int n) {
    __shared__ uint32_t part[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1 + 1023) / 1024;
    const int a = tid * per, b = a + per < n + 1 ? a + per : n + 1;
    uint32_t sum = 0;
    for (int k = a; k < b; ++k) sum += icnt[k];
    part[tid] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t add = tid >= o ? part[tid - o] : 0u;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    uint32_t run = part[tid] - sum;  // exclusive prefix of this thread's chunk
    for (int k = a; k < b; ++k) {
        run += icnt[k];
        icnt[k] = run;  // inclusive over icnt[0..k] = offset of item k (icnt[k] counts item k - 1)
        cursor[k] = run;
    }
}

// one wave per user: each list entry lands in its item's range with its position and list end
__global__ __launch_bounds__(256) void k_item_scatter(const uint32_t* __restrict__ off, const int32_t* __restrict__ items,
                                                      int32_t n_users, uint32_t* __restrict__ cursor,
                                                      uint32_t* __restrict__ ient, uint32_t* __restrict__ iend,
                                                      float* __restrict__ work) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
    for (int64_t u = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); u < n_users;
         u += nwaves) {
        const uint32_t e = off[u + 1];
        for (uint32_t t = off[u] + lane; t < e; t += 64) {
            const int32_t it = items[t];
            const uint32_t p = atomicAdd(&cursor[it], 1u);
            ient[p] = t;
            iend[p] = e;
            atomicAdd(&work[it], static_cast<float>(e - t));  // the row's pair adds (ordering only)
        }
    }
}

__global__ __launch_bounds__(kRowThreads) void k_supp_rows(const int32_t* __restrict__ items, const uint32_t* __restrict__ ioff,
                                                   const uint32_t* __restrict__ ient, const uint32_t* __restrict__ iend,
                                                   const int32_t* __restrict__ order, float* __restrict__ supp, int ip) {
    __shared__ uint32_t cnt[kRowChunk];
    // rows with the most pair adds first: their blocks set the kernel's length
    const int i = order[blockIdx.x], tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t q0 = ioff[i], q1 = ioff[i + 1];
    float* row = supp + static_cast<size_t>(i) * ip;
    for (int c0 = 0; c0 < ip; c0 += kRowChunk) {
        const int w = ip - c0 < kRowChunk ? ip - c0 : kRowChunk;
        for (int j = tid; j < w; j += kRowThreads) cnt[j] = 0u;
        __syncthreads();
        if (c0 + w > i) {  // the tails hold items >= i only
            for (uint32_t q = q0 + wv; q < q1; q += kRowThreads / 64) {
                const uint32_t t = ient[q], e = iend[q];
                for (uint32_t s = t + lane; s < e; s += 64) {
                    const int j = items[s] - c0;
                    if (j >= 0 && j < w) atomicAdd(&cnt[j], (j + c0 == i && s != t) ? 2u : 1u);
                }
            }
        }
        __syncthreads();
        for (int j = tid; j < w; j += kRowThreads) row[c0 + j] = static_cast<float>(cnt[j]);
        __syncthreads();
    }
}

// Lower triangle from the upper one, 32 x 32 tiles through LDS (tile (bx, by)
// with bx <= by reads upper tile (bx, by) and writes its transpose at (by, bx)).
__global__ __launch_bounds__(256) void k_mirror(float* __restrict__ S, int ip) {
    __shared__ float t[32][33];
    const int bx = blockIdx.x, by = blockIdx.y;
    if (bx > by) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) t[r][tx] = S[static_cast<size_t>(bx * 32 + r) * ip + by * 32 + tx];
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int row = by * 32 + r, col = bx * 32 + tx;  // lower element (row > col) or the diagonal tile's
        if (row > col) S[static_cast<size_t>(row) * ip + col] = t[tx][r];
    }
}

__global__ __launch_bounds__(256) void k_diag(const float* __restrict__ supp, int ip, float* __restrict__ diag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ip) diag[i] = supp[static_cast<size_t>(i) * ip + i];
}

// KnnsModel::sim (src/knns.rs:64-70), in place over the support matrix:
// phi = n * N_USERS / (supp_ii supp_jj) (0 if the product is 0), then
// phi * n / (n + shrinkage).  N_USERS is the reference's constant 480,189.
__global__ __launch_bounds__(256) void k_knns_sim(float* __restrict__ S, const float* __restrict__ diag, int ip,
                                                  float shrinkage) {
#pragma clang fp contract(off)
    const int64_t total = static_cast<int64_t>(ip) * ip;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(e / ip), j = static_cast<int>(e % ip);
        const float n = S[e];
        const float den = diag[i] * diag[j];
        const float phi = den > 0.0f ? n * 480189.0f / den : 0.0f;
        S[e] = phi * n / (n + shrinkage);
    }
}

size_t support_scratch_words(int64_t n_rows, int ip) {
    return 2 * (static_cast<size_t>(ip) + 1) + 2 * static_cast<size_t>(n_rows) + 2 * static_cast<size_t>(ip);
}

// Blocks on the stream once: the rows' work is read back to order the rows.
hipError_t launch_support(const uint32_t* off, const int32_t* items, int32_t n_users, int64_t n_rows, float* S,
                          float* diag, int ip, uint32_t* scratch, hipStream_t st) {
    (void)hipGetLastError();
    uint32_t* icnt = scratch;
    uint32_t* cursor = icnt + ip + 1;
    uint32_t* ient = cursor + ip + 1;
    uint32_t* iend = ient + n_rows;
    float* work = reinterpret_cast<float*>(iend + n_rows);
    int32_t* order = reinterpret_cast<int32_t*>(work + ip);
    hipError_t e = hipMemsetAsync(icnt, 0, sizeof(uint32_t) * (static_cast<size_t>(ip) + 1), st);
    if (e == hipSuccess) e = hipMemsetAsync(work, 0, sizeof(float) * static_cast<size_t>(ip), st);
    if (e != hipSuccess) return e;
    int64_t blocks = (n_rows + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks > 0) hipLaunchKernelGGL(k_item_hist, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, items, n_rows, icnt);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, st, icnt, cursor, ip);
    int64_t ublocks = (static_cast<int64_t>(n_users) + 3) / 4;
    if (ublocks > 65536) ublocks = 65536;
    if (ublocks > 0)
        hipLaunchKernelGGL(k_item_scatter, dim3(static_cast<unsigned>(ublocks)), dim3(256), 0, st, off, items, n_users,
                           cursor, ient, iend, work);
    std::vector<float> w(static_cast<size_t>(ip));
    e = hipMemcpyAsync(w.data(), work, sizeof(float) * w.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    std::vector<int32_t> rows(static_cast<size_t>(ip));
    for (int r = 0; r < ip; ++r) rows[r] = r;
    std::stable_sort(rows.begin(), rows.end(), [&](int32_t a, int32_t b) { return w[a] > w[b]; });
    e = hipMemcpyAsync(order, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_supp_rows, dim3(ip), dim3(kRowThreads), 0, st, items, icnt, ient, iend, order, S, ip);
    hipLaunchKernelGGL(k_mirror, dim3(ip / 32, ip / 32), dim3(256), 0, st, S, ip);
    hipLaunchKernelGGL(k_diag, dim3((ip + 255) / 256), dim3(256), 0, st, S, ip, diag);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // rows is a host temporary
    return e;
}

// KnnModel (src/knn.rs:321-334): diff = stat1 / (stat0 + shrinkage), in place over stat1
__global__ __launch_bounds__(256) void k_stat_diff(const float* __restrict__ S0, float* __restrict__ D, int ip,
                                                   float shrinkage) {
#pragma clang fp contract(off)
    const int64_t total = static_cast<int64_t>(ip) * ip;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        D[e] = D[e] / (S0[e] + shrinkage);
}

hipError_t launch_stat_sim(float* S, float* D, float* diag, int ip, float shrinkage, hipStream_t st) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_stat_diff, dim3(65536), dim3(256), 0, st, S, D, ip, shrinkage);
    hipLaunchKernelGGL(k_diag, dim3((ip + 255) / 256), dim3(256), 0, st, S, ip, diag);
    // KnnModel::support_sim (src/knn.rs:105-113) is KnnsModel::sim's formula over stat0
    hipLaunchKernelGGL(k_knns_sim, dim3(65536), dim3(256), 0, st, S, diag, ip, shrinkage);
    return hipGetLastError();
}

hipError_t launch_knns_sim(const uint32_t* off, const int32_t* items, int32_t n_users, int64_t n_rows, float* S,
                           float* diag, int ip, float shrinkage, uint32_t* scratch, hipStream_t st) {
    hipError_t e = launch_support(off, items, n_users, n_rows, S, diag, ip, scratch, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_knns_sim, dim3(65536), dim3(256), 0, st, S, diag, ip, shrinkage);
    return hipGetLastError();
}

hipError_t launch_norms(const float* F, int fp, int rows, float* norm, hipStream_t st) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_norms, dim3((rows + 3) / 4), dim3(256), 0, st, F, fp, rows, norm);
    return hipGetLastError();
}

hipError_t launch_cosine(const float* F, int fp, const float* norm, float* S, int ip, hipStream_t st) {
    if (ip % 64 != 0 || fp % 32 != 0) return hipErrorInvalidValue;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_cosine, dim3(ip / 64, ip / 64), dim3(256), 0, st, F, fp, norm, S, ip);
    return hipGetLastError();
}

hipError_t launch_knnf_predict(const KnnView& v, int64_t n, const int32_t* us, const int32_t* is, const int16_t* ds,
                               float* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_knnf_predict, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, v, n, us, is, ds, out);
    return hipGetLastError();
}

}  // namespace knn
