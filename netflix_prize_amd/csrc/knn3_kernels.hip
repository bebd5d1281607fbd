// Knn3Model on the GPU (src/knn3.rs): per item, the neighbour candidates by
// shrunk lift of co-rating counts, the FP64 normal equations of the two
// regression signals over the item's users, the non-negative least-squares
// weights, and the batched prediction sum.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "knn_launch.hpp"

namespace knn {

namespace {

constexpr int kMaxK = 64;       // candidates per item (one wave of lanes)
constexpr int kMapMax = 32768;  // items a k3_cov LDS map covers (one byte each)
constexpr int kCh = 32;         // users staged per k3_cov round
constexpr int kSlots = kCh / 4; // users per wave per round
constexpr uint32_t kScanMax = 256;  // longer user lists are binary-searched

__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// src/knn3.rs:201-209, FP32 as written: phi = n_mj N_USERS / (n_m n_j),
// prox = phi n_mj / (n_mj + shrinkage).  N_USERS is the reference's 480,189.
__device__ __forceinline__ float knn3_prox(float nmj, float nm, float nj, float shrink) {
#pragma clang fp contract(off)
    const float phi = nmj * 480189.0f / (nm * nj);
    return phi * nmj / (nmj + shrink);
}

// Candidate j of item m as an order-preserving key (prox > 0, so its bits),
// 0 when j is no candidate: j == m, no co-rating, or no users (:197-205).
__device__ __forceinline__ uint32_t cand_key(const Knn3View& v, const float* row, int m, float nm, int j) {
    const float c = row[j];
    if (j == m || !(c > 0.0f)) return 0u;
    const float nj = v.diag[j];
    if (!(nj >= 1.0f)) return 0u;
    return __float_as_uint(knn3_prox(c, nm, nj, v.shrinkage));
}

__device__ __forceinline__ double wave_min_f64(double x) {
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(x, o, 64);
        x = y < x ? y : x;
    }
    return x;
}

}  // namespace

// ---------------------------------------------------------------------------
// Candidates of item m (src/knn3.rs:195-216), one 256-thread block per item:
// K = clamp(#{prox >= threshold}, k_min, k_max), at most the candidate count;
// the K largest prox by a 4 x 8-bit radix select of the K-th key, equal keys
// at the K-th place taken by smaller item index; then ordered by prox, largest
// first, smaller index first among equals.  The reference orders with an
// unstable sort, so the order among equal prox is this build's choice.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k3_select(Knn3View v) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_len, s_above, s_digit, s_need, s_ngt;
    __shared__ uint32_t wave_cnt[4];
    __shared__ uint32_t sel_key[kMaxK];
    __shared__ int32_t sel_idx[kMaxK];
    const int m = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const float nm = v.diag[m];
    if (!(nm >= 1.0f)) {
        if (tid == 0) v.nb_cnt[m] = 0;
        return;
    }
    const float* row = v.cnt + static_cast<size_t>(m) * v.ip;
    if (tid == 0) {
        s_len = 0;
        s_above = 0;
        s_ngt = 0;
    }
    __syncthreads();
    uint32_t len = 0, above = 0;
    for (int j = tid; j < v.n_items; j += 256) {
        const uint32_t k = cand_key(v, row, m, nm, j);
        if (k) {
            ++len;
            if (__uint_as_float(k) >= v.threshold) ++above;
        }
    }
    atomicAdd(&s_len, len);
    atomicAdd(&s_above, above);
    __syncthreads();
    uint32_t K = s_above > static_cast<uint32_t>(v.k_min) ? s_above : static_cast<uint32_t>(v.k_min);
    if (K > static_cast<uint32_t>(v.k_max)) K = static_cast<uint32_t>(v.k_max);
    if (K > s_len) K = s_len;
    if (K == 0) {
        if (tid == 0) v.nb_cnt[m] = 0;
        return;
    }
    // the K-th largest key: digits from the top, counting from the largest bin
    uint32_t prefix = 0, mask = 0, need = K;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[tid] = 0;
        __syncthreads();
        for (int j = tid; j < v.n_items; j += 256) {
            const uint32_t k = cand_key(v, row, m, nm, j);
            if (k && (k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t cum = 0;
            int d = 255;
            for (; d > 0; --d) {
                if (cum + hist[d] >= need) break;
                cum += hist[d];
            }
            s_digit = static_cast<uint32_t>(d);
            s_need = need - cum;
        }
        __syncthreads();
        prefix |= s_digit << shift;
        mask |= 255u << shift;
        need = s_need;
        __syncthreads();
    }
    const uint32_t T = prefix;
    // every key above T, in any order (sorted below)
    for (int j = tid; j < v.n_items; j += 256) {
        const uint32_t k = cand_key(v, row, m, nm, j);
        if (k > T) {
            const uint32_t p = atomicAdd(&s_ngt, 1u);
            sel_key[p] = k;
            sel_idx[p] = j;
        }
    }
    __syncthreads();
    // the first `need` keys equal to T by item index
    const uint32_t base = K - need;
    uint32_t taken = 0;
    for (int j0 = 0; j0 < v.n_items && taken < need; j0 += 256) {
        const int j = j0 + tid;
        const bool eq = j < v.n_items && cand_key(v, row, m, nm, j) == T;
        const uint64_t bal = __ballot(eq);
        if (lane == 0) wave_cnt[wv] = static_cast<uint32_t>(__builtin_popcountll(bal));
        __syncthreads();
        uint32_t rank = static_cast<uint32_t>(__builtin_popcountll(bal & lanes_below(lane)));
        uint32_t total = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wv) rank += wave_cnt[w];
            total += wave_cnt[w];
        }
        if (eq && taken + rank < need) {
            sel_key[base + taken + rank] = T;
            sel_idx[base + taken + rank] = j;
        }
        taken += total;
        __syncthreads();
    }
    __syncthreads();
    if (wv == 0) {
        const bool live = lane < static_cast<int>(K);
        const uint32_t mk = live ? sel_key[lane] : 0u;
        const int32_t mi = live ? sel_idx[lane] : 0;
        int rank = 0;
        for (int o = 0; o < static_cast<int>(K); ++o) {
            const uint32_t ko = sel_key[o];
            const int32_t io = sel_idx[o];
            rank += (ko > mk || (ko == mk && io < mi)) ? 1 : 0;
        }
        if (live) v.nb_item[static_cast<size_t>(m) * kMaxK + rank] = mi;
        if (lane == 0) v.nb_cnt[m] = static_cast<int32_t>(K);
    }
}

// ---------------------------------------------------------------------------
// Normal equations of item m (src/knn3.rs:222-286), one 256-thread block per
// item.  Users of m are staged kCh at a time as dense rows over the K
// candidates (zero where the user has no rating), err and r_prime; thread
// (a = lane, wave g) owns C[a][16g .. 16g+15] of both signals, wave 0 also
// d[a].  Every sum runs over the users in ascending order of exact FP64
// products of FP32 values, so it equals the reference's sum bit for bit (its
// skipped users add nothing; here they add +0).  A user list longer than
// kScanMax is probed by binary search per candidate (a wave's long lists
// searched together), a shorter one scanned through an LDS map item ->
// candidate.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k3_cov(Knn3View v) {
#pragma clang fp contract(off)
    __shared__ uint8_t map[kMapMax];
    __shared__ int32_t nbi[kMaxK];
    __shared__ float Xe[kCh][kMaxK], Xr[kCh][kMaxK];
    __shared__ float me[kCh], mr[kCh];
    // items with the most users first: their blocks set the kernel's length
    const int m = v.heavy_first[blockIdx.x], tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int K = v.nb_cnt[m];
    if (K == 0) return;
    for (int j = tid; j < v.ip; j += 256) map[j] = 0xFF;
    if (tid < K) nbi[tid] = v.nb_item[static_cast<size_t>(m) * kMaxK + tid];
    __syncthreads();
    if (tid < K) map[nbi[tid]] = static_cast<uint8_t>(tid);
    const bool cols = wv * 16 < K;
    double ce[16], cb[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) ce[b] = cb[b] = 0.0;
    double de = 0.0, db = 0.0;
    const uint32_t b0 = v.i_off[m], b1 = v.i_off[m + 1];
    for (uint32_t base = b0; base < b1; base += kCh) {
        const int nr = b1 - base < static_cast<uint32_t>(kCh) ? static_cast<int>(b1 - base) : kCh;
        for (int e = tid; e < kCh * kMaxK; e += 256) {
            (&Xe[0][0])[e] = 0.0f;
            (&Xr[0][0])[e] = 0.0f;
        }
        if (tid < nr) {
            me[tid] = v.i_err[base + tid];
            mr[tid] = v.i_rp[base + tid];
        }
        __syncthreads();
        // the wave's users r = wv, wv + 4, ...: short lists scanned one after the
        // other, long ones searched together (lane a looks for candidate a in
        // every long list at once, kSlots independent load chains)
        uint32_t lo[kSlots], hi[kSlots], en[kSlots];
        const int32_t want = lane < K ? nbi[lane] : 0;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
            const int r = wv + 4 * k;
            lo[k] = hi[k] = en[k] = 0;
            if (r < nr) {
                const int32_t u = v.i_user[base + r];
                const uint32_t s = v.u_off[u], e = v.u_off[u + 1];
                if (e - s > kScanMax) {
                    en[k] = e;
                    if (lane < K) {
                        lo[k] = s;
                        hi[k] = e;
                    }
                } else {
                    // 256 entries a batch, every load unconditional (clamped into the list)
                    for (uint32_t t0 = s; t0 < e; t0 += 256) {
                        int32_t it[4];
                        float ee[4], rr[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const uint32_t t = t0 + q * 64 + lane;
                            const uint32_t tc = t < e ? t : s;
                            it[q] = v.u_item[tc];
                            ee[q] = v.u_err[tc];
                            rr[q] = v.u_rp[tc];
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const uint8_t a = map[it[q]];
                            if (t0 + q * 64 + lane < e && a != 0xFF) {
                                Xe[r][a] = ee[q];
                                Xr[r][a] = rr[q];
                            }
                        }
                    }
                }
            }
        }
        for (;;) {
            bool live = false;
#pragma unroll
            for (int k = 0; k < kSlots; ++k) live |= lo[k] < hi[k];
            if (!__ballot(live)) break;
            int32_t x[kSlots];
#pragma unroll
            for (int k = 0; k < kSlots; ++k) x[k] = v.u_item[lo[k] < hi[k] ? (lo[k] + hi[k]) >> 1 : 0u];
#pragma unroll
            for (int k = 0; k < kSlots; ++k) {
                if (lo[k] < hi[k]) {
                    const uint32_t mid = (lo[k] + hi[k]) >> 1;
                    if (x[k] < want) lo[k] = mid + 1;
                    else hi[k] = mid;
                }
            }
        }
        if (lane < K) {
#pragma unroll
            for (int k = 0; k < kSlots; ++k) {
                const uint32_t t = lo[k] < en[k] ? lo[k] : 0u;
                const bool hit = lo[k] < en[k] && v.u_item[t] == want;
                if (hit) {
                    Xe[wv + 4 * k][lane] = v.u_err[t];
                    Xr[wv + 4 * k][lane] = v.u_rp[t];
                }
            }
        }
        __syncthreads();
        if (cols) {
            for (int r = 0; r < nr; ++r) {
                const double xa = Xe[r][lane], xra = Xr[r][lane];
#pragma unroll
                for (int b = 0; b < 16; ++b) {
                    ce[b] = fma(xa, static_cast<double>(Xe[r][wv * 16 + b]), ce[b]);
                    cb[b] = fma(xra, static_cast<double>(Xr[r][wv * 16 + b]), cb[b]);
                }
                if (wv == 0) {
                    de = fma(xa, static_cast<double>(me[r]), de);
                    db = fma(xra, static_cast<double>(mr[r]), db);
                }
            }
        }
        __syncthreads();
    }
    // M = x C_err + (1 - x) C_bias + reg I, rhs = x d_err + (1 - x) d_bias
    if (cols && lane < K) {
        const double x = v.x, y = 1.0 - v.x;
        double* srow = v.sys + static_cast<size_t>(m) * kMaxK * kMaxK + static_cast<size_t>(lane) * kMaxK;
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const int col = wv * 16 + b;
            if (col < K) {
                double s = x * ce[b] + y * cb[b];
                if (col == lane) s = s + v.reg;
                srow[col] = s;
            }
        }
        if (wv == 0) v.rhs[static_cast<size_t>(m) * kMaxK + lane] = x * de + y * db;
    }
}

// ---------------------------------------------------------------------------
// solve_nnls (src/knn3.rs:79-120) for item m, one wave: on the active set,
// lane r owns row r; left-looking Cholesky (diagonal not > 0: all weights 0),
// forward substitution, and back substitution with each row's dot product
// summed sequentially.  The reference's linear-algebra crate may accumulate
// those dot products in another order, so the weights agree with the
// reference within FP64 rounding, not bit for bit; a weight within rounding
// of zero can then leave or stay in the active set differently.  The most
// negative component (the first of equals) leaves the set.  The positive
// weights are kept as FP32 in candidate order (:290-294).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k3_nnls(Knn3View v) {
#pragma clang fp contract(off)
    __shared__ double Lp[kMaxK * (kMaxK + 1) / 2];  // the lower triangle, row after row
    auto L = [&](int r, int c) -> double& { return Lp[r * (r + 1) / 2 + c]; };
    __shared__ double prod[kMaxK];
    __shared__ double resw[kMaxK];
    __shared__ int idxs[kMaxK];
    const int m = blockIdx.x, lane = threadIdx.x;
    const int K = v.nb_cnt[m];
    if (K == 0) {
        if (lane == 0) v.w_cnt[m] = 0;
        return;
    }
    const double* sys = v.sys + static_cast<size_t>(m) * kMaxK * kMaxK;
    const double* rhs = v.rhs + static_cast<size_t>(m) * kMaxK;
    bool act = lane < K;
    bool ok = false;
    resw[lane] = 0.0;
    for (int it = 0; it < K; ++it) {
        const uint64_t am = __ballot(act);
        const int k = __builtin_popcountll(am);
        if (k == 0) break;
        if (act) idxs[__builtin_popcountll(am & lanes_below(lane))] = lane;
        __syncthreads();
        double b = 0.0;
        if (lane < k) {
            const int gi = idxs[lane];
            b = rhs[gi];
            for (int rj = 0; rj <= lane; ++rj) L(lane, rj) = sys[static_cast<size_t>(gi) * kMaxK + idxs[rj]];
        }
        __syncthreads();
        bool fail = false;
        for (int j = 0; j < k; ++j) {
            double s = 0.0;
            if (lane >= j && lane < k) {
                s = L(lane, j);
                for (int kk = 0; kk < j; ++kk) s = s + (-L(j, kk)) * L(lane, kk);
            }
            const double diag = __shfl(s, j, 64);
            if (!(diag > 0.0)) {
                fail = true;
                break;
            }
            const double den = sqrt(diag);
            if (lane == j) L(j, j) = den;
            else if (lane > j && lane < k) L(lane, j) = s / den;
            __syncthreads();
        }
        if (fail) break;
        for (int i = 0; i < k; ++i) {  // L y = b
            const double y = __shfl(b, i, 64) / L(i, i);
            if (lane == i) b = y;
            else if (lane > i && lane < k) b = b + (-y) * L(lane, i);
        }
        for (int i = k - 1; i >= 0; --i) {  // L^T x = y
            prod[lane] = (lane > i && lane < k) ? L(lane, i) * b : 0.0;
            __syncthreads();
            if (lane == i) {
                double dot = 0.0;
                for (int r = i + 1; r < k; ++r) dot = dot + prod[r];
                b = (b - dot) / L(i, i);
            }
            __syncthreads();
        }
        const double wmin = wave_min_f64(lane < k && b < 0.0 ? b : 0.0);
        if (!(wmin < 0.0)) {
            if (lane < k) resw[idxs[lane]] = b;
            ok = true;
            break;
        }
        const uint64_t wm = __ballot(lane < k && b == wmin);
        const int out = idxs[__builtin_ctzll(wm)];
        __syncthreads();
        if (lane == out) act = false;
    }
    __syncthreads();
    const double w = ok && lane < K ? resw[lane] : 0.0;
    const bool keep = lane < K && w > 0.0;
    const uint64_t km = __ballot(keep);
    if (keep) {
        const size_t o = static_cast<size_t>(m) * kMaxK + static_cast<size_t>(__builtin_popcountll(km & lanes_below(lane)));
        v.w_item[o] = v.nb_item[static_cast<size_t>(m) * kMaxK + lane];
        v.w_val[o] = static_cast<float>(w);
    }
    if (lane == 0) v.w_cnt[m] = __builtin_popcountll(km);
}

// ---------------------------------------------------------------------------
// Knn3Model::predict (src/knn3.rs:379-389), one wave per query: lane c looks
// up the c-th weighted neighbour in the user's item-sorted list (UserErrors,
// src/knn3.rs:37-77) and forms w * err; the products of the found neighbours
// are summed in neighbour order in FP32 (0 with no neighbours).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k3_predict(Knn3View v, int64_t n, const int32_t* __restrict__ us,
                                                  const int32_t* __restrict__ is, float* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ float vals[4][kMaxK];
    const int lane = threadIdx.x & 63, ws = threadIdx.x >> 6;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + ws; q < n; q += nwaves) {
        const int32_t u = us[q], i = is[q];
        const int nn = v.w_cnt[i];
        bool hit = false;
        float val = 0.0f;
        if (lane < nn) {
            const size_t o = static_cast<size_t>(i) * kMaxK + lane;
            const int32_t want = v.w_item[o];
            uint32_t lo = v.u_off[u], hi = v.u_off[u + 1];
            const uint32_t e = hi;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (v.u_item[mid] < want) lo = mid + 1;
                else hi = mid;
            }
            if (lo < e && v.u_item[lo] == want) {
                hit = true;
                val = v.w_val[o] * v.u_err[lo];
            }
        }
        vals[ws][lane] = val;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t hm = __ballot(hit);
        float pred = 0.0f;
        while (hm) {
            const int c = __builtin_ctzll(hm);
            pred = pred + vals[ws][c];
            hm &= hm - 1;
        }
        if (lane == 0) out[q] = pred;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

hipError_t launch_knn3_build(const Knn3View& v, hipStream_t st) {
    (void)hipGetLastError();
    if (v.n_items < 1 || v.ip > kMapMax || v.k_max > kMaxK) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k3_select, dim3(v.n_items), dim3(256), 0, st, v);
    hipLaunchKernelGGL(k3_cov, dim3(v.n_items), dim3(256), 0, st, v);
    hipLaunchKernelGGL(k3_nnls, dim3(v.n_items), dim3(64), 0, st, v);
    return hipGetLastError();
}

hipError_t launch_knn3_predict(const Knn3View& v, int64_t n, const int32_t* us, const int32_t* is, float* out,
                               hipStream_t st) {
    (void)hipGetLastError();
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k3_predict, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, v, n, us, is, out);
    return hipGetLastError();
}

}  // namespace knn
