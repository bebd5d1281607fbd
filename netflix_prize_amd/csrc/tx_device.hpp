// Device code shared by the translation units of libtsvd_tx.so: small
// helpers and the parallel SGD kernel template k_sgd_par (instantiated per
// schedule variant in tx_sgd_*.hip so the variants compile in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tx_types.hpp"

namespace tx {


// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

template <int TEAM>
__device__ __forceinline__ float team_sum(float v) {
#pragma unroll
    for (int o = TEAM / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, TEAM);
    return v;
}

template <int TEAM>
__device__ __forceinline__ double team_sum_d(double v) {
#pragma unroll
    for (int o = TEAM / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, TEAM);
    return v;
}

template <int VEC>
struct Vec {
    float v[VEC];
};

template <int VEC>
__device__ __forceinline__ void load_vec(const float* __restrict__ p, float (&out)[VEC]) {
    if constexpr (VEC % 4 == 0) {
#pragma unroll
        for (int c = 0; c < VEC; c += 4) {
            const float4 x = *reinterpret_cast<const float4*>(p + c);
            out[c] = x.x; out[c + 1] = x.y; out[c + 2] = x.z; out[c + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < VEC; ++c) out[c] = p[c];
    }
}

template <int VEC>
__device__ __forceinline__ void store_vec(float* __restrict__ p, const float (&in)[VEC]) {
    if constexpr (VEC % 4 == 0) {
#pragma unroll
        for (int c = 0; c < VEC; c += 4)
            *reinterpret_cast<float4*>(p + c) = make_float4(in[c], in[c + 1], in[c + 2], in[c + 3]);
    } else {
#pragma unroll
        for (int c = 0; c < VEC; ++c) p[c] = in[c];
    }
}

// Strided element mapping: lane l of a team holds elements l, l + TEAM, ...,
// so each per-element instruction of a team touches TEAM consecutive floats.
// VEC consecutive BF16 values widened to FP32 (exact).
template <int VEC>
__device__ __forceinline__ void load_bf16(const uint16_t* __restrict__ p, float (&out)[VEC]) {
#pragma unroll
    for (int c = 0; c < VEC; ++c) out[c] = __uint_as_float(static_cast<uint32_t>(p[c]) << 16);
}

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint16_t to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

template <int VEC, int TEAM>
__device__ __forceinline__ void load_str(const float* __restrict__ p, float (&out)[VEC]) {
#pragma unroll
    for (int c = 0; c < VEC; ++c) out[c] = p[c * TEAM];
}

template <int VEC, int TEAM>
__device__ __forceinline__ void store_str(float* __restrict__ p, const float (&in)[VEC]) {
#pragma unroll
    for (int c = 0; c < VEC; ++c) p[c * TEAM] = in[c];
}

__device__ __forceinline__ void atomic_add_dev(float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// time_bin (src/tx.rs:285-289)
__device__ __forceinline__ int time_bin_of(int day, int B, int day_range) {
    const long long b = (static_cast<long long>(day) * B) / day_range;
    return static_cast<int>(b < B - 1 ? b : B - 1);
}

// dev (src/tx.rs:292-300); powf here is the device libm (within 1-2 ulp of the
// host's), used only for arbitrary predict queries.  Training ratings carry a
// host-evaluated dev.
__device__ __forceinline__ float dev_of(float tu_mean, int day, float beta) {
    const float dt = static_cast<float>(day) - tu_mean;
    if (dt == 0.0f) return 0.0f;
    const float s = dt > 0.0f ? 1.0f : -1.0f;
    return s * powf(fabsf(dt), beta);
}

// SparseUD::index (src/tx.rs:76-83): binary search of the user's days.
__device__ __forceinline__ int64_t sud_find(const DevModel& m, int u, int16_t day) {
    uint32_t lo = m.sud_start[u], hi = m.sud_start[u + 1];
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        const int16_t d = m.sud_date[mid];
        if (d < day) lo = mid + 1;
        else if (d > day) hi = mid;
        else return mid;
    }
    return -1;
}

__device__ __forceinline__ float sigmoid_f(float x) { return 1.0f / (1.0f + expf(-x)); }

// OrdinalHead::grad (src/lib.rs:565-599); returns d loss / d score and the
// threshold gradient.
__device__ __forceinline__ float ordinal_grad(const float th[4], float score, int y, float g_t[4]) {
    float F[6], a[6];
    F[0] = 0.0f; F[5] = 1.0f;
#pragma unroll
    for (int k = 1; k < 5; ++k) F[k] = sigmoid_f(th[k - 1] - score);
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = F[k] * (1.0f - F[k]);
    float p_y = F[y] - F[y - 1];
    if (!(p_y >= 1e-6f)) p_y = 1e-6f;
    const float g_ds = (a[y] - a[y - 1]) / p_y;
#pragma unroll
    for (int k = 0; k < 4; ++k) g_t[k] = 0.0f;
    if (y >= 2) g_t[y - 2] += a[y - 1] / p_y;
    if (y <= 4) g_t[y - 1] -= a[y] / p_y;
    return g_ds;
}

// OrdinalHead::predict_probs expectation (src/lib.rs:536-559, src/tx.rs:717-720)
__device__ __forceinline__ float ordinal_expect(const float th[4], float score) {
    float F[6];
    F[0] = 0.0f; F[5] = 1.0f;
#pragma unroll
    for (int k = 1; k < 5; ++k) F[k] = sigmoid_f(th[k - 1] - score);
    float p[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) p[k] = F[k + 1] - F[k];
    return 1.0f * p[0] + 2.0f * p[1] + 3.0f * p[2] + 4.0f * p[3] + 5.0f * p[4];
}

// ---------------------------------------------------------------------------
// Parallel schedule.  One team of TEAM lanes owns one user segment; lane l
// holds factor elements [l*VEC, l*VEC + VEC).  Users of a cell are pulled from
// a per-(stratum, cell) queue in chunks; cell = block index mod n_cells, and
// workgroups are dispatched to the 8 XCDs round-robin, so with n_cells = 8 a
// cell's workgroups all run on one XCD and share its L2.
// ---------------------------------------------------------------------------

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the TEAM lanes of a team; every lane of the team gets the result.
// Within a 16-lane row: quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror (each lane ends with the row sum); across rows
// one bpermute (TEAM = 32) or four readlanes (TEAM = 64).
template <int TEAM>
__device__ __forceinline__ float team_sum_dpp(float v) {
    if constexpr (TEAM >= 2) v += dpp_f<0xB1>(v);
    if constexpr (TEAM >= 4) v += dpp_f<0x4E>(v);
    if constexpr (TEAM >= 8) v += dpp_f<0x141>(v);
    if constexpr (TEAM >= 16) v += dpp_f<0x140>(v);
    if constexpr (TEAM == 32) v += __shfl_xor(v, 16, 32);
    if constexpr (TEAM == 64) {
        const int b = __float_as_int(v);
        v = (__int_as_float(__builtin_amdgcn_readlane(b, 0)) + __int_as_float(__builtin_amdgcn_readlane(b, 16))) +
            (__int_as_float(__builtin_amdgcn_readlane(b, 32)) + __int_as_float(__builtin_amdgcn_readlane(b, 48)));
    }
    return v;
}

// Value of lane j (team-relative, uniform across the team) to the whole team.
template <int TEAM>
__device__ __forceinline__ uint32_t team_bcast(uint32_t v, uint32_t j) {
    if constexpr (TEAM == 64) return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(j)));
    else return static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>(j), TEAM));
}

template <int TEAM>
__device__ __forceinline__ float team_bcast_f(float v, uint32_t j) {
    return __uint_as_float(team_bcast<TEAM>(__float_as_uint(v), j));
}

// Apply an SGD delta to a shared row element (see tx_schedule.update_flags):
// mode 2 writes old + delta, modes 0 and 1 add atomically (gfx950 float
// atomics carry no scope bits, so both compile to the same instruction).
__device__ __forceinline__ void apply_delta(float* p, float old, float delta, int mode) {
    if (mode == 2) *p = old + delta;
    else __hip_atomic_fetch_add(p, delta, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The mode test is hoisted out of the element loop: both sides issue VEC
// stores, so the wait counters stay exact after the branch joins.
template <int VEC, int TEAM>
__device__ __forceinline__ void apply_vec(float* p, const float (&old)[VEC], const float (&delta)[VEC], int mode) {
    if (mode == 2) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) p[c * TEAM] = old[c] + delta[c];
    } else {
#pragma unroll
        for (int c = 0; c < VEC; ++c) __hip_atomic_fetch_add(p + c * TEAM, delta[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// A copy of a wave-uniform value that the compiler must keep in a VGPR.
__device__ __forceinline__ float in_vgpr(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

// A wave-uniform pointer the compiler must keep in VGPRs: loads through it
// land in VGPRs (and free the scalar registers the value would hold).
template <class T>
__device__ __forceinline__ T* in_vgpr_ptr(T* p) {
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    uint32_t lo, hi;
    asm("v_mov_b32 %0, %1" : "=v"(lo) : "s"(static_cast<uint32_t>(a)));
    asm("v_mov_b32 %0, %1" : "=v"(hi) : "s"(static_cast<uint32_t>(a >> 32)));
    return reinterpret_cast<T*>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Walks rows 0..n-1 of a user's item list in order.  Row r's item index is
// *idx_ptr(r): the indices arrive TEAM at a time (lane l loads index b0 + l),
// a batch ahead.  The rows are read contiguously (lane l: elements l*VEC ..),
// GB at a time and double-buffered, so use() of one group overlaps the loads
// of the next and a store issued by use() is never waited for by a later
// load.  Every load is unconditional, indexes past n clamped to row n-1 (read,
// never used).  batch(b0, ib) runs once per index batch: lane l holds the
// index of row b0 + l in ib.
template <int TEAM, int VEC, int GB, class IdxPtr, class Batch, class Use>
__device__ __forceinline__ void for_rows(uint32_t n, int lane, const float* yfeat, int fs, IdxPtr idx_ptr,
                                         Batch batch, Use use) {
    static_assert(TEAM % (2 * GB) == 0, "two row groups must tile an index batch");
    if (n == 0) return;
    auto idx_load = [&](uint32_t r) { return *idx_ptr(r < n ? r : n - 1); };
    const int coff = lane * VEC;
    uint32_t b0 = 0;
    uint32_t ib = idx_load(static_cast<uint32_t>(lane)), ibn = idx_load(TEAM + static_cast<uint32_t>(lane));
    auto jrow = [&](uint32_t r) -> uint32_t {  // item of row r, r - b0 < 2*TEAM
        const uint32_t d = r - b0;
        if (d < TEAM) return team_bcast<TEAM>(ib, d);
        asm volatile("");  // keep the branch: a select would wait for ibn
        return team_bcast<TEAM>(ibn, d - TEAM);
    };
    float ya[GB][VEC], yb[GB][VEC];
    uint32_t ja[GB], jb[GB];
    auto load = [&](uint32_t r0, float (&y)[GB][VEC], uint32_t (&jj)[GB]) {
#pragma unroll
        for (int k = 0; k < GB; ++k) {
            jj[k] = jrow(r0 + k);
            load_vec<VEC>(yfeat + static_cast<size_t>(jj[k]) * fs + coff, y[k]);
        }
    };
    batch(b0, ib);
    load(0, ya, ja);
    for (uint32_t r0 = 0;; r0 += 2 * GB) {
        load(r0 + GB, yb, jb);
#pragma unroll
        for (int k = 0; k < GB; ++k)
            if (r0 + k < n) use(r0 + k, ja[k], ya[k]);
        if (r0 + GB >= n) break;
        load(r0 + 2 * GB, ya, ja);
#pragma unroll
        for (int k = 0; k < GB; ++k)
            if (r0 + GB + k < n) use(r0 + GB + k, jb[k], yb[k]);
        if (r0 + 2 * GB >= n) break;
        if (r0 + 2 * GB - b0 == TEAM) {
            b0 += TEAM;
            ib = ibn;
            ibn = idx_load(b0 + TEAM + static_cast<uint32_t>(lane));
            batch(b0, ib);
        }
    }
}

// Rows per group of for_rows: 32 floats of row data per lane per group.
template <int TEAM, int VEC>
constexpr int row_group() {
    return (TEAM / 2 < 32 / VEC) ? TEAM / 2 : 32 / VEC;
}

// for_rows over a row stream whose entries may name rows of different
// tables: entry j's row starts at row_ptr(j); use(j, row) sees the entry and
// the row in the contiguous layout.  Same pipeline: indices a batch ahead,
// rows GB at a time double-buffered, every load unconditional (indexes past n
// clamped).  (Row memory is the natural element order in every layout, so a
// row this walker stores contiguously reads back strided elsewhere.)
template <int TEAM, int VEC, int GB, class RowPtr, class Use>
__device__ __forceinline__ void for_rows_at(uint32_t n, int lane, const uint32_t* idx, RowPtr row_ptr, Use use) {
    static_assert(TEAM % (2 * GB) == 0, "two row groups must tile an index batch");
    if (n == 0) return;
    auto idx_load = [&](uint32_t r) { return idx[r < n ? r : n - 1]; };
    uint32_t b0 = 0;
    uint32_t ib = idx_load(static_cast<uint32_t>(lane)), ibn = idx_load(TEAM + static_cast<uint32_t>(lane));
    auto jrow = [&](uint32_t r) -> uint32_t {
        const uint32_t d = r - b0;
        if (d < TEAM) return team_bcast<TEAM>(ib, d);
        asm volatile("");
        return team_bcast<TEAM>(ibn, d - TEAM);
    };
    float ya[GB][VEC], yb[GB][VEC];
    uint32_t ja[GB], jb[GB];
    auto load = [&](uint32_t r0, float (&y)[GB][VEC], uint32_t (&jj)[GB]) {
#pragma unroll
        for (int k = 0; k < GB; ++k) {
            jj[k] = jrow(r0 + k);
            load_vec<VEC>(row_ptr(jj[k]) + lane * VEC, y[k]);
        }
    };
    load(0, ya, ja);
    for (uint32_t r0 = 0;; r0 += 2 * GB) {
        load(r0 + GB, yb, jb);
#pragma unroll
        for (int k = 0; k < GB; ++k)
            if (r0 + k < n) use(ja[k], ya[k]);
        if (r0 + GB >= n) break;
        load(r0 + 2 * GB, ya, ja);
#pragma unroll
        for (int k = 0; k < GB; ++k)
            if (r0 + GB + k < n) use(jb[k], yb[k]);
        if (r0 + 2 * GB >= n) break;
        if (r0 + 2 * GB - b0 == TEAM) {
            b0 += TEAM;
            ib = ibn;
            ibn = idx_load(b0 + TEAM + static_cast<uint32_t>(lane));
        }
    }
}

// Workgroup barrier that orders LDS only: the waves of a multi-wave team meet
// here once per rating, and a release fence at workgroup scope would also wait
// for every global load in flight (the whole prefetch ring).
__device__ __forceinline__ void wg_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Day-stream entries with this bit set are day headers (index = SparseUD day).
constexpr uint32_t kDayHead = 0x80000000u;

// The next training rating and everything read for it before its update.
// The rating's scalars arrive in one load spread over team lanes:
//   sc  lane 0 b_i, 1 b_i,bin, 2 b_u,day, 3 b_i,f(0)
//   sc2 lane 0 c_u,day, 1 b_u,bin
// A user-owned bias written by an earlier rating after this one was fetched
// is forwarded in bud/but (has_bud/has_but set) rather than written into
// sc/sc2: those registers may still be waiting for their load, and writing
// them would stall the wave until every younger memory operation returned.
template <int VEC, bool DAY, int NB = 0>
struct Pending {
    uint32_t i, ud, fb;
    float dev, r;
    int b, raw;
    float q[VEC], qf[VEC];
    // neighbourhood terms: c(i, j) and w(i, j) of this lane's two global
    // neighbour slots
    float cv[NB ? 2 : 1], wv[NB ? 2 : 1];
    // day terms: p_u,day, the day's NSVD1 sum over y_day (strided) and the run norm
    float pd[DAY ? VEC : 1], sd[DAY ? VEC : 1];
    float nd;
    float sc, sc2;
    float* sc_addr;  // this lane's slot of sc (read at fetch, updated after the rating)
    float bud, but;
    bool has_bud, has_but;
};

#ifndef TX_DAY_GB
#define TX_DAY_GB 16
#endif

#ifndef TX_PREFETCH
#define TX_PREFETCH 3
#endif

#ifndef TX_SGD_WAVES
#define TX_SGD_WAVES 1
#endif

// PASS 0: one SGD epoch of TxModel::fit_epoch (src/tx.rs:734-1115).
// PASS 1: the user pass of Tsvdx4pModel (src/tsvdx4p.rs:390-639): item tables
//         frozen, user parameters trained, each user's NSVD1 gradient over its
//         norm stored per user (sv.gacc, sv.gacc_b) for k_nsvd_gather.
// PASS 2: the same user pass without the accumulators (fold-in of users
//         against fixed item tables).
// DAY:    the day terms of a model that is not low-memory (src/tx.rs:846-871,
//         1025-1036), one stratum only.  A user's training ratings are sorted
//         by day, so each SparseUD day is one same-day run.  At user start one
//         row stream (m.dg_item) sums every training day's y_day rows (its
//         training items, plus its held-out items with full_su) into
//         sv.day_acc; a run divides that sum by the run norm, carries p_u,day
//         and its gradient sum in registers and leaves them in ufeat_day and
//         sv.day_acc; at user end a second stream (m.dy_item) applies the y_day
//         step of every run.  A user's days have disjoint item sets (each item
//         is rated once), so only other users' concurrent steps separate this
//         order from the reference's gather-at-run-start, step-at-run-end.
// MW:     a team of W = blockDim.x / 64 waves per user, W <= MW (4 or 8; 0 =
//         one wave; f > 1024, one team per block, one stratum, PASS 0).
//         Wave w owns elements [w*64*VEC,
//         (w+1)*64*VEC) of every row and runs the one-wave code on that chunk;
//         the waves meet only in the rating's dot product (per-wave sums
//         through LDS and a barrier, added in wave order) and the queue grab.
//         Every wave computes the same scalars; the item-bias atomics and
//         the shared NSVD1-bias writes are issued by wave 0 alone, user-owned
//         slots by every wave (each wave's own later loads then see them).
// SPLIT:  a team of two waves per user stream (128-thread workgroups, one
//         stratum, plain y rows, low-memory models without ordinal head or
//         neighbourhood terms).  Wave 0 runs the rating loops; wave 1 takes
//         the users from the queue, gathers each user's NSVD1 sum before its
//         rating loop and applies its y step during the next user's loop.
//         The sums and gradients pass through LDS slots (by user parity) and
//         the two waves meet at one workgroup barrier per user, so the rating
//         wave never waits for the row streams.  Against one wave per user,
//         a user's y step lands one user later and its NSVD1 sum is taken one
//         user earlier; the arithmetic is the same.
template <int TEAM, int VEC, bool QF, bool P2, bool ORD, int PASS, bool DAY, int MW, int NB = 0, bool SPLIT = false>
__global__ __launch_bounds__(MW == 8 ? 512 : 256, MW ? 1 : TX_SGD_WAVES) void k_sgd_par(DevModel m, StrataView sv, int stratum) {
    static_assert(NB == 0 || (PASS == 0 && TEAM == 64), "neighbourhood terms: SGD epoch, one-wave lanes");
    static_assert(!SPLIT || (TEAM == 64 && PASS == 0 && !DAY && !MW && NB == 0 && !ORD),
                  "split teams: the one-wave SGD epoch of low-memory models");
    static_assert(PASS == 0 || !ORD, "the user passes have no ordinal head");
    static_assert(PASS == 0 || !DAY, "the user passes train low-memory models");
    static_assert(!MW || TEAM == 64, "multi-wave teams: 64-lane waves");
    constexpr int CHUNK = 4;
    // rows per group of the row streams: day-term models stream one-element
    // rows (VEC 1) in groups of 16, which keeps the kernel within the
    // registers of three waves per SIMD
    constexpr int GBS = (DAY && VEC == 1) ? TX_DAY_GB : row_group<TEAM, VEC>();
    // up to 4 waves share a SIMD's register file with nothing else; with 8,
    // two waves share each SIMD and one rating ahead keeps them closer to it
    constexpr int DEPTH = MW == 8 ? 1 : TX_PREFETCH;
    __shared__ float xbuf[(MW == 8 ? 512 : 256) * VEC];  // per-wave transpose slot, TEAM*VEC floats each
    __shared__ float dred[2][MW ? 9 : 1];            // per-wave dot products + wave 0's b_i(t) (MW), double-buffered
    __shared__ uint32_t qgrab[1];                    // the team's queue grab (MW)
    // SPLIT: per user parity the NSVD1 sum (contiguous row order) and its bias
    // part, the gradient over the norm (row order) and its bias part, the user
    constexpr uint32_t kNoUser = 0xffffffffu;
    __shared__ float sp_s[SPLIT ? 2 : 1][SPLIT ? TEAM * VEC : 1];
    __shared__ float sp_g[SPLIT ? 2 : 1][SPLIT ? TEAM * VEC : 1];
    __shared__ float sp_sb[2], sp_gb[2];
    __shared__ uint32_t sp_user[2];
    const int lane = threadIdx.x % TEAM;
    const int wv = MW ? static_cast<int>(threadIdx.x / 64) : 0;  // wave of the team
    const int nwv = MW ? static_cast<int>(blockDim.x / 64) : 1;
    const int cbase = wv * TEAM * VEC;                            // first element of the wave's chunk
    const bool lead = wv == 0;                                    // issues the shared one-off writes
    uint32_t dpar = 0;
    const int S = sv.n_strata;
    const int cell = sv.n_cells > 1 ? static_cast<int>(blockIdx.x % sv.n_cells) : 0;
    // with length buckets in one launch the workgroup's range names its queue
    // (bq_first nondecreasing: the count of bucket starts at or below it)
    const int32_t bx = static_cast<int32_t>(blockIdx.x);
    const int qid = sv.n_bq > 0 ? (bx >= sv.bq_first[1]) + (bx >= sv.bq_first[2]) + (bx >= sv.bq_first[3])
                                : stratum * sv.n_cells + cell;
    const uint32_t qbeg = sv.q_start[qid];
    const uint32_t qlen = sv.q_start[qid + 1] - qbeg;
    const int fs = m.fs, F = m.F, B = m.B;
    // The rating loop's hyper-parameters live in VGPRs (an opaque copy, so the
    // compiler cannot fold them back): VALU operands either way, and the
    // scalar registers stay free for the loop state instead of spilling.
    Hyper h = m.h;
    h.lr_ub = in_vgpr(h.lr_ub); h.lr_ta = in_vgpr(h.lr_ta); h.lr_tu = in_vgpr(h.lr_tu);
    h.lr_ubd = in_vgpr(h.lr_ubd); h.lr_ib = in_vgpr(h.lr_ib); h.lr_ti = in_vgpr(h.lr_ti);
    h.lr_ibf = in_vgpr(h.lr_ibf); h.lr_cu = in_vgpr(h.lr_cu); h.reg_cu = in_vgpr(h.reg_cu);
    h.lr_u = in_vgpr(h.lr_u); h.reg_u = in_vgpr(h.reg_u); h.lr_i = in_vgpr(h.lr_i); h.reg_i = in_vgpr(h.reg_i);
    h.w_bias = in_vgpr(h.w_bias); h.w_factor = in_vgpr(h.w_factor);
    if constexpr (P2) { h.lr_u2 = in_vgpr(h.lr_u2); h.reg_u2 = in_vgpr(h.reg_u2); }
    if constexpr (QF) { h.lr_iqf = in_vgpr(h.lr_iqf); h.reg_iqf = in_vgpr(h.reg_iqf); }
    const bool has_ud = h.lr_ud != 0.0f;  // p_u,day trained (src/tx.rs:1005-1012)
    if constexpr (DAY) { h.lr_ud = in_vgpr(h.lr_ud); h.reg_ud = in_vgpr(h.reg_ud); }
    const int off = cbase + lane;  // strided: element cbase + c*TEAM + lane
    const bool first = (stratum == 0);
    const bool last = (stratum == S - 1);
    const int qm = sv.q_mode, ym = sv.y_mode, nm = sv.nb_mode;
    const bool bfz = PASS != 0 && m.bf16_frozen != 0;  // q, q_f read from BF16 copies

    // Per-lane slot of the rating's scalars (sc): lane 0 b_i, 1 b_i,bin,
    // 2 b_u,day, 3 b_i,f(0); other lanes, and lane 3 when b_i,f is skipped,
    // re-read b_i.  Address = base + i*imul + (b & bm) + (ud & um); all
    // lane-dependent, so they live in VGPRs and cost no scalar registers.
    const bool l3 = lane == 3 && !m.skip_ibf;
    const bool sc_item = lane == 0 || lane == 1 || l3;
    float* const sc_base = lane == 1 ? m.bit_bin : lane == 2 ? m.ubias_day : l3 ? m.ibias_freq : m.ibias;
    const uint32_t ibs = static_cast<uint32_t>(m.ibs);  // b_i and b_i,bin share one row per item
    const uint32_t sc_imul = lane == 1 ? ibs : lane == 2 ? 0u : l3 ? static_cast<uint32_t>(F) : ibs;
    const uint32_t sc_bm = lane == 1 ? 0xffffffffu : 0u;
    const uint32_t sc_um = lane == 2 ? 0xffffffffu : 0u;
    const uint32_t sc_fm = (DAY && l3) ? 0xffffffffu : 0u;  // b_i,f(t): the day's freq bin
    const float sc_lr = lane == 0 ? h.lr_ib : lane == 1 ? h.lr_ti : h.lr_ibf;
    float* const q_base = m.ifeat + off;
    // the run norms load into VGPRs (day-term kernels)
    const float* const run_norm = DAY ? in_vgpr_ptr(m.t_run_norm) : m.t_run_norm;

    float th[4] = {0.f, 0.f, 0.f, 0.f};
    if constexpr (ORD) {
#pragma unroll
        for (int k = 0; k < 4; ++k) th[k] = m.thresholds[k];
    }

    if constexpr (SPLIT) {
        if (threadIdx.x >= TEAM) {
            // the streaming wave: queue, NSVD1 gathers, y steps
            const int coff = lane * VEC;
            auto grab = [&]() -> uint32_t {
                uint32_t w = 0;
                if (lane == 0) w = __hip_atomic_fetch_add(&sv.q_head[qid], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                w = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(w)));
                return w < qlen ? sv.q_users[qbeg + w] : kNoUser;
            };
            auto gather = [&](uint32_t u, int slot) {  // src/tx.rs:801-822
                float acc[VEC];
                float sb = 0.0f;
#pragma unroll
                for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
                const uint32_t ts = m.t_off[u], te = m.t_off[u + 1];
                const uint32_t nh = m.full_su ? m.h_start[u + 1] - m.h_start[u] : 0;
                const uint32_t ntr = te - ts, ntot = ntr + nh;
                const uint32_t* titems = m.t_item + ts;
                const uint32_t* hitems = m.h_item + m.h_start[u] - ntr;  // row r >= ntr is hitems[r]
                const bool yb_on = !m.skip_yb;
                for_rows<TEAM, VEC, GBS>(
                    ntot, lane, m.yfeat + cbase, fs, [&](uint32_t r) { return r < ntr ? titems + r : hitems + r; },
                    [&](uint32_t b0, uint32_t ib) {
                        if (yb_on && b0 + static_cast<uint32_t>(lane) < ntot) sb += m.ybias[ib];
                    },
                    [&](uint32_t, uint32_t, const float (&y)[VEC]) {
#pragma unroll
                        for (int c = 0; c < VEC; ++c) acc[c] += y[c];
                    });
                const float inv = 1.0f / m.norm_train[u];
#pragma unroll
                for (int c = 0; c < VEC; ++c) sp_s[slot][coff + c] = acc[c] * inv;
                const float sbv = m.skip_yb ? 0.0f : team_sum_dpp<TEAM>(sb) * inv;
                if (lane == 0) sp_sb[slot] = sbv;
            };
            auto ystep = [&](uint32_t u, int slot) {  // src/tx.rs:1099-1106, plain rows
                const uint32_t ts = m.t_off[u], te = m.t_off[u + 1];
                float gc[VEC];
#pragma unroll
                for (int c = 0; c < VEC; ++c) gc[c] = sp_g[slot][coff + c];
                const float gb = h.lr_yb * sp_gb[slot];
                const uint32_t* titems = m.t_item + ts;
                for_rows<TEAM, VEC, GBS>(
                    te - ts, lane, m.yfeat + cbase, fs, [&](uint32_t r) { return titems + r; },
                    [&](uint32_t b0, uint32_t ib) {
                        if (gb != 0.0f && b0 + static_cast<uint32_t>(lane) < te - ts) m.ybias[ib] = m.ybias[ib] - gb;
                    },
                    [&](uint32_t, uint32_t j, const float (&y)[VEC]) {
                        float ny[VEC];
#pragma unroll
                        for (int c = 0; c < VEC; ++c) ny[c] = y[c] - h.lr_y * (gc[c] + h.reg_y * y[c]);
                        store_vec<VEC>(m.yfeat + static_cast<size_t>(j) * fs + cbase + coff, ny);
                    });
            };
            // user k's sum is in slot k & 1 before barrier k - 1; its gradient
            // is in slot k & 1 after barrier k; one barrier per user, and one
            // more after the last (the runner's count is the same)
            const uint32_t u0 = grab();
            if (lane == 0) sp_user[0] = u0;
            if (u0 != kNoUser) gather(u0, 0);
            wg_barrier();
            uint32_t uprev = kNoUser;
            for (uint32_t k = 0;; ++k) {
                const uint32_t uk =
                    static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(sp_user[k & 1])));
                if (uprev != kNoUser) ystep(uprev, static_cast<int>((k - 1) & 1));
                const uint32_t nx = uk != kNoUser ? grab() : kNoUser;
                if (lane == 0) sp_user[(k + 1) & 1] = nx;
                if (nx != kNoUser) gather(nx, static_cast<int>((k + 1) & 1));
                uprev = uk;
                wg_barrier();
                if (uk == kNoUser) break;
            }
            return;
        }
        wg_barrier();  // the first user's sum is ready
    }
    uint32_t split_k = 0;  // SPLIT: index of the runner's current user
    for (;;) {
        uint32_t w0 = 0;
        uint32_t split_u = 0;
        if constexpr (SPLIT) {
            split_u = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(sp_user[split_k & 1])));
        } else if constexpr (MW) {
            // one grab for the team, passed to its waves through LDS
            wg_barrier();
            if (lead && lane == 0)
                qgrab[0] = __hip_atomic_fetch_add(&sv.q_head[qid], static_cast<uint32_t>(CHUNK), __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
            wg_barrier();
            w0 = qgrab[0];
        } else if (lane == 0) {
            w0 = __hip_atomic_fetch_add(&sv.q_head[qid], static_cast<uint32_t>(CHUNK), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        }
        // a full-wave team takes the value with readfirstlane so the compiler
        // sees it (and the user, segment bounds and loop trip counts derived
        // from it) as wave-uniform: scalar loop control, no exec-mask loops
        if constexpr (TEAM == 64) w0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(w0)));
        else w0 = __shfl(w0, 0, TEAM);
        uint32_t w1;
        if constexpr (SPLIT) {
            w0 = 0;
            w1 = split_u != kNoUser ? 1u : 0u;
        } else {
            if (w0 >= qlen) break;
            w1 = w0 + CHUNK < qlen ? w0 + CHUNK : qlen;
        }
        for (uint32_t w = w0; w < w1; ++w) {
            const uint32_t u = SPLIT ? split_u : sv.q_users[qbeg + w];
            const uint32_t s0 = sv.seg_off[static_cast<size_t>(u) * (S + 1) + stratum];
            const uint32_t s1 = sv.seg_off[static_cast<size_t>(u) * (S + 1) + stratum + 1];
            const size_t prow = static_cast<size_t>(u) * fs + off;
            const uint32_t ts = m.t_off[u], te = m.t_off[u + 1];
            const float norm = m.norm_train[u];
            const size_t ubrow = static_cast<size_t>(u) * B;

            // the user's global neighbours (src/tx.rs:762-796), this epoch's
            // sample or every one: lane l holds neighbours l and l + 64 (at
            // most 128): item, c scale (rated_scale or impl_scale), and for a
            // rated one (r - b~) * rated_scale.  Multi-wave teams: wave 0 alone.
            uint32_t nj[2] = {0u, 0u};
            float ncs[2] = {0.0f, 0.0f}, nwc[2] = {0.0f, 0.0f};
            bool nok[2] = {false, false}, nrt[2] = {false, false};
            float nb_nn = 0.0f, nb_nr = 0.0f;
            if constexpr (NB > 0) {
                if (lead) {
                    const uint32_t i0 = m.imp_start[u];
                    const uint32_t nr_full = te - ts, ni_full = m.imp_start[u + 1] - i0;
                    const uint32_t cw = m.smp_cnt[u];
                    const bool smp = cw != 0xffffffffu;
                    const uint32_t rb = smp ? (cw & 0xffffu) : nr_full, ib = smp ? (cw >> 16) : ni_full;
                    const float rs = smp ? static_cast<float>(nr_full) / static_cast<float>(rb) : 1.0f;
                    const float is = (smp && ib > 0) ? static_cast<float>(ni_full) / static_cast<float>(ib) : 1.0f;
                    const uint16_t* pos = m.smp + static_cast<size_t>(u) * m.nb_max;
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const uint32_t k = static_cast<uint32_t>(lane) + 64u * q;
                        if (k < rb) {
                            const uint32_t t2 = ts + (smp ? pos[k] : k);
                            nj[q] = m.t_item[t2];
                            nwc[q] = m.t_rmb[t2] * rs;
                            ncs[q] = rs;
                            nok[q] = nrt[q] = true;
                        } else if (k < rb + ib) {
                            nj[q] = m.imp_item[i0 + (smp ? pos[k] : k - rb)];
                            ncs[q] = is;
                            nok[q] = true;
                        }
                    }
                    nb_nn = m.norm_nu[u];
                    nb_nr = m.norm_ru[u];
                }
            }

            // issue the loads one rating needs, from its packed record
            auto fetch = [&](uint4 rc, Pending<VEC, DAY, NB>& n) {
                n.i = rc.x & 0xFFFFFu;
                n.raw = static_cast<int>((rc.x >> 20) & 15u);
                n.b = static_cast<int>(rc.x >> 24);
                // the freq bin is packed above the day index for day-term models
                // only (low-memory records carry bin 0, so y is the index)
                n.ud = DAY ? (rc.y & 0x0FFFFFFFu) : rc.y;
                n.fb = DAY ? (rc.y >> 28) : 0u;
                // day-term kernels hold a slot's dev and target in VGPRs: their
                // rating loop has more live scalars than the scalar file holds
                n.dev = DAY ? in_vgpr(__uint_as_float(rc.z)) : __uint_as_float(rc.z);
                n.r = DAY ? in_vgpr(__uint_as_float(rc.w)) : __uint_as_float(rc.w);
                n.has_bud = n.has_but = false;
                n.bud = n.but = 0.0f;
                if (bfz) {
                    // the user passes hold q and q_f frozen: their BF16 copies, the
                    // raw 16 bits kept in the slot and widened where they are used
                    // (a conversion here would wait for the load)
                    const uint16_t* bq = m.b_ifeat + static_cast<size_t>(n.i) * fs + off;
#pragma unroll
                    for (int c = 0; c < VEC; ++c) n.q[c] = __uint_as_float(static_cast<uint32_t>(bq[c * TEAM]));
                } else {
                    load_str<VEC, TEAM>(q_base + static_cast<size_t>(n.i) * fs, n.q);
                }
                if constexpr (QF) {
                    const size_t qfr = static_cast<size_t>(n.i) * F + (DAY ? n.fb : 0u);
                    if (bfz) {
                        const uint16_t* bq = m.b_ifeat_freq + qfr * fs + off;
#pragma unroll
                        for (int c = 0; c < VEC; ++c) n.qf[c] = __uint_as_float(static_cast<uint32_t>(bq[c * TEAM]));
                    } else {
                        load_str<VEC, TEAM>(m.ifeat_freq + qfr * fs + off, n.qf);
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < VEC; ++c) n.qf[c] = 0.0f;
                }
                if constexpr (DAY) {
                    load_str<VEC, TEAM>(m.ufeat_day + static_cast<size_t>(n.ud) * fs + off, n.pd);
                    load_str<VEC, TEAM>(sv.day_acc + static_cast<size_t>(n.ud) * fs + off, n.sd);
                }
                if constexpr (NB > 0) {
                    // c(i, j) and w(i, j) of the lane's neighbour slots (empty
                    // slots read column 0, never used)
                    const size_t nrow = static_cast<size_t>(n.i) * static_cast<uint32_t>(m.I);
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        n.cv[q] = m.nc[nrow + nj[q]];
                        n.wv[q] = m.nw[nrow + nj[q]];
                    }
                }
                n.sc_addr = sc_base + static_cast<size_t>(n.i) * sc_imul + (static_cast<uint32_t>(n.b) & sc_bm) +
                            (n.ud & sc_um) + (n.fb & sc_fm);
                n.sc = *n.sc_addr;
                {
                    // always one load: with both terms skipped it re-reads the sc
                    // slot just fetched (cache-hot, ignored).  A branch that set
                    // sc2 = 0 instead would write a register whose earlier load
                    // may be in flight, and the wait for it drains the pipeline.
                    const float* a = m.skip_cut ? &m.but_bin[ubrow + n.b] : &m.cut[n.ud];
                    a = (lane == 1 && !m.skip_but) ? &m.but_bin[ubrow + n.b] : a;
                    a = (m.skip_cut && m.skip_but) ? n.sc_addr : a;
                    n.sc2 = *a;
                }
            };
            float p[VEC], p2[VEC], s[VEC], G[VEC];
            load_str<VEC, TEAM>(m.ufeat + prow, p);
            if constexpr (P2) {
                load_str<VEC, TEAM>(m.ufeat2 + prow, p2);
            } else {
#pragma unroll
                for (int c = 0; c < VEC; ++c) p2[c] = 0.0f;
            }
            float su_b;
            if constexpr (SPLIT) {
                // the NSVD1 sum the streaming wave gathered (row order -> strided)
#pragma unroll
                for (int c = 0; c < VEC; ++c) {
                    s[c] = sp_s[split_k & 1][c * TEAM + lane];
                    G[c] = 0.0f;
                }
                su_b = sp_sb[split_k & 1];
            } else if (PASS != 0 && first && sv.su_cached) {
                // y is frozen for the whole user pass and ycache was rebuilt from
                // it over the same rows: the NSVD1 sum without a second gather
                load_str<VEC, TEAM>(m.ycache + prow, s);
                su_b = m.ycache_bias[u];
#pragma unroll
                for (int c = 0; c < VEC; ++c) G[c] = 0.0f;
            } else if (first) {
                // NSVD1 user aggregation from the current y (src/tx.rs:801-822):
                // rows summed in the contiguous layout, then transposed to the
                // strided layout through the team's LDS slot
                float acc[VEC];
                float sb = 0.0f;
#pragma unroll
                for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
                const uint32_t nh = m.full_su ? m.h_start[u + 1] - m.h_start[u] : 0;
                const uint32_t ntr = te - ts;
                const uint32_t ntot = ntr + nh;
                const uint32_t* titems = m.t_item + ts;
                const uint32_t* hitems = m.h_item + m.h_start[u] - ntr;  // row r >= ntr is hitems[r]
                const int coff = lane * VEC;
                const bool yb_on = !m.skip_yb;
                for_rows<TEAM, VEC, GBS>(
                    ntot, lane, m.yfeat + cbase, fs,
                    [&](uint32_t r) { return r < ntr ? titems + r : hitems + r; },
                    [&](uint32_t b0, uint32_t ib) {
                        if (yb_on && b0 + static_cast<uint32_t>(lane) < ntot) sb += m.ybias[ib];
                    },
                    [&](uint32_t, uint32_t, const float (&y)[VEC]) {
#pragma unroll
                        for (int c = 0; c < VEC; ++c) acc[c] += y[c];
                    });
                const float inv = 1.0f / norm;
                float* x = xbuf + (threadIdx.x / TEAM) * (TEAM * VEC);
#pragma unroll
                for (int c = 0; c < VEC; ++c) x[coff + c] = acc[c] * inv;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int c = 0; c < VEC; ++c) s[c] = x[c * TEAM + lane];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                su_b = m.skip_yb ? 0.0f : team_sum_dpp<TEAM>(sb) * inv;
                if constexpr (DAY) {
                    // every training day's sum over y_day of its training items and
                    // (full_su) held-out items, in one stream of the user's days:
                    // [rows of day d][header d] ...  (src/tx.rs:846-871)
                    const uint32_t g0 = m.dg_off[u];
                    float dacc[VEC];
#pragma unroll
                    for (int c = 0; c < VEC; ++c) dacc[c] = 0.0f;
                    for_rows_at<TEAM, VEC, GBS>(
                        m.dg_off[u + 1] - g0, lane, m.dg_item + g0,
                        [&](uint32_t j) {
                            return (j & kDayHead) ? sv.day_acc + static_cast<size_t>(j & ~kDayHead) * fs + cbase
                                                  : m.yfeat_day + static_cast<size_t>(j) * fs + cbase;
                        },
                        [&](uint32_t j, const float (&y)[VEC]) {
                            if (j & kDayHead) {
                                store_vec<VEC>(sv.day_acc + static_cast<size_t>(j & ~kDayHead) * fs + cbase + lane * VEC, dacc);
#pragma unroll
                                for (int c = 0; c < VEC; ++c) dacc[c] = 0.0f;
                            } else {
#pragma unroll
                                for (int c = 0; c < VEC; ++c) dacc[c] += y[c];
                            }
                        });
                }
                if (S > 1) {
                    store_str<VEC, TEAM>(m.ycache + prow, s);
                    if (lane == 0) m.ycache_bias[u] = su_b;
                }
#pragma unroll
                for (int c = 0; c < VEC; ++c) G[c] = 0.0f;
            } else {
                load_str<VEC, TEAM>(m.ycache + prow, s);
                su_b = m.ycache_bias[u];
                load_str<VEC, TEAM>(sv.gacc + prow, G);
            }
            float Gb = first ? 0.0f : sv.gacc_b[u];
            float ubias = m.ubias[u], alpha = m.alpha_u[u], cu = m.cu[u];


            // records arrive TEAM at a time (lane l holds record tb + l); the two
            // batches after the current one are in flight, so the copies at a
            // batch boundary only touch registers loaded a batch or more ago
            // (copying the batch just requested would wait for it).  Every load in
            // the rating loop is unconditional, indexes past the segment end
            // being clamped to its last record (re-read, never consumed): a
            // path that skipped them would leave the hardware's in-order wait
            // counter nothing younger than the prefetched rows, and each use
            // of a prefetched row would then wait for all memory traffic.
            auto load_recs = [&](uint32_t tb) {
                const uint32_t t = tb + static_cast<uint32_t>(lane);
                return sv.p_rec[t < s1 ? t : s1 - 1];
            };
            if (s0 < s1) {
            uint4 cb = load_recs(s0), nb = load_recs(s0 + TEAM), fb = load_recs(s0 + 2 * TEAM);
            uint32_t tb = s0;
            // record t of the current batch or the next.  A branch, not a select:
            // a select would read nb at every rating and wait for it; the empty
            // asm keeps the compiler from turning the branch back into a select.
            auto rec_at = [&](uint32_t t) {
                const uint32_t j = t - tb;
                uint4 r;
                if (j < TEAM) {
                    r = make_uint4(team_bcast<TEAM>(cb.x, j), team_bcast<TEAM>(cb.y, j), team_bcast<TEAM>(cb.z, j),
                                   team_bcast<TEAM>(cb.w, j));
                } else {
                    asm volatile("");
                    const uint32_t l = j - TEAM;
                    r = make_uint4(team_bcast<TEAM>(nb.x, l), team_bcast<TEAM>(nb.y, l), team_bcast<TEAM>(nb.z, l),
                                   team_bcast<TEAM>(nb.w, l));
                }
                return r;
            };
            // ring of the next DEPTH ratings, loaded before the current rating's
            // updates so their latency overlaps DEPTH ratings of compute.  The
            // ring has DEPTH+1 slots and the loop body covers all of them, so a
            // load lands in the slot consumed one step earlier, whose value is
            // dead: no register copy (and no wait for the load) at the back-edge.
            constexpr int RING = DEPTH + 1;
            Pending<VEC, DAY, NB> ring[RING];
            // day terms of the current same-day run (DAY): p_u,day, the day's
            // NSVD1 sum and the run's gradient sum, and the run's day
            float pud[VEC], sdr[VEC], Gd[VEC];
            uint32_t run_ud = 0xffffffffu;
#pragma unroll
            for (int c = 0; c < VEC; ++c) pud[c] = sdr[c] = Gd[c] = 0.0f;
#pragma unroll
            for (int k = 0; k < RING; ++k) {  // slots not yet fetched match no rating
                ring[k].b = -1;
                ring[k].ud = 0xffffffffu;
            }
#pragma unroll
            for (int k = 0; k < DEPTH; ++k) {
                fetch(rec_at(s0 + k < s1 ? s0 + k : s1 - 1), ring[k]);
                if constexpr (DAY) ring[k].nd = run_norm[s0 + k < s1 ? s0 + k : s1 - 1];
            }
#ifdef TX_PROLOGUE_PAD
            {
                // re-read the q row of the first rating TX_PROLOGUE_PAD times (L2-hot):
                // the first steps then see as many younger memory operations as
                // the steady state, so the compiler's merged wait counts rise
                const volatile float* pq = q_base + static_cast<size_t>(ring[0].i) * fs;
#pragma unroll
                for (int k = 0; k < TX_PROLOGUE_PAD; ++k) (void)pq[(k & (VEC - 1)) * TEAM];
            }
#endif

            static_assert(TEAM % RING == 0, "the prefetch depth plus one must divide the team size");
            for (; tb < s1; tb += TEAM) {
            for (uint32_t t0 = tb; t0 < tb + TEAM && t0 < s1; t0 += RING) {
#pragma unroll
                for (int k = 0; k < RING; ++k) {
                const uint32_t t = t0 + k;
                if (t >= s1) break;
                const Pending<VEC, DAY, NB>& cur = ring[k];
                fetch(rec_at(t + DEPTH < s1 ? t + DEPTH : s1 - 1), ring[(k + DEPTH) % RING]);
                if constexpr (DAY) ring[(k + DEPTH) % RING].nd = run_norm[t + DEPTH < s1 ? t + DEPTH : s1 - 1];
                const uint32_t i = cur.i;
                const float dev = cur.dev;
                const float c_ib = team_bcast_f<TEAM>(cur.sc, 0), c_bt = team_bcast_f<TEAM>(cur.sc, 1);
                const float c_ibf = m.skip_ibf ? 0.0f : team_bcast_f<TEAM>(cur.sc, 3);
                const float c_bud = cur.has_bud ? cur.bud : team_bcast_f<TEAM>(cur.sc, 2);
                const float c_cut = m.skip_cut ? 0.0f : team_bcast_f<TEAM>(cur.sc2, 0);
                const float c_but = cur.has_but ? cur.but : (m.skip_but ? 0.0f : team_bcast_f<TEAM>(cur.sc2, 1));
                float bi_t = c_ib + c_bt + c_ibf;
                // the rating's q and q_f rows (widened from BF16 in the user passes)
                float qv[VEC], qfv[VEC];
#pragma unroll
                for (int c = 0; c < VEC; ++c) {
                    qv[c] = bfz ? __uint_as_float(__float_as_uint(cur.q[c]) << 16) : cur.q[c];
                    qfv[c] = (QF && bfz) ? __uint_as_float(__float_as_uint(cur.qf[c]) << 16) : cur.qf[c];
                }
                const float cu_t = cu + c_cut;
                const float bu_t = ubias + c_but + alpha * dev + c_bud + su_b;
                if constexpr (DAY) {
                    // a new run takes the prefetched day rows (they arrived with
                    // this rating's q row): p_u,day, and the day's y_day sum divided
                    // by the run norm (src/tx.rs:859-866)
                    if (cur.ud != run_ud) {
#pragma unroll
                        for (int c = 0; c < VEC; ++c) {
                            pud[c] = cur.pd[c];
                            sdr[c] = cur.sd[c] / cur.nd;
                            Gd[c] = 0.0f;
                        }
                    }
                    run_ud = cur.ud;
                }

                float part = 0.0f;
                if constexpr (DAY) {
#pragma unroll
                    for (int c = 0; c < VEC; ++c) {
                        float pu = p[c] + s[c] + dev * p2[c];
                        pu += pud[c];
                        pu += sdr[c];
                        part += (qv[c] + qfv[c]) * pu;
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < VEC; ++c) part += (qv[c] + qfv[c]) * (p[c] + s[c] + dev * p2[c]);
                }
                // neighbourhood terms (src/tx.rs:903-956) join the dot product's
                // lane partials: nn c(i,j) scale + nr (r - b~) w(i,j) scale per
                // global neighbour j != i, then the same-day neighbours of the
                // rating's day (NB 2: read now, counted for their norms)
                float sd_nn = 0.0f, sd_nr = 0.0f;
                uint32_t sd_a = 0, sd_b = 0;
                if constexpr (NB > 0) {
                    if (lead) {
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (nok[q] && nj[q] != i) part += nb_nn * (ncs[q] * cur.cv[q]) + nb_nr * (nwc[q] * cur.wv[q]);
                    }
                }
                if constexpr (NB == 2) {
                    if (lead) {
                        sd_a = m.sd_start[cur.ud];
                        sd_b = m.sd_start[cur.ud + 1];
                        const size_t irow = static_cast<size_t>(i) * static_cast<uint32_t>(m.I);
                        float cds = 0.0f, wds = 0.0f, nt = 0.0f, nr = 0.0f;
                        for (uint32_t e = sd_a + static_cast<uint32_t>(lane); e < sd_b; e += 64u) {
                            const uint32_t it = m.sd_item[e], j = it & 0x7fffffffu;
                            if (j == i) continue;
                            nt += 1.0f;
                            cds += m.nc_day[irow + j];
                            if (it >> 31) {
                                nr += 1.0f;
                                wds += m.sd_rmb[e] * m.nw_day[irow + j];
                            }
                        }
                        const int n_tot = static_cast<int>(team_sum_dpp<TEAM>(nt));
                        const int n_rat = static_cast<int>(team_sum_dpp<TEAM>(nr));
                        sd_nn = m.inv_sqrt_n[n_tot < m.n_isq ? n_tot : m.n_isq - 1];
                        sd_nr = m.inv_sqrt_n[n_rat < m.n_isq ? n_rat : m.n_isq - 1];
                        part += sd_nn * cds + sd_nr * wds;
                    }
                }
                float dot = team_sum_dpp<TEAM>(part);
                if constexpr (MW) {
                    // the waves' chunk sums, added in wave order by every wave, and
                    // wave 0's item biases (other teams' atomics may land between
                    // the waves' loads; one value keeps the waves' scalars equal)
                    if (lane == 0) dred[dpar][wv] = dot;
                    if (lane == 0 && lead) dred[dpar][8] = bi_t;
                    wg_barrier();
                    float d = 0.0f;
                    for (int w = 0; w < nwv; ++w) d += dred[dpar][w];
                    dot = d;
                    bi_t = dred[dpar][8];
                    dpar ^= 1u;
                }
                const float score = m.gbias + bu_t + bi_t * cu_t + dot;
                float err;
                if constexpr (ORD) {
                    float g_t[4];
                    err = ordinal_grad(th, score, cur.raw, g_t);
                } else {
                    err = score - cur.r;
                }
                // the user passes weight bias and factor errors alike (src/tsvdx4p.rs:560-600)
                const float err_b = PASS == 0 ? err * h.w_bias : err;
                const float err_d = PASS == 0 ? err * h.w_factor : err;
                if constexpr (NB > 0) {
                    // neighbourhood steps (src/tx.rs:1038-1095) as deltas of the
                    // values read for the forward pass
                    if (lead) {
                        const float err_n = err * h.w_nbr;
                        const size_t irow = static_cast<size_t>(i) * static_cast<uint32_t>(m.I);
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            if (nok[q] && nj[q] != i) {
                                const size_t idx = irow + nj[q];
                                if (nrt[q])
                                    apply_delta(m.nw + idx, cur.wv[q],
                                                -h.lr_w * (err_n * nb_nr * nwc[q] + h.reg_w * cur.wv[q]), nm);
                                apply_delta(m.nc + idx, cur.cv[q],
                                            -h.lr_c * (err_n * nb_nn * ncs[q] + h.reg_c * cur.cv[q]), nm);
                            }
                        }
                        if constexpr (NB == 2) {
                            for (uint32_t e = sd_a + static_cast<uint32_t>(lane); e < sd_b; e += 64u) {
                                const uint32_t it = m.sd_item[e], j = it & 0x7fffffffu;
                                if (j == i) continue;
                                const size_t idx = irow + j;
                                if (it >> 31) {
                                    const float wd = m.nw_day[idx];
                                    apply_delta(m.nw_day + idx, wd,
                                                -h.lr_w_day * (err_n * sd_nr * m.sd_rmb[e] + h.reg_w_day * wd), nm);
                                }
                                const float cd = m.nc_day[idx];
                                apply_delta(m.nc_day + idx, cd, -h.lr_c_day * (err_n * sd_nn + h.reg_c_day * cd), nm);
                            }
                        }
                    }
                }

                // user-owned biases: registers, or written back now and forwarded
                // to every prefetched rating that reads the same slot
                ubias -= h.lr_ub * err_b;
                alpha -= h.lr_ta * err_b * dev;
                const float but_new = c_but - h.lr_tu * err_b;
                const float bud_new = c_bud - h.lr_ubd * err_b;
#pragma unroll
                for (int jj = 1; jj <= DEPTH; ++jj) {
                    Pending<VEC, DAY, NB>& nx = ring[(k + jj) % RING];  // ratings t+1 .. t+DEPTH
                    if (nx.b == cur.b) { nx.has_but = true; nx.but = but_new; }
                    if (nx.ud == cur.ud) { nx.has_bud = true; nx.bud = bud_new; }
                }
                // item biases in one instruction (lanes 0, 1, 3), b_u,day in one (lane 2)
                if (PASS == 0 && sc_item) {
                    if (lead) apply_delta(cur.sc_addr, cur.sc, -sc_lr * (err_b * cu_t), qm);
                } else if (lane == 2) {
                    *cur.sc_addr = bud_new;
                }
                if (lane == 1 && !m.skip_but) m.but_bin[ubrow + cur.b] = but_new;
                cu -= h.lr_cu * (err_b * bi_t + h.reg_cu * (cu - 1.0f));
                Gb += (PASS == 0 && m.sum_err_bug) ? err_b * static_cast<float>(m.f) : err_b;
                float dq[VEC], dqf[VEC];
#pragma unroll
                for (int c = 0; c < VEC; ++c) {
                    const float qe = qv[c] + qfv[c];
                    G[c] += err_d * qe;
                    const float pe = DAY ? p[c] + dev * p2[c] + (has_ud ? pud[c] : 0.0f) + s[c] + sdr[c]
                                         : p[c] + dev * p2[c] + s[c];
                    if constexpr (DAY) {
                        Gd[c] += err_d * qe;
                        const float pd_new = pud[c] - h.lr_ud * (err_d * qe + h.reg_ud * pud[c]);
                        pud[c] = has_ud ? pd_new : pud[c];
                    }
                    p[c] -= h.lr_u * (err_d * qe + h.reg_u * p[c]);
                    if constexpr (P2) p2[c] -= h.lr_u2 * (err_d * qe * dev + h.reg_u2 * p2[c]);
                    dq[c] = -h.lr_i * (err_d * pe + h.reg_i * cur.q[c]);
                    if constexpr (QF) dqf[c] = -h.lr_iqf * (err_d * pe + h.reg_iqf * cur.qf[c]);
                }
                if constexpr (PASS == 0) {
                    apply_vec<VEC, TEAM>(q_base + static_cast<size_t>(i) * fs, cur.q, dq, qm);
                    if constexpr (QF) {
                        const size_t qfr = static_cast<size_t>(i) * F + (DAY ? cur.fb : 0u);
                        apply_vec<VEC, TEAM>(m.ifeat_freq + qfr * fs + off, cur.qf, dqf, qm);
                    }
                }
                if constexpr (DAY) {
                    // end of the run: p_u,day and the run's gradient sum to memory
                    // (the next rating's record is already decoded in the ring)
                    if (t + 1 >= s1 || ring[(k + 1) % RING].ud != cur.ud) {
                        float gd[VEC];
#pragma unroll
                        for (int c = 0; c < VEC; ++c) gd[c] = Gd[c] / cur.nd;
                        store_str<VEC, TEAM>(m.ufeat_day + static_cast<size_t>(cur.ud) * fs + off, pud);
                        store_str<VEC, TEAM>(sv.day_acc + static_cast<size_t>(cur.ud) * fs + off, gd);
                    }
                }
                }
            }
            cb = nb;
            nb = fb;
            fb = load_recs(tb + 3 * TEAM);
            }
            }  // s0 < s1

            store_str<VEC, TEAM>(m.ufeat + prow, p);
            if constexpr (P2) store_str<VEC, TEAM>(m.ufeat2 + prow, p2);
            if (lane == 0 && lead) {
                m.ubias[u] = ubias;
                m.alpha_u[u] = alpha;
                m.cu[u] = cu;
            }
            if constexpr (PASS == 1) {
                // the user's NSVD1 gradient over its norm, once per user: the
                // item-side gather (k_nsvd_gather) adds it into the step of each
                // of the user's training items (NsvdDelta, src/tsvdx4p.rs:622-630)
                // instead of a row of atomics per rating
                float gq[VEC];
#pragma unroll
                for (int c = 0; c < VEC; ++c) gq[c] = G[c] / norm;
                store_str<VEC, TEAM>(sv.gacc + prow, gq);
                if (lane == 0) sv.gacc_b[u] = Gb / norm;
            } else if (PASS == 2) {
                // fold-in: nothing beyond the user's own parameters
            } else if (SPLIT) {
                // the gradient over the norm, for the streaming wave's y step
                const float inv = 1.0f / norm;
#pragma unroll
                for (int c = 0; c < VEC; ++c) sp_g[split_k & 1][c * TEAM + lane] = G[c] * inv;
                if (lane == 0) sp_gb[split_k & 1] = Gb * inv;
            } else if (last && S > 1) {
                // stratified epoch: the user's whole-epoch NSVD1 gradient over its
                // norm, for the item-side gather after the last stratum (the
                // y rows stay fixed through the epoch's strata; k_nsvd_gather
                // applies every user's step at once, divided by n_contrib^exp as
                // in apply_nsvd_deltas, src/tsvdx4p.rs:341-385)
                const float inv = 1.0f / norm;
                float gq[VEC];
#pragma unroll
                for (int c = 0; c < VEC; ++c) gq[c] = G[c] * inv;
                store_str<VEC, TEAM>(sv.gacc + prow, gq);
                if (lane == 0) sv.gacc_b[u] = Gb * inv;
            } else if (last) {
                // end-of-user NSVD1 step over R(u) (src/tx.rs:1099-1106)
                const float inv = 1.0f / norm;
                float gq[VEC];
#pragma unroll
                for (int c = 0; c < VEC; ++c) gq[c] = G[c] * inv;
                const float gb = h.lr_yb * (Gb * inv);
                if (ym == 2) {
                    // lock-free rows: contiguous float4 read-modify-write, with the
                    // user's gradient moved to the contiguous layout through LDS
                    const int coff = lane * VEC;
                    float* x = xbuf + (threadIdx.x / TEAM) * (TEAM * VEC);
#pragma unroll
                    for (int c = 0; c < VEC; ++c) x[c * TEAM + lane] = gq[c];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    float gc[VEC];
#pragma unroll
                    for (int c = 0; c < VEC; ++c) gc[c] = x[coff + c];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t* titems = m.t_item + ts;
                    for_rows<TEAM, VEC, GBS>(
                        te - ts, lane, m.yfeat + cbase, fs, [&](uint32_t r) { return titems + r; },
                        [&](uint32_t b0, uint32_t ib) {
                            if (gb != 0.0f && lead && b0 + static_cast<uint32_t>(lane) < te - ts)
                                m.ybias[ib] = m.ybias[ib] - gb;
                        },
                        [&](uint32_t, uint32_t j, const float (&y)[VEC]) {
                            float ny[VEC];
#pragma unroll
                            for (int c = 0; c < VEC; ++c) ny[c] = y[c] - h.lr_y * (gc[c] + h.reg_y * y[c]);
                            store_vec<VEC>(m.yfeat + static_cast<size_t>(j) * fs + cbase + coff, ny);
                        });
                } else {
                // rows of batch b+1 are loaded before batch b's updates are issued
                    float ya[4][VEC], yb[4][VEC];
                    uint32_t ja[4], jb[4];
                    auto load_batch = [&](uint32_t t0, float (&y)[4][VEC], uint32_t (&jj)[4]) {
    #pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            jj[k] = t0 + k < te ? m.t_item[t0 + k] : 0xffffffffu;
                            if (jj[k] != 0xffffffffu) load_str<VEC, TEAM>(m.yfeat + static_cast<size_t>(jj[k]) * fs + off, y[k]);
                        }
                    };
                    auto update_batch = [&](uint32_t t0, const float (&y)[4][VEC], const uint32_t (&jj)[4]) {
    #pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            if (jj[k] == 0xffffffffu) continue;
                            float dy[VEC];
    #pragma unroll
                            for (int c = 0; c < VEC; ++c) dy[c] = -h.lr_y * (gq[c] + h.reg_y * y[k][c]);
                            apply_vec<VEC, TEAM>(m.yfeat + static_cast<size_t>(jj[k]) * fs + off, y[k], dy, ym);
                            if (gb != 0.0f && lead && lane == ((t0 + k) & (TEAM - 1)))
                                apply_delta(&m.ybias[jj[k]], m.ybias[jj[k]], -gb, ym);
                        }
                    };
                    load_batch(ts, ya, ja);
                    for (uint32_t t0 = ts; t0 < te; t0 += 8) {
                        if (t0 + 4 < te) load_batch(t0 + 4, yb, jb);
                        update_batch(t0, ya, ja);
                        if (t0 + 4 >= te) break;
                        if (t0 + 8 < te) load_batch(t0 + 8, ya, ja);
                        update_batch(t0 + 4, yb, jb);
                    }
                }
            } else {
                store_str<VEC, TEAM>(sv.gacc + prow, G);
                if (lane == 0) sv.gacc_b[u] = Gb;
            }
            if constexpr (DAY) {
                // y_day step of every run (src/tx.rs:1025-1036) in one stream of
                // the user's days: [header d][training items of d] ...; the header
                // row is the run's gradient sum already divided by its norm, and
                // y_j -= lr_yd (g + reg_yd y_j), rows written like the NSVD1 rows
                // (plain or atomic per the y update mode)
                const uint32_t y0 = m.dy_off[u];
                float g[VEC];
#pragma unroll
                for (int c = 0; c < VEC; ++c) g[c] = 0.0f;
                for_rows_at<TEAM, VEC, GBS>(
                    m.dy_off[u + 1] - y0, lane, m.dy_item + y0,
                    [&](uint32_t j) {
                        return (j & kDayHead) ? sv.day_acc + static_cast<size_t>(j & ~kDayHead) * fs + cbase
                                              : m.yfeat_day + static_cast<size_t>(j) * fs + cbase;
                    },
                    [&](uint32_t j, const float (&y)[VEC]) {
                        if (j & kDayHead) {
#pragma unroll
                            for (int c = 0; c < VEC; ++c) g[c] = y[c];
                        } else {
                            float* row = m.yfeat_day + static_cast<size_t>(j) * fs + cbase + lane * VEC;
                            if (ym == 2) {
                                float ny[VEC];
#pragma unroll
                                for (int c = 0; c < VEC; ++c) ny[c] = y[c] - h.lr_yd * (g[c] + h.reg_yd * y[c]);
                                store_vec<VEC>(row, ny);
                            } else {
#pragma unroll
                                for (int c = 0; c < VEC; ++c)
                                    apply_delta(row + c, y[c], -h.lr_yd * (g[c] + h.reg_yd * y[c]), ym);
                            }
                        }
                    });
            }
        }
        if constexpr (SPLIT) {
            wg_barrier();  // the gradient is in its slot; the next sum is in the other
            if (split_u == kNoUser) break;
            ++split_k;
        }
    }
}

// ---------------------------------------------------------------------------
// k_sgd_par launch helpers (each variant family is instantiated in its own
// translation unit: tx_sgd_lm.hip, tx_sgd_day.hip, tx_sgd_user.hip,
// tx_sgd_mw.hip)
// ---------------------------------------------------------------------------

// Dispatch on (TEAM, VEC): the shapes the host picks for a padded row stride
// (one wave per user: four 16-lane teams in a wave would run their users'
// loops to the longest of the four).  The per-day rebuild sums a few rows per
// day and takes four elements per lane instead (TX_DAY_SHAPES).
// (A build may predefine TX_SHAPES to a subset, e.g. to inspect one kernel.)
#ifndef TX_SHAPES
#define TX_SHAPES(X)  \
    X(64, 1)          \
    X(64, 2)          \
    X(64, 4)          \
    X(64, 8)          \
    X(64, 16)
#define TX_DAY_SHAPES(X) \
    X(16, 4)             \
    X(32, 4)             \
    X(64, 4)             \
    X(64, 8)             \
    X(64, 16)
#endif
#ifndef TX_DAY_SHAPES
#define TX_DAY_SHAPES(X) TX_SHAPES(X)
#endif


// one-wave teams: PASS and DAY fixed, QF / P2 / ORD chosen at run time
template <int TEAM, int VEC, int PASS, bool DAY>
inline hipError_t sgd_par_tvp(const DevModel& m, const StrataView& sv, int stratum, int blocks, bool qf, bool p2,
                              bool ord, hipStream_t st) {
    (void)hipGetLastError();
#define TX_PAR(QF, P2, ORD) \
    hipLaunchKernelGGL((k_sgd_par<TEAM, VEC, QF, P2, ORD, PASS, DAY, 0>), dim3(blocks),                          \
                       dim3(sv.block_threads > 0 ? sv.block_threads : 256), 0, st, m, sv, stratum)
    if constexpr (PASS == 0) {
        if (ord) {
            if (qf) { if (p2) TX_PAR(true, true, true); else TX_PAR(true, false, true); }
            else { if (p2) TX_PAR(false, true, true); else TX_PAR(false, false, true); }
            return hipGetLastError();
        }
    } else {
        if (ord) return hipErrorInvalidValue;
    }
    if (qf) { if (p2) TX_PAR(true, true, false); else TX_PAR(true, false, false); }
    else { if (p2) TX_PAR(false, true, false); else TX_PAR(false, false, false); }
#undef TX_PAR
    return hipGetLastError();
}

// neighbourhood models (NB 1: global w/c, NB 2: plus same-day w_day/c_day):
// q_f and p2 always on (all-zero tables leave every value as without them),
// ORD chosen at run time
template <int TEAM, int VEC, bool DAY, int NB>
inline hipError_t sgd_par_nb_tvp(const DevModel& m, const StrataView& sv, int stratum, int blocks, bool ord,
                                 hipStream_t st) {
    (void)hipGetLastError();
    const dim3 bt(sv.block_threads > 0 ? sv.block_threads : 256);
    if (ord)
        hipLaunchKernelGGL((k_sgd_par<TEAM, VEC, true, true, true, 0, DAY, 0, NB>), dim3(blocks), bt, 0, st, m, sv,
                           stratum);
    else
        hipLaunchKernelGGL((k_sgd_par<TEAM, VEC, true, true, false, 0, DAY, 0, NB>), dim3(blocks), bt, 0, st, m, sv,
                           stratum);
    return hipGetLastError();
}

// the same with multi-wave teams (f > 1024, no ordinal head)
template <bool DAY, int NB>
inline hipError_t sgd_par_nb_mw(const DevModel& m, const StrataView& sv, int teams, int waves, hipStream_t st) {
    (void)hipGetLastError();
    if (waves <= 4)
        hipLaunchKernelGGL((k_sgd_par<64, 16, true, true, false, 0, DAY, 4, NB>), dim3(teams), dim3(64 * waves), 0, st,
                           m, sv, 0);
    else
        hipLaunchKernelGGL((k_sgd_par<64, 16, true, true, false, 0, DAY, 8, NB>), dim3(teams), dim3(64 * waves), 0, st,
                           m, sv, 0);
    return hipGetLastError();
}

}  // namespace tx
