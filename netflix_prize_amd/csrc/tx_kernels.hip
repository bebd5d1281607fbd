// HIP kernels of the MI355X tsvdx4 (timeSVD++) path, gfx950.
//
//   k_sgd_serial     one wavefront replays TxModel::fit_epoch
//                    (src/tx.rs:734-1115) user by user in the reference's
//                    order with the reference's float operation order.
//   k_sgd_par        all users in parallel: one team of TEAM lanes per user,
//                    VEC factor elements per lane, user state in registers,
//                    item updates as device-scope atomic deltas; run once
//                    per stratum of the rotating user-block x item-block
//                    schedule.
//   k_rebuild_ycache[_day]  epoch-end NSVD1 caches (src/tx.rs:391-454).
//   k_predict        predict / predict_subscores / calc_rmse
//                    (src/tx.rs:303-339, 663-724; src/lib.rs:659-673).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "tx_launch.hpp"
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

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        y[e] = to_bf16(x[e]);
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
// Serial schedule: the reference trainer, one wavefront.
// Lane l owns factor elements k = l, l+64, ...; scalars are evaluated by every
// lane from the same loads and stored by lane 0.  The dot product is summed
// sequentially over k by lane 0, so every value matches the reference bit for
// bit (contraction is off for this kernel).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(64) void k_sgd_serial(DevModel m, const int64_t* __restrict__ order,
                                                   float* __restrict__ scratch) {
#pragma clang fp contract(off)
    extern __shared__ float prod[];  // [fs + 1]
    const int lane = threadIdx.x;
    const int f = m.f, fs = m.fs, F = m.F, B = m.B;
    const bool low_mem = m.low_mem != 0;
    const Hyper h = m.h;
    float* su = scratch;
    float* G = scratch + fs;
    float* Gd = scratch + 2 * fs;
    float* sud = scratch + 3 * fs;
    const bool has_ud = !low_mem && h.lr_ud != 0.0f;

    for (int ui = 0; ui < m.U; ++ui) {
        const int u = static_cast<int>(order[ui]);
        const uint32_t start = m.t_off[u], end = m.t_off[u + 1];
        if (start == end) continue;
        const float norm = m.norm_train[u];

        // NSVD1 user aggregation (src/tx.rs:801-822)
        for (int k = lane; k < f; k += 64) {
            float acc = 0.0f;
            for (uint32_t t = start; t < end; ++t) acc += m.yfeat[static_cast<size_t>(m.t_item[t]) * fs + k];
            if (m.full_su)
                for (uint32_t q = m.h_start[u]; q < m.h_start[u + 1]; ++q)
                    acc += m.yfeat[static_cast<size_t>(m.h_item[q]) * fs + k];
            acc = acc / norm;
            su[k] = acc;
            m.ycache[static_cast<size_t>(u) * fs + k] = acc;
            G[k] = 0.0f;
            Gd[k] = 0.0f;
            sud[k] = 0.0f;
        }
        float su_bias = 0.0f;
        for (uint32_t t = start; t < end; ++t) su_bias += m.ybias[m.t_item[t]];
        if (m.full_su)
            for (uint32_t q = m.h_start[u]; q < m.h_start[u + 1]; ++q) su_bias += m.ybias[m.h_item[q]];
        su_bias = su_bias / norm;
        if (lane == 0) m.ycache_bias[u] = su_bias;

        float sum_err = 0.0f;
        float norm_day = 0.0f;
        int cur_fb = 0;
        uint32_t run_begin = start;

        for (uint32_t t = start; t < end; ++t) {
            const uint32_t i = m.t_item[t];
            const float r = m.t_r[t];
            const int b = m.t_bin[t];
            const float dev = m.t_dev[t];
            const uint32_t ud = m.t_ud[t];

            if (!low_mem && (t == start || m.t_run_stop[t - 1] != m.t_run_stop[t])) {
                // start of a same-day run (src/tx.rs:846-871)
                run_begin = t;
                const uint32_t stop = m.t_run_stop[t];
                cur_fb = m.sud_fbin[ud];
                for (int k = lane; k < f; k += 64) {
                    Gd[k] = 0.0f;
                    float acc = 0.0f;
                    for (uint32_t td = t; td <= stop; ++td)
                        acc += m.yfeat_day[static_cast<size_t>(m.t_item[td]) * fs + k];
                    if (m.full_su)
                        for (uint32_t q = m.hd_start[ud]; q < m.hd_start[ud + 1]; ++q)
                            acc += m.yfeat_day[static_cast<size_t>(m.hd_item[q]) * fs + k];
                    sud[k] = acc / m.t_run_norm[t];
                }
                norm_day = m.t_run_norm[t];
            }
            const int fb = low_mem ? 0 : cur_fb;

            // forward (src/tx.rs:878-901)
            const float bu_day = m.ubias_day[ud];
            const float bu_t = m.ubias[u] + m.but_bin[static_cast<size_t>(u) * B + b] + m.alpha_u[u] * dev + bu_day
                               + m.ycache_bias[u];
            const float bi_t = m.ibias[static_cast<size_t>(i) * m.ibs] + m.bit_bin[static_cast<size_t>(i) * m.ibs + b]
                               + m.ibias_freq[static_cast<size_t>(i) * F + fb];
            const float cu_t = m.cu[u] + m.cut[ud];
            const size_t qrow = static_cast<size_t>(i) * fs;
            const size_t qfrow = (static_cast<size_t>(i) * F + fb) * fs;
            const size_t prow = static_cast<size_t>(u) * fs;
            for (int k = lane; k < f; k += 64) {
                const float qi_eff = m.ifeat[qrow + k] + m.ifeat_freq[qfrow + k];
                float pu_eff = m.ufeat[prow + k] + su[k] + dev * m.ufeat2[prow + k];
                if (!low_mem) {
                    pu_eff += m.ufeat_day[static_cast<size_t>(ud) * fs + k];
                    pu_eff += sud[k] + 0.0f;
                }
                prod[k] = pu_eff * qi_eff;
            }
            __syncthreads();
            if (lane == 0) {
                float d = 0.0f;
                for (int k = 0; k < f; ++k) d += prod[k];
                prod[fs] = d;
            }
            __syncthreads();
            const float dot = prod[fs];
            __syncthreads();
            const float bias_part = m.gbias + bu_t + bi_t * cu_t;
            const float score = bias_part + dot + 0.0f;

            float err;
            if (m.has_ord) {
                float th[4], g_t[4];
                for (int k = 0; k < 4; ++k) th[k] = m.thresholds[k];
                err = ordinal_grad(th, score, static_cast<int>(m.t_raw[t]), g_t);
                if (lane == 0)
                    for (int k = 0; k < 4; ++k) m.thresholds[k] = th[k] - h.lr_t * (g_t[k] + h.reg_t * th[k]);
            } else {
                err = score - r;
            }
            const float err_b = err * h.w_bias;
            const float err_d = err * h.w_factor;

            // bias steps (src/tx.rs:981-990)
            if (lane == 0) {
                m.ubias[u] = m.ubias[u] - h.lr_ub * err_b;
                float* bb = &m.but_bin[static_cast<size_t>(u) * B + b];
                *bb = *bb - h.lr_tu * err_b;
                m.alpha_u[u] = m.alpha_u[u] - h.lr_ta * err_b * dev;
                m.ubias_day[ud] = m.ubias_day[ud] - h.lr_ubd * err_b;
                m.ibias[static_cast<size_t>(i) * m.ibs] = m.ibias[static_cast<size_t>(i) * m.ibs] - h.lr_ib * (err_b * cu_t);
                float* tb = &m.bit_bin[static_cast<size_t>(i) * m.ibs + b];
                *tb = *tb - h.lr_ti * (err_b * cu_t);
                float* fbp = &m.ibias_freq[static_cast<size_t>(i) * F + fb];
                *fbp = *fbp - h.lr_ibf * (err_b * cu_t);
                m.cu[u] = m.cu[u] - h.lr_cu * (err_b * bi_t + h.reg_cu * (m.cu[u] - 1.0f));
            }

            // factor steps (src/tx.rs:995-1023)
            if (!m.sum_err_bug) sum_err += err_b;
            else
                for (int k = 0; k < f; ++k) sum_err += err_b;
            for (int k = lane; k < f; k += 64) {
                const float qi_k = m.ifeat[qrow + k];
                const float qf_k = m.ifeat_freq[qfrow + k];
                const float qi_eff_k = qi_k + qf_k;
                const float pu_k = m.ufeat[prow + k];
                const float pu2_k = m.ufeat2[prow + k];
                G[k] += err_d * qi_eff_k;
                if (!low_mem) Gd[k] += err_d * qi_eff_k;
                m.ufeat[prow + k] = pu_k - h.lr_u * (err_d * qi_eff_k + h.reg_u * pu_k);
                m.ufeat2[prow + k] = pu2_k - h.lr_u2 * (err_d * qi_eff_k * dev + h.reg_u2 * pu2_k);
                float pud_k = 0.0f;
                if (has_ud) {
                    float* pd = &m.ufeat_day[static_cast<size_t>(ud) * fs + k];
                    const float pud = *pd;
                    *pd = pud - h.lr_ud * (err_d * qi_eff_k + h.reg_ud * pud);
                    pud_k = pud;
                }
                const float pu_eff = pu_k + dev * pu2_k + pud_k + su[k] + (low_mem ? 0.0f : sud[k] + 0.0f);
                m.ifeat[qrow + k] = qi_k - h.lr_i * (err_d * pu_eff + h.reg_i * qi_k);
                m.ifeat_freq[qfrow + k] = qf_k - h.lr_iqf * (err_d * pu_eff + h.reg_iqf * qf_k);
            }

            // end of a same-day run (src/tx.rs:1025-1036)
            if (!low_mem && t == m.t_run_stop[t]) {
                for (uint32_t td = run_begin; td <= t; ++td) {
                    const size_t yr = static_cast<size_t>(m.t_item[td]) * fs;
                    for (int k = lane; k < f; k += 64) {
                        const float yj = m.yfeat_day[yr + k];
                        m.yfeat_day[yr + k] = yj - h.lr_yd * (Gd[k] / norm_day + h.reg_yd * yj);
                    }
                }
            }
        }

        // end of user (src/tx.rs:1099-1106)
        for (uint32_t t = start; t < end; ++t) {
            const uint32_t j = m.t_item[t];
            const size_t yr = static_cast<size_t>(j) * fs;
            for (int k = lane; k < f; k += 64) {
                const float yj = m.yfeat[yr + k];
                m.yfeat[yr + k] = yj - h.lr_y * (G[k] / norm + h.reg_y * yj);
            }
            if (lane == 0) m.ybias[j] = m.ybias[j] - h.lr_yb * (sum_err / norm);
        }
    }
}

// ---------------------------------------------------------------------------
// Plain schedule: the per-user wavefront kernel with nothing else.  Wave w of
// the grid takes users order[w], order[w + waves], ... (the epoch's user order,
// no queue); per user it sums the NSVD1 rows of N(u), then walks the ratings
// in file order: loads the q row and the scalars the rating needs and waits for
// them, a wave-reduced dot product, bias steps (user-owned in registers, item
// biases by atomic add), p in registers, q by atomic add of its SGD delta; and
// at the user's end the y step over R(u) (src/tx.rs:734-1115).  No prefetch,
// no length buckets, no graph: the reference point of the batched kernel.
// Low-memory models without ordinal head; lane l holds elements l, l + 64, ...
// ---------------------------------------------------------------------------

template <int VEC>
__global__ __launch_bounds__(64) void k_sgd_plain(DevModel m, const int64_t* __restrict__ order, int qf_on) {
    const int lane = threadIdx.x;
    const int fs = m.fs, F = m.F, B = m.B;
    const Hyper h = m.h;
    // qf_on: the q_f rows (bin 0 in low-memory training) are not all zero
    for (int64_t w = blockIdx.x; w < m.U; w += gridDim.x) {
        const int u = static_cast<int>(order[w]);
        const uint32_t ts = m.t_off[u], te = m.t_off[u + 1];
        if (ts == te) continue;
        const float norm = m.norm_train[u];
        const size_t prow = static_cast<size_t>(u) * fs + lane;
        float p[VEC], p2[VEC], s[VEC], G[VEC];
        float sb = 0.0f;
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            p[c] = m.ufeat[prow + c * 64];
            p2[c] = m.ufeat2[prow + c * 64];
            s[c] = 0.0f;
            G[c] = 0.0f;
        }
        // NSVD1 sum over the training items and (full_su) the held-out items
        const uint32_t hs = m.full_su ? m.h_start[u] : 0u, he = m.full_su ? m.h_start[u + 1] : 0u;
        for (uint32_t r = ts; r < te + (he - hs); ++r) {
            const uint32_t j = r < te ? m.t_item[r] : m.h_item[hs + (r - te)];
            const float* y = m.yfeat + static_cast<size_t>(j) * fs + lane;
#pragma unroll
            for (int c = 0; c < VEC; ++c) s[c] += y[c * 64];
            sb += m.ybias[j];
        }
        const float inv = 1.0f / norm;
#pragma unroll
        for (int c = 0; c < VEC; ++c) s[c] *= inv;
        const float su_b = sb * inv;
        float ubias = m.ubias[u], alpha = m.alpha_u[u], cu = m.cu[u], Gb = 0.0f;
        for (uint32_t t = ts; t < te; ++t) {
            const uint32_t i = m.t_item[t], ud = m.t_ud[t];
            const int b = m.t_bin[t];
            const float dev = m.t_dev[t], r = m.t_r[t];
            const float* qrow = m.ifeat + static_cast<size_t>(i) * fs + lane;
            const float* frow = m.ifeat_freq + static_cast<size_t>(i) * F * fs + lane;
            float q[VEC], qf[VEC];
            float part = 0.0f;
#pragma unroll
            for (int c = 0; c < VEC; ++c) {
                q[c] = qrow[c * 64];
                qf[c] = qf_on ? frow[c * 64] : 0.0f;
                part += (q[c] + qf[c]) * (p[c] + s[c] + dev * p2[c]);
            }
            const float dot = team_sum<64>(part);
            const float c_bud = m.ubias_day[ud];
            const float c_but = m.but_bin[static_cast<size_t>(u) * B + b];
            const float bi_t = m.ibias[static_cast<size_t>(i) * m.ibs] + m.bit_bin[static_cast<size_t>(i) * m.ibs + b] + m.ibias_freq[static_cast<size_t>(i) * F];
            const float cu_t = cu + m.cut[ud];
            const float bu_t = ubias + c_but + alpha * dev + c_bud + su_b;
            const float err = m.gbias + bu_t + bi_t * cu_t + dot - r;
            const float err_b = err * h.w_bias, err_d = err * h.w_factor;
            ubias -= h.lr_ub * err_b;
            alpha -= h.lr_ta * err_b * dev;
            if (lane == 0) {
                m.ubias_day[ud] = c_bud - h.lr_ubd * err_b;
                m.but_bin[static_cast<size_t>(u) * B + b] = c_but - h.lr_tu * err_b;
                atomic_add_dev(&m.ibias[static_cast<size_t>(i) * m.ibs], -h.lr_ib * (err_b * cu_t));
                atomic_add_dev(&m.bit_bin[static_cast<size_t>(i) * m.ibs + b], -h.lr_ti * (err_b * cu_t));
                atomic_add_dev(&m.ibias_freq[static_cast<size_t>(i) * F], -h.lr_ibf * (err_b * cu_t));
            }
            cu -= h.lr_cu * (err_b * bi_t + h.reg_cu * (cu - 1.0f));
            Gb += m.sum_err_bug ? err_b * static_cast<float>(m.f) : err_b;
#pragma unroll
            for (int c = 0; c < VEC; ++c) {
                const float qe = q[c] + qf[c];
                G[c] += err_d * qe;
                const float pe = p[c] + dev * p2[c] + s[c];
                p[c] -= h.lr_u * (err_d * qe + h.reg_u * p[c]);
                p2[c] -= h.lr_u2 * (err_d * qe * dev + h.reg_u2 * p2[c]);
                atomic_add_dev(m.ifeat + static_cast<size_t>(i) * fs + lane + c * 64, -h.lr_i * (err_d * pe + h.reg_i * q[c]));
                if (qf_on)
                    atomic_add_dev(m.ifeat_freq + static_cast<size_t>(i) * F * fs + lane + c * 64,
                                   -h.lr_iqf * (err_d * pe + h.reg_iqf * qf[c]));
            }
        }
        // end of user: the y step over R(u) (src/tx.rs:1099-1106)
        const float gb = h.lr_yb * (Gb * inv);
        for (uint32_t t = ts; t < te; ++t) {
            const uint32_t j = m.t_item[t];
            float* y = m.yfeat + static_cast<size_t>(j) * fs + lane;
#pragma unroll
            for (int c = 0; c < VEC; ++c) y[c * 64] = y[c * 64] - h.lr_y * (G[c] * inv + h.reg_y * y[c * 64]);
            if (lane == 0 && gb != 0.0f) m.ybias[j] = m.ybias[j] - gb;
        }
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            m.ufeat[prow + c * 64] = p[c];
            m.ufeat2[prow + c * 64] = p2[c];
        }
        if (lane == 0) {
            m.ubias[u] = ubias;
            m.alpha_u[u] = alpha;
            m.cu[u] = cu;
        }
    }
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
template <int VEC, bool DAY>
struct Pending {
    uint32_t i, ud, fb;
    float dev, r;
    int b, raw;
    float q[VEC], qf[VEC];
    // day terms: p_u,day, the day's NSVD1 sum over y_day (strided) and the run norm
    float pd[DAY ? VEC : 1], sd[DAY ? VEC : 1];
    float nd;
    float sc, sc2;
    float* sc_addr;  // this lane's slot of sc (read at fetch, updated after the rating)
    float bud, but;
    bool has_bud, has_but;
};

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
template <int TEAM, int VEC, bool QF, bool P2, bool ORD, int PASS, bool DAY, int MW>
__global__ __launch_bounds__(MW == 8 ? 512 : 256, MW ? 1 : TX_SGD_WAVES) void k_sgd_par(DevModel m, StrataView sv, int stratum) {
    static_assert(PASS == 0 || !ORD, "the user passes have no ordinal head");
    static_assert(PASS == 0 || !DAY, "the user passes train low-memory models");
    static_assert(!MW || TEAM == 64, "multi-wave teams: 64-lane waves");
    constexpr int CHUNK = 4;
    // up to 4 waves share a SIMD's register file with nothing else; with 8,
    // two waves share each SIMD and one rating ahead keeps them closer to it
    constexpr int DEPTH = MW == 8 ? 1 : TX_PREFETCH;
    __shared__ float xbuf[(MW == 8 ? 512 : 256) * VEC];  // per-wave transpose slot, TEAM*VEC floats each
    __shared__ float dred[2][MW ? 9 : 1];            // per-wave dot products + wave 0's b_i(t) (MW), double-buffered
    __shared__ uint32_t qgrab[1];                    // the team's queue grab (MW)
    const int lane = threadIdx.x % TEAM;
    const int wv = MW ? static_cast<int>(threadIdx.x / 64) : 0;  // wave of the team
    const int nwv = MW ? static_cast<int>(blockDim.x / 64) : 1;
    const int cbase = wv * TEAM * VEC;                            // first element of the wave's chunk
    const bool lead = wv == 0;                                    // issues the shared one-off writes
    uint32_t dpar = 0;
    const int S = sv.n_strata;
    const int cell = sv.n_cells > 1 ? static_cast<int>(blockIdx.x % sv.n_cells) : 0;
    const int qid = stratum * sv.n_cells + cell;
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
    const int qm = sv.q_mode, ym = sv.y_mode;
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

    float th[4] = {0.f, 0.f, 0.f, 0.f};
    if constexpr (ORD) {
#pragma unroll
        for (int k = 0; k < 4; ++k) th[k] = m.thresholds[k];
    }

    for (;;) {
        uint32_t w0 = 0;
        if constexpr (MW) {
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
        if (w0 >= qlen) break;
        const uint32_t w1 = w0 + CHUNK < qlen ? w0 + CHUNK : qlen;
        for (uint32_t w = w0; w < w1; ++w) {
            const uint32_t u = sv.q_users[qbeg + w];
            const uint32_t s0 = sv.seg_off[static_cast<size_t>(u) * (S + 1) + stratum];
            const uint32_t s1 = sv.seg_off[static_cast<size_t>(u) * (S + 1) + stratum + 1];
            const size_t prow = static_cast<size_t>(u) * fs + off;
            const uint32_t ts = m.t_off[u], te = m.t_off[u + 1];
            const float norm = m.norm_train[u];
            const size_t ubrow = static_cast<size_t>(u) * B;

            // issue the loads one rating needs, from its packed record
            auto fetch = [&](uint4 rc, Pending<VEC, DAY>& n) {
                n.i = rc.x & 0xFFFFFu;
                n.raw = static_cast<int>((rc.x >> 20) & 15u);
                n.b = static_cast<int>(rc.x >> 24);
                // the freq bin is packed above the day index for day-term models
                // only (low-memory records carry bin 0, so y is the index)
                n.ud = DAY ? (rc.y & 0x0FFFFFFFu) : rc.y;
                n.fb = DAY ? (rc.y >> 28) : 0u;
                n.dev = __uint_as_float(rc.z);
                n.r = __uint_as_float(rc.w);
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
            if (PASS != 0 && first && sv.su_cached) {
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
                for_rows<TEAM, VEC, row_group<TEAM, VEC>()>(
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
                    for_rows_at<TEAM, VEC, row_group<TEAM, VEC>()>(
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
            Pending<VEC, DAY> ring[RING];
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
                if constexpr (DAY) ring[k].nd = m.t_run_norm[s0 + k < s1 ? s0 + k : s1 - 1];
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
                const Pending<VEC, DAY>& cur = ring[k];
                fetch(rec_at(t + DEPTH < s1 ? t + DEPTH : s1 - 1), ring[(k + DEPTH) % RING]);
                if constexpr (DAY) ring[(k + DEPTH) % RING].nd = m.t_run_norm[t + DEPTH < s1 ? t + DEPTH : s1 - 1];
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

                // user-owned biases: registers, or written back now and forwarded
                // to every prefetched rating that reads the same slot
                ubias -= h.lr_ub * err_b;
                alpha -= h.lr_ta * err_b * dev;
                const float but_new = c_but - h.lr_tu * err_b;
                const float bud_new = c_bud - h.lr_ubd * err_b;
#pragma unroll
                for (int jj = 1; jj <= DEPTH; ++jj) {
                    Pending<VEC, DAY>& nx = ring[(k + jj) % RING];  // ratings t+1 .. t+DEPTH
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
                    for_rows<TEAM, VEC, row_group<TEAM, VEC>()>(
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
                for_rows_at<TEAM, VEC, row_group<TEAM, VEC>()>(
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
    }
}

// ---------------------------------------------------------------------------
// Tsvdx4pModel: NSVD1 delta apply and the item pass
// ---------------------------------------------------------------------------

// The NSVD1 step of the tsvdx4p user pass (apply_nsvd_deltas,
// src/tsvdx4p.rs:341-385) as a gather: item j's summed gradient is the sum of
// the stored gradients G_u / norm_u of the users who rated it, read in the
// by-item order (the same order every epoch, so the step is deterministic and
// needs no atomics), then y_j -= lr_y (D_j / s_j + reg_y y_j) and ybias_j -=
// lr_yb Db_j / s_j with s_j = n_contrib^nsvd_norm_exp (0: no contributors,
// no step).  One wave per item; rows wider than a wave (f > 1024) a chunk of
// 64 * VEC elements at a time.
template <int VEC>
__global__ __launch_bounds__(256) void k_nsvd_gather(DevModel m, ItemView iv, const float* __restrict__ gacc,
                                                     const float* __restrict__ gacc_b, const float* __restrict__ scale) {
#pragma clang fp contract(off)
    constexpr int TEAM = 64;
    const int lane = threadIdx.x % TEAM;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / TEAM;
    const int nwaves = gridDim.x * blockDim.x / TEAM;
    const int fs = m.fs;
    for (int j = wave; j < m.I; j += nwaves) {
        const uint32_t r0 = iv.bi_start[j], n = iv.bi_start[j + 1] - r0;
        const float sc = scale[j];
        if (n == 0 || !(sc > 0.0f)) continue;
        for (int cb = 0; cb < fs; cb += TEAM * VEC) {
            const bool first = cb == 0;
            float acc[VEC];
            float sb = 0.0f;
#pragma unroll
            for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
            for_rows<TEAM, VEC, row_group<TEAM, VEC>()>(
                n, lane, gacc + cb, fs, [&](uint32_t r) { return iv.bi_user + r0 + r; },
                [&](uint32_t b0, uint32_t ib) {
                    if (first && b0 + static_cast<uint32_t>(lane) < n) sb += gacc_b[ib];
                },
                [&](uint32_t, uint32_t, const float (&g)[VEC]) {
#pragma unroll
                    for (int c = 0; c < VEC; ++c) acc[c] += g[c];
                });
            float* y = m.yfeat + static_cast<size_t>(j) * fs + cb + lane * VEC;
            float yv[VEC];
            load_vec<VEC>(y, yv);
#pragma unroll
            for (int c = 0; c < VEC; ++c)
                if (cb + lane * VEC + c < m.f) yv[c] = yv[c] - m.h.lr_y * (acc[c] / sc + m.h.reg_y * yv[c]);
            store_vec<VEC>(y, yv);
            if (first) {
                const float db = team_sum_dpp<TEAM>(sb);
                if (lane == 0) m.ybias[j] = m.ybias[j] - m.h.lr_yb * (db / sc);
            }
        }
    }
}

// VEC contiguous FP32 words as raw bits.
template <int VEC>
__device__ __forceinline__ void load_raw(const float* __restrict__ p, uint32_t (&out)[VEC]) {
    if constexpr (VEC % 4 == 0) {
#pragma unroll
        for (int c = 0; c < VEC; c += 4) {
            const uint4 x = *reinterpret_cast<const uint4*>(p + c);
            out[c] = x.x; out[c + 1] = x.y; out[c + 2] = x.z; out[c + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < VEC; ++c) out[c] = __float_as_uint(p[c]);
    }
}

// VEC contiguous BF16 values as raw words, two per word in out[0 .. VEC/2)
// (one value in out[0] for VEC = 1).
template <int VEC>
__device__ __forceinline__ void load_bf16_raw(const uint16_t* __restrict__ p, uint32_t (&out)[VEC]) {
    if constexpr (VEC == 1) {
        out[0] = p[0];
    } else if constexpr (VEC == 2) {
        out[0] = *reinterpret_cast<const uint32_t*>(p);
    } else if constexpr (VEC == 4) {
        const uint2 x = *reinterpret_cast<const uint2*>(p);
        out[0] = x.x; out[1] = x.y;
    } else {
#pragma unroll
        for (int c = 0; c < VEC / 2; c += 4) {
            const uint4 x = *reinterpret_cast<const uint4*>(p + 2 * c);
            out[c] = x.x; out[c + 1] = x.y; out[c + 2] = x.z; out[c + 3] = x.w;
        }
    }
}

// Element c of a row loaded by load_raw (bf = false) or load_bf16_raw (bf = true).
template <int VEC>
__device__ __forceinline__ float widen(const uint32_t (&raw)[VEC], int c, bool bf) {
    if (!bf) return __uint_as_float(raw[c]);
    const uint32_t w = raw[c >> 1];
    return __uint_as_float((c & 1) ? (w & 0xffff0000u) : (w << 16));
}

// One rating of the item pass and its frozen user-side inputs.
//   sc:  lane 0 b_u, 1 b_u,bin, 2 alpha_u, 3 b_u,day
//   sc2: lane 0 ycache_bias, 1 c_u, 2 c_u,day
template <int VEC, bool P2, bool QF>
struct ItemPending {
    uint32_t b, fb;
    float dev, r;
    // the rating's user rows as loaded, contiguous layout: p, ycache (and p2)
    // as FP32 words or, from the BF16 copies, two values per word in the
    // first VEC/2 words; widened where they are used, so no instruction here
    // waits for their loads
    uint32_t p[VEC], s[VEC], p2[P2 ? VEC : 1];
    float sc, sc2;
    // q_f row of the rating's freq bin as loaded, and the row written by an
    // earlier rating of the item after that load (has_qf), forwarded here so
    // the loaded registers are never written while their load is in flight
    float qf[QF ? VEC : 1], qfo[QF ? VEC : 1];
    bool has_qf;
};

// fit_epoch_item (src/tsvdx4p.rs:644-762): one team per item (longest first;
// items are independent), the item's ratings in training-file order with every
// user-side value frozen.  q_i stays in registers and the item's bias slots
// (b_i, b_{i,bin}[B], b_{i,f}[F]) in the team's LDS for the whole item; both
// are written back once.  Low-memory form; with QF the q_f rows of the day's
// freq bin (the real bin here, bin 0 in the user pass: src/tsvdx4p.rs:497,
// 708) are read and written per rating.
// MW (4 or 8; 0 = one wave): a team of blockDim.x / 64 waves per item, wave w
// holding elements [1024w, 1024w + 1024) of the rows (f > 1024); the waves
// join each rating's dot product through LDS, keep their own copies of the
// item's bias slots (they compute the same values) and wave 0 writes them back.
template <int TEAM, int VEC, bool P2, bool QF, int MW>
__global__ __launch_bounds__(MW == 8 ? 512 : 256) void k_item_pass(DevModel m, ItemView iv) {
    static_assert(!MW || TEAM == 64, "multi-wave teams: 64-lane waves");
    constexpr int DEPTH = MW == 8 ? 1 : TX_PREFETCH;
    constexpr int RING = DEPTH + 1;
    static_assert(TEAM % RING == 0, "the prefetch depth plus one must divide the team size");
    extern __shared__ float ish[];
    __shared__ float dred[2][MW ? 8 : 1];
    __shared__ uint32_t qgrab[1];
    const int lane = threadIdx.x % TEAM;
    const int wv = MW ? static_cast<int>(threadIdx.x / 64) : 0;
    const int nwv = MW ? static_cast<int>(blockDim.x / 64) : 1;
    const bool lead = wv == 0;
    uint32_t dpar = 0;
    const int B = m.B, F = m.F, fs = m.fs;
    float* bt = ish + (threadIdx.x / TEAM) * (B + F);  // per wave: b_{i,bin}[B], then b_{i,f}[F]
    float* bq = bt + B;
    const int coff = wv * TEAM * VEC + lane * VEC;
    const Hyper h = m.h;
    const bool bfz = m.bf16_frozen != 0;  // p, p2, ycache read from BF16 copies
    // lane-indexed user scalar slots: base + u*umul + (b & bm) + (ud & um)
    const float* sc_base = lane == 1 ? m.but_bin : lane == 2 ? m.alpha_u : lane == 3 ? m.ubias_day : m.ubias;
    const uint32_t sc_umul = lane == 1 ? static_cast<uint32_t>(B) : lane == 3 ? 0u : 1u;
    const uint32_t sc_bm = lane == 1 ? 0xffffffffu : 0u;
    const uint32_t sc_um = lane == 3 ? 0xffffffffu : 0u;
    const float* sc2_base = lane == 1 ? m.cu : lane == 2 ? m.cut : m.ycache_bias;
    const uint32_t sc2_umul = lane == 2 ? 0u : 1u;
    const uint32_t sc2_um = lane == 2 ? 0xffffffffu : 0u;

    for (;;) {
        uint32_t w = 0;
        if constexpr (MW) {
            wg_barrier();
            if (lead && lane == 0) qgrab[0] = __hip_atomic_fetch_add(iv.q_head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wg_barrier();
            w = qgrab[0];
        } else if (lane == 0) {
            w = __hip_atomic_fetch_add(iv.q_head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if constexpr (TEAM == 64) w = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(w)));
        else w = __shfl(w, 0, TEAM);
        if (w >= iv.n_items) break;
        const uint32_t i = iv.q_items[w];
        const uint32_t r0 = iv.bi_start[i], r1 = iv.bi_start[i + 1];
        float q[VEC];
        load_vec<VEC>(m.ifeat + static_cast<size_t>(i) * fs + coff, q);
        float ib = m.ibias[static_cast<size_t>(i) * m.ibs];
        for (int l = lane; l < B; l += TEAM) bt[l] = m.bit_bin[static_cast<size_t>(i) * m.ibs + l];
        for (int l = lane; l < F; l += TEAM) bq[l] = m.ibias_freq[static_cast<size_t>(i) * F + l];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        auto fetch = [&](uint4 rc, ItemPending<VEC, P2, QF>& n) {
            const uint32_t u = rc.x & 0xFFFFFu;
            n.b = (rc.x >> 20) & 0xFFu;
            n.fb = rc.x >> 28;
            if constexpr (QF) {
                load_vec<VEC>(m.ifeat_freq + (static_cast<size_t>(i) * F + n.fb) * fs + coff, n.qf);
                n.has_qf = false;
            }
            n.dev = __uint_as_float(rc.z);
            n.r = __uint_as_float(rc.w);
            const size_t urow = static_cast<size_t>(u) * fs + coff;
            if (bfz) {
                load_bf16_raw<VEC>(m.b_ufeat + urow, n.p);
                load_bf16_raw<VEC>(m.b_ycache + urow, n.s);
                if constexpr (P2) load_bf16_raw<VEC>(m.b_ufeat2 + urow, n.p2);
            } else {
                load_raw<VEC>(m.ufeat + urow, n.p);
                load_raw<VEC>(m.ycache + urow, n.s);
                if constexpr (P2) load_raw<VEC>(m.ufeat2 + urow, n.p2);
            }
            n.sc = sc_base[static_cast<size_t>(u) * sc_umul + (n.b & sc_bm) + (rc.y & sc_um)];
            n.sc2 = sc2_base[static_cast<size_t>(u) * sc2_umul + (rc.y & sc2_um)];
        };
        if (r0 < r1) {
            // the item's rating records TEAM at a time (lane l holds record
            // tb + l), two batches ahead: a fetch takes its record from a
            // register loaded a batch or more ago instead of loading it and
            // waiting for it (which would drain the whole prefetch ring)
            auto load_recs = [&](uint32_t tb) {
                const uint32_t t = tb + static_cast<uint32_t>(lane);
                return iv.bi_rec[t < r1 ? t : r1 - 1];
            };
            uint4 cb = load_recs(r0), nb = load_recs(r0 + TEAM), fb = load_recs(r0 + 2 * TEAM);
            uint32_t tb = r0;
            auto rec_at = [&](uint32_t t) {  // record t of the current batch or the next
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
            ItemPending<VEC, P2, QF> ring[RING];
#pragma unroll
            for (int k = 0; k < DEPTH; ++k) fetch(rec_at(r0 + k < r1 ? r0 + k : r1 - 1), ring[k]);
            for (; tb < r1; tb += TEAM) {
            for (uint32_t t0 = tb; t0 < tb + TEAM && t0 < r1; t0 += RING) {
#pragma unroll
                for (int k = 0; k < RING; ++k) {
                    const uint32_t t = t0 + k;
                    if (t >= r1) break;
                    const ItemPending<VEC, P2, QF>& cur = ring[k];
                    fetch(rec_at(t + DEPTH < r1 ? t + DEPTH : r1 - 1), ring[(k + DEPTH) % RING]);
                    float pe[VEC];  // p + ycache + dev * p2
#pragma unroll
                    for (int c = 0; c < VEC; ++c) {
                        const float pv = widen<VEC>(cur.p, c, bfz), sv = widen<VEC>(cur.s, c, bfz);
                        if constexpr (P2) pe[c] = pv + sv + cur.dev * widen<VEC>(cur.p2, c, bfz);
                        else pe[c] = pv + sv;
                    }
                    const float c_ub = team_bcast_f<TEAM>(cur.sc, 0), c_but = team_bcast_f<TEAM>(cur.sc, 1);
                    const float c_al = team_bcast_f<TEAM>(cur.sc, 2), c_bud = team_bcast_f<TEAM>(cur.sc, 3);
                    const float c_ycb = team_bcast_f<TEAM>(cur.sc2, 0), c_cu = team_bcast_f<TEAM>(cur.sc2, 1);
                    const float c_cut = team_bcast_f<TEAM>(cur.sc2, 2);
                    const float c_bt = bt[cur.b], c_bq = bq[cur.fb];
                    const float bu_t = c_ub + c_but + c_al * cur.dev + c_bud + c_ycb;
                    const float bi_t = ib + c_bt + c_bq;
                    const float cu_t = c_cu + c_cut;
                    float qf[VEC];  // the q_f row (zero without QF)
                    if constexpr (QF) {
#pragma unroll
                        for (int c = 0; c < VEC; ++c) qf[c] = cur.has_qf ? cur.qfo[c] : cur.qf[c];
                    } else {
#pragma unroll
                        for (int c = 0; c < VEC; ++c) qf[c] = 0.0f;
                    }
                    float part = 0.0f;
#pragma unroll
                    for (int c = 0; c < VEC; ++c) part += pe[c] * (q[c] + qf[c]);
                    float dot = team_sum_dpp<TEAM>(part);
                    if constexpr (MW) {
                        if (lane == 0) dred[dpar][wv] = dot;
                        wg_barrier();
                        float d = 0.0f;
                        for (int k2 = 0; k2 < nwv; ++k2) d += dred[dpar][k2];
                        dot = d;
                        dpar ^= 1u;
                    }
                    const float err = m.gbias + bu_t + bi_t * cu_t + dot - cur.r;
                    const float gi = err * cu_t;
                    ib -= h.lr_ib * gi;
                    bt[cur.b] = c_bt - h.lr_ti * gi;   // every lane writes the same value
                    bq[cur.fb] = c_bq - h.lr_ibf * gi;
#pragma unroll
                    for (int c = 0; c < VEC; ++c) q[c] -= h.lr_i * (err * pe[c] + h.reg_i * q[c]);
                    if constexpr (QF) {
                        float nq[VEC];
#pragma unroll
                        for (int c = 0; c < VEC; ++c) nq[c] = qf[c] - h.lr_iqf * (err * pe[c] + h.reg_iqf * qf[c]);
                        store_vec<VEC>(m.ifeat_freq + (static_cast<size_t>(i) * F + cur.fb) * fs + coff, nq);
                        // ratings t+1 .. t+DEPTH were loaded before this store
#pragma unroll
                        for (int jj = 1; jj <= DEPTH; ++jj) {
                            ItemPending<VEC, P2, QF>& nx = ring[(k + jj) % RING];
                            if (nx.fb == cur.fb) {
                                nx.has_qf = true;
#pragma unroll
                                for (int c = 0; c < VEC; ++c) nx.qfo[c] = nq[c];
                            }
                        }
                    }
                }
            }
            cb = nb;
            nb = fb;
            fb = load_recs(tb + 3 * TEAM);
            }
        }
        store_vec<VEC>(m.ifeat + static_cast<size_t>(i) * fs + coff, q);
        if (lane == 0 && lead) m.ibias[static_cast<size_t>(i) * m.ibs] = ib;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lead) {
            for (int l = lane; l < B; l += TEAM) m.bit_bin[static_cast<size_t>(i) * m.ibs + l] = bt[l];
            for (int l = lane; l < F; l += TEAM) m.ibias_freq[static_cast<size_t>(i) * F + l] = bq[l];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// Epoch-end caches.  Per element the sums run over the same rows in the same
// order as the reference (train rows, then held-out rows), so the result is
// bit-identical.
// ---------------------------------------------------------------------------

template <int TEAM, int VEC>
__global__ __launch_bounds__(256) void k_rebuild_ycache(DevModel m) {
#pragma clang fp contract(off)
    const int team = (blockIdx.x * blockDim.x + threadIdx.x) / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int nteams = gridDim.x * blockDim.x / TEAM;
    const int fs = m.fs, off = lane * VEC;
    for (int u = team; u < m.U; u += nteams) {
    // rows wider than the team (f > 1024) are summed a TEAM*VEC chunk at a time
    for (int cb = 0; cb < fs; cb += TEAM * VEC) {
        const bool yb_on = !m.skip_yb && cb == 0;
        float acc[VEC];
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
        float sb = 0.0f;
        const uint32_t ts = m.t_off[u], ntr = m.t_off[u + 1] - ts;
        const uint32_t n = ntr + (m.h_start[u + 1] - m.h_start[u]);
        const uint32_t* titems = m.t_item + ts;
        const uint32_t* hitems = m.h_item + m.h_start[u] - ntr;  // row r >= ntr is hitems[r]
        // rows streamed in order (training rows, then held-out rows); every
        // per-element sum and the ybias sum run sequentially in that order
        for_rows<TEAM, VEC, row_group<TEAM, VEC>()>(
            n, lane, m.yfeat + cb, fs, [&](uint32_t r) { return r < ntr ? titems + r : hitems + r; },
            [&](uint32_t b0, uint32_t ib) {
                if (!yb_on) return;
                const float ybl = b0 + static_cast<uint32_t>(lane) < n ? m.ybias[ib] : 0.0f;
                const uint32_t nb = n - b0 < static_cast<uint32_t>(TEAM) ? n - b0 : static_cast<uint32_t>(TEAM);
                for (uint32_t k = 0; k < nb; ++k) sb += team_bcast_f<TEAM>(ybl, k);
            },
            [&](uint32_t, uint32_t, const float (&y)[VEC]) {
#pragma unroll
                for (int c = 0; c < VEC; ++c) acc[c] += y[c];
            });
        const float nrm = m.norm_rebuild[u];
        if (nrm > 0.0f) {
#pragma unroll
            for (int c = 0; c < VEC; ++c) acc[c] = acc[c] / nrm;
            sb = sb / nrm;
        }
        store_vec<VEC>(m.ycache + static_cast<size_t>(u) * fs + cb + off, acc);
        if (lane == 0 && cb == 0) m.ycache_bias[u] = sb;
    }
    }
}

template <int TEAM, int VEC>
__global__ __launch_bounds__(256) void k_rebuild_ycache_day(DevModel m, int64_t n_ud) {
#pragma clang fp contract(off)
    const int64_t team = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int64_t nteams = static_cast<int64_t>(gridDim.x) * blockDim.x / TEAM;
    const int fs = m.fs;
    for (int64_t d = team; d < n_ud; d += nteams) {
    for (int off = lane * VEC; off < fs; off += TEAM * VEC) {  // chunks of rows wider than the team
        float acc[VEC];
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
        for (uint32_t q = m.hd_start[d]; q < m.hd_start[d + 1]; ++q) {
            float y[VEC];
            load_vec<VEC>(m.yfeat_day + static_cast<size_t>(m.hd_item[q]) * fs + off, y);
#pragma unroll
            for (int c = 0; c < VEC; ++c) acc[c] += y[c];
        }
        for (uint32_t q = m.td_start[d]; q < m.td_start[d + 1]; ++q) {
            float y[VEC];
            load_vec<VEC>(m.yfeat_day + static_cast<size_t>(m.td_item[q]) * fs + off, y);
#pragma unroll
            for (int c = 0; c < VEC; ++c) acc[c] += y[c];
        }
        const float nrm = m.sud_norm[d];
        if (nrm > 0.0f) {
#pragma unroll
            for (int c = 0; c < VEC; ++c) acc[c] = acc[c] / nrm;
        }
        store_vec<VEC>(m.ycache_day + static_cast<size_t>(d) * fs + off, acc);
    }
    }
}

// ---------------------------------------------------------------------------
// Predict: one team per (user, item, day) query.
//   MODE 0: score (ordinal expectation when the head is on)
//   MODE 1: subscores bias / mf / nsvd1
//   MODE 2: squared error vs. target, FP64 sum into out_d[0], count in out_d[1]
// ---------------------------------------------------------------------------

template <int TEAM, int VEC, int MODE>
__global__ __launch_bounds__(256) void k_predict(DevModel m, int64_t n, const int32_t* __restrict__ us,
                                                 const int32_t* __restrict__ is, const int16_t* __restrict__ ds,
                                                 float* __restrict__ out0, float* __restrict__ out1,
                                                 float* __restrict__ out2, const float* __restrict__ target,
                                                 const int8_t* __restrict__ is_test, double* __restrict__ out_d) {
    const int64_t team = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int64_t nteams = static_cast<int64_t>(gridDim.x) * blockDim.x / TEAM;
    const int fs = m.fs, F = m.F, B = m.B, off = lane * VEC;
    double sse = 0.0, cnt = 0.0;
    for (int64_t qn = team; qn < n; qn += nteams) {
        if (MODE == 2 && is_test != nullptr && is_test[qn] != 0) continue;
        const int u = us[qn], i = is[qn];
        const int day = ds[qn];
        const int b = time_bin_of(day, B, m.day_range);
        const float dev = dev_of(m.tu_mean[u], day, m.h.beta);
        const int64_t ud = sud_find(m, u, static_cast<int16_t>(day));
        const int fb = ud >= 0 ? m.sud_fbin[ud] : 0;
        const float bu_day = ud >= 0 ? m.ubias_day[ud] : 0.0f;
        const float bi_t = m.ibias[static_cast<size_t>(i) * m.ibs] + m.bit_bin[static_cast<size_t>(i) * m.ibs + b] + m.ibias_freq[static_cast<size_t>(i) * F + fb];
        const float cu_t = m.cu[u] + (ud >= 0 ? m.cut[ud] : 0.0f);
        const bool day_terms = !m.low_mem && ud >= 0;
        // partial sums over the row, a TEAM*VEC chunk at a time (one chunk up to f=1024)
        float part = 0.0f, mf = 0.0f, ns = 0.0f;
        for (int cb = 0; cb < fs; cb += TEAM * VEC) {
        const int o = cb + off;
        float q[VEC], qf[VEC], p[VEC], p2[VEC], s[VEC];
        if (m.bf16_pred) {
            load_bf16<VEC>(m.b_ifeat + static_cast<size_t>(i) * fs + o, q);
            load_bf16<VEC>(m.b_ifeat_freq + (static_cast<size_t>(i) * F + fb) * fs + o, qf);
            load_bf16<VEC>(m.b_ufeat + static_cast<size_t>(u) * fs + o, p);
            load_bf16<VEC>(m.b_ufeat2 + static_cast<size_t>(u) * fs + o, p2);
            load_bf16<VEC>(m.b_ycache + static_cast<size_t>(u) * fs + o, s);
        } else {
            load_vec<VEC>(m.ifeat + static_cast<size_t>(i) * fs + o, q);
            load_vec<VEC>(m.ifeat_freq + (static_cast<size_t>(i) * F + fb) * fs + o, qf);
            load_vec<VEC>(m.ufeat + static_cast<size_t>(u) * fs + o, p);
            load_vec<VEC>(m.ufeat2 + static_cast<size_t>(u) * fs + o, p2);
            load_vec<VEC>(m.ycache + static_cast<size_t>(u) * fs + o, s);
        }
        float pd[VEC], sd[VEC];
        if (day_terms) {
            load_vec<VEC>(m.ufeat_day + static_cast<size_t>(ud) * fs + o, pd);
            load_vec<VEC>(m.ycache_day + static_cast<size_t>(ud) * fs + o, sd);
        } else {
#pragma unroll
            for (int c = 0; c < VEC; ++c) { pd[c] = 0.0f; sd[c] = 0.0f; }
        }
        if (MODE == 1) {
#pragma unroll
            for (int c = 0; c < VEC; ++c) {
                const float qe = q[c] + qf[c];
                mf += (p[c] + dev * p2[c] + pd[c]) * qe;
                ns += (s[c] + sd[c]) * qe;
            }
        } else {
#pragma unroll
            for (int c = 0; c < VEC; ++c) part += (q[c] + qf[c]) * (p[c] + s[c] + dev * p2[c] + pd[c] + sd[c]);
        }
        }
        if (MODE == 1) {
            mf = team_sum<TEAM>(mf);
            ns = team_sum<TEAM>(ns);
            if (lane == 0) {
                const float bias = m.gbias + m.ubias[u] + m.ibias[static_cast<size_t>(i) * m.ibs] + m.ibias_freq[static_cast<size_t>(i) * F + fb]
                                   + m.but_bin[static_cast<size_t>(u) * B + b] + m.alpha_u[u] * dev + bu_day
                                   + m.bit_bin[static_cast<size_t>(i) * m.ibs + b] + bi_t * (cu_t - 1.0f) + m.ycache_bias[u];
                out0[qn] = bias;
                out1[qn] = mf;
                out2[qn] = ns;
            }
            continue;
        }
        const float dot = team_sum<TEAM>(part);
        const float bu_t = m.ubias[u] + m.but_bin[static_cast<size_t>(u) * B + b] + m.alpha_u[u] * dev + bu_day
                           + m.ycache_bias[u];
        float score = m.gbias + bu_t + bi_t * cu_t + dot;
        if (m.has_ord) {
            const float th[4] = {m.thresholds[0], m.thresholds[1], m.thresholds[2], m.thresholds[3]};
            score = ordinal_expect(th, score);
        }
        if (MODE == 0) {
            if (lane == 0) out0[qn] = score;
        } else {
            const double e = static_cast<double>(score - target[qn]);
            sse += e * e;
            cnt += 1.0;
        }
    }
    if (MODE == 2) {
        // lane 0 of every team holds the same sums as its team-mates; reduce
        // over the block's teams, then one FP64 atomic per block.
        __shared__ double red[2][256];
        red[0][threadIdx.x] = lane == 0 ? sse : 0.0;
        red[1][threadIdx.x] = lane == 0 ? cnt : 0.0;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if (static_cast<int>(threadIdx.x) < o) {
                red[0][threadIdx.x] += red[0][threadIdx.x + o];
                red[1][threadIdx.x] += red[1][threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            atomicAdd(&out_d[0], red[0][0]);
            atomicAdd(&out_d[1], red[1][0]);
        }
    }
}

// ---------------------------------------------------------------------------
// fills (used instead of memsets so a captured epoch holds only kernels)
// ---------------------------------------------------------------------------

__global__ void k_fill_f32(float* p, float v, int64_t n) {
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[k] = v;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[k] = v;
}

__global__ void k_fill_f64(double* p, double v, int64_t n) {
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[k] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline int grid_for(int64_t teams, int team) {
    const int64_t threads = teams * team;
    int64_t blocks = (threads + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    return static_cast<int>(blocks < 1 ? 1 : blocks);
}

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

hipError_t launch_sgd_serial(const DevModel& m, const int64_t* order, float* scratch, hipStream_t st) {
    const size_t shm = sizeof(float) * (static_cast<size_t>(m.fs) + 1);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_sgd_serial, dim3(1), dim3(64), shm, st, m, order, scratch);
    return hipGetLastError();
}

template <int TEAM, int VEC, int PASS, bool DAY>
static hipError_t sgd_par_tvp(const DevModel& m, const StrataView& sv, int stratum, int blocks, bool qf, bool p2,
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

template <int TEAM, int VEC>
static hipError_t sgd_par_tv(const DevModel& m, const StrataView& sv, int stratum, int blocks, bool qf, bool p2,
                             bool ord, int pass, bool day, hipStream_t st) {
    if (day) {
        if (pass != 0 || sv.n_strata != 1 || !sv.day_acc) return hipErrorInvalidValue;
        return sgd_par_tvp<TEAM, VEC, 0, true>(m, sv, stratum, blocks, qf, p2, ord, st);
    }
    if (pass == 0) return sgd_par_tvp<TEAM, VEC, 0, false>(m, sv, stratum, blocks, qf, p2, ord, st);
    if (pass == 1) return sgd_par_tvp<TEAM, VEC, 1, false>(m, sv, stratum, blocks, qf, p2, ord, st);
    if (pass == 2) return sgd_par_tvp<TEAM, VEC, 2, false>(m, sv, stratum, blocks, qf, p2, ord, st);
    return hipErrorInvalidValue;
}

hipError_t launch_sgd_plain(const DevModel& m, const int64_t* order, int vec, int blocks, bool qf, hipStream_t st) {
    (void)hipGetLastError();
    switch (vec) {
        case 1: hipLaunchKernelGGL(k_sgd_plain<1>, dim3(blocks), dim3(64), 0, st, m, order, qf ? 1 : 0); break;
        case 2: hipLaunchKernelGGL(k_sgd_plain<2>, dim3(blocks), dim3(64), 0, st, m, order, qf ? 1 : 0); break;
        case 4: hipLaunchKernelGGL(k_sgd_plain<4>, dim3(blocks), dim3(64), 0, st, m, order, qf ? 1 : 0); break;
        case 8: hipLaunchKernelGGL(k_sgd_plain<8>, dim3(blocks), dim3(64), 0, st, m, order, qf ? 1 : 0); break;
        case 16: hipLaunchKernelGGL(k_sgd_plain<16>, dim3(blocks), dim3(64), 0, st, m, order, qf ? 1 : 0); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sgd_par(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                          bool qf, bool p2, bool ord, int pass, bool day, hipStream_t st) {
#define X(T, V) \
    if (team == T && vec == V) return sgd_par_tv<T, V>(m, sv, stratum, blocks, qf, p2, ord, pass, day, st);
    TX_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

template <bool DAY, int MW, int PASS = 0>
static hipError_t sgd_par_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool qf, bool p2, bool ord,
                             hipStream_t st) {
    (void)hipGetLastError();
#define TX_MW(QF, P2, ORD)                                                                                       \
    hipLaunchKernelGGL((k_sgd_par<64, 16, QF, P2, ORD, PASS, DAY, MW>), dim3(teams), dim3(64 * waves), 0, st, m, \
                       sv, 0)
    if constexpr (PASS != 0) {
        if (ord || DAY) return hipErrorInvalidValue;
        if (qf) { if (p2) TX_MW(true, true, false); else TX_MW(true, false, false); }
        else { if (p2) TX_MW(false, true, false); else TX_MW(false, false, false); }
        return hipGetLastError();
    } else if (ord) {
        if (qf) { if (p2) TX_MW(true, true, true); else TX_MW(true, false, true); }
        else { if (p2) TX_MW(false, true, true); else TX_MW(false, false, true); }
    } else {
        if (qf) { if (p2) TX_MW(true, true, false); else TX_MW(true, false, false); }
        else { if (p2) TX_MW(false, true, false); else TX_MW(false, false, false); }
    }
#undef TX_MW
    return hipGetLastError();
}

hipError_t launch_sgd_par_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool qf, bool p2, bool ord,
                             bool day, hipStream_t st, int pass) {
    if (waves < 2 || waves > 8 || m.fs != 1024 * waves || sv.n_strata != 1 || (day && !sv.day_acc))
        return hipErrorInvalidValue;
    if (pass == 1 || pass == 2) {
        if (day || ord) return hipErrorInvalidValue;
        if (waves <= 4)
            return pass == 1 ? sgd_par_mw<false, 4, 1>(m, sv, teams, waves, qf, p2, false, st)
                             : sgd_par_mw<false, 4, 2>(m, sv, teams, waves, qf, p2, false, st);
        return pass == 1 ? sgd_par_mw<false, 8, 1>(m, sv, teams, waves, qf, p2, false, st)
                         : sgd_par_mw<false, 8, 2>(m, sv, teams, waves, qf, p2, false, st);
    }
    if (pass != 0) return hipErrorInvalidValue;
    if (waves <= 4)
        return day ? sgd_par_mw<true, 4>(m, sv, teams, waves, qf, p2, ord, st)
                   : sgd_par_mw<false, 4>(m, sv, teams, waves, qf, p2, ord, st);
    return day ? sgd_par_mw<true, 8>(m, sv, teams, waves, qf, p2, ord, st)
               : sgd_par_mw<false, 8>(m, sv, teams, waves, qf, p2, ord, st);
}

hipError_t launch_nsvd_gather(const DevModel& m, const ItemView& iv, const float* gacc, const float* gacc_b,
                              const float* scale, int vec, hipStream_t st) {
    if (m.I <= 0) return hipSuccess;
    const int blocks = std::min(2048, (m.I + 3) / 4);
    (void)hipGetLastError();
    switch (vec) {
        case 1: hipLaunchKernelGGL(k_nsvd_gather<1>, dim3(blocks), dim3(256), 0, st, m, iv, gacc, gacc_b, scale); break;
        case 2: hipLaunchKernelGGL(k_nsvd_gather<2>, dim3(blocks), dim3(256), 0, st, m, iv, gacc, gacc_b, scale); break;
        case 4: hipLaunchKernelGGL(k_nsvd_gather<4>, dim3(blocks), dim3(256), 0, st, m, iv, gacc, gacc_b, scale); break;
        case 8: hipLaunchKernelGGL(k_nsvd_gather<8>, dim3(blocks), dim3(256), 0, st, m, iv, gacc, gacc_b, scale); break;
        case 16: hipLaunchKernelGGL(k_nsvd_gather<16>, dim3(blocks), dim3(256), 0, st, m, iv, gacc, gacc_b, scale); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_item_pass(const DevModel& m, const ItemView& iv, int team, int vec, bool p2, bool qf,
                            int blocks, hipStream_t st, int waves) {
    if (waves > 1) {
        // multi-wave teams (f > 1024): one item per workgroup of 64*waves threads
        if (team != 64 || vec != 16 || waves > 8 || m.fs != 1024 * waves) return hipErrorInvalidValue;
        const size_t shm = sizeof(float) * static_cast<size_t>(m.B + m.F) * waves;
        (void)hipGetLastError();
#define TX_IPW(P2, QF, MW) \
    hipLaunchKernelGGL((k_item_pass<64, 16, P2, QF, MW>), dim3(blocks), dim3(64 * waves), shm, st, m, iv)
        if (waves <= 4) {
            if (qf) { if (p2) TX_IPW(true, true, 4); else TX_IPW(false, true, 4); }
            else { if (p2) TX_IPW(true, false, 4); else TX_IPW(false, false, 4); }
        } else {
            if (qf) { if (p2) TX_IPW(true, true, 8); else TX_IPW(false, true, 8); }
            else { if (p2) TX_IPW(true, false, 8); else TX_IPW(false, false, 8); }
        }
#undef TX_IPW
        return hipGetLastError();
    }
    const size_t shm = sizeof(float) * static_cast<size_t>(m.B + m.F) * (256 / team);
#define TX_IP(T, V, P2, QF) \
    hipLaunchKernelGGL((k_item_pass<T, V, P2, QF, 0>), dim3(blocks), dim3(256), shm, st, m, iv)
#define X(T, V)                                                                     \
    if (team == T && vec == V) {                                                     \
        (void)hipGetLastError();                                                     \
        if (qf) { if (p2) TX_IP(T, V, true, true); else TX_IP(T, V, false, true); }   \
        else { if (p2) TX_IP(T, V, true, false); else TX_IP(T, V, false, false); }    \
        return hipGetLastError();                                                    \
    }
    TX_SHAPES(X)
#undef X
#undef TX_IP
    return hipErrorInvalidValue;
}

hipError_t launch_rebuild(const DevModel& m, int64_t n_ud, int team, int vec, hipStream_t st) {
    bool done = false;
#define X(T, V)                                                                                   \
    if (!done && team == T && vec == V) {                                                          \
        (void)hipGetLastError();                                                                   \
        hipLaunchKernelGGL((k_rebuild_ycache<T, V>), dim3(grid_for(m.U, T)), dim3(256), 0, st, m); \
        if (hipError_t e = hipGetLastError()) return e;                                            \
        done = true;                                                                               \
    }
    TX_SHAPES(X)
#undef X
    if (!done) return hipErrorInvalidValue;
    if (m.low_mem || n_ud <= 0) return hipSuccess;
    // per-day sums: four elements per lane, up to 64 lanes (wider rows in
    // chunks of 64 x 16)
    const int dt = m.fs / 4 <= 64 ? m.fs / 4 : 64, dv = m.fs / dt <= 16 ? m.fs / dt : 16;
#define X(T, V)                                                                                                 \
    if (dt == T && dv == V) {                                                                                    \
        (void)hipGetLastError();                                                                                 \
        hipLaunchKernelGGL((k_rebuild_ycache_day<T, V>), dim3(grid_for(n_ud, T)), dim3(256), 0, st, m, n_ud);    \
        return hipGetLastError();                                                                                \
    }
    TX_DAY_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_predict(const DevModel& m, int mode, int64_t n, const int32_t* us, const int32_t* is,
                          const int16_t* ds, float* o0, float* o1, float* o2, const float* target,
                          const int8_t* is_test, double* out_d, int team, int vec, hipStream_t st) {
    if (n <= 0) return hipSuccess;
#define X(T, V)                                                                                            \
    if (team == T && vec == V) {                                                                            \
        const int g = grid_for(n, T);                                                                       \
        (void)hipGetLastError();                                                                            \
        if (mode == 0)                                                                                      \
            hipLaunchKernelGGL((k_predict<T, V, 0>), dim3(g), dim3(256), 0, st, m, n, us, is, ds, o0, o1, o2, \
                               target, is_test, out_d);                                                     \
        else if (mode == 1)                                                                                 \
            hipLaunchKernelGGL((k_predict<T, V, 1>), dim3(g), dim3(256), 0, st, m, n, us, is, ds, o0, o1, o2, \
                               target, is_test, out_d);                                                     \
        else                                                                                                \
            hipLaunchKernelGGL((k_predict<T, V, 2>), dim3(g), dim3(256), 0, st, m, n, us, is, ds, o0, o1, o2, \
                               target, is_test, out_d);                                                     \
        return hipGetLastError();                                                                           \
    }
    TX_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_to_bf16, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 1) > 4096 ? 4096 : grid_for(n, 1)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 1) > 4096 ? 4096 : grid_for(n, 1)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// tx_predict_noalloc's scratch: a file-scope device array, so that entry point
// never allocates.  Every device that runs the library holds its own copy
// until the process ends; all calls on one device share it.
static __device__ __align__(256) char scratch[kNoAllocBytes];

hipError_t noalloc_scratch(int device, void** out) {
    static std::mutex mu;
    static void* ptr[64] = {};
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptr[device] = p;
    }
    *out = ptr[device];
    return hipSuccess;
}

hipError_t launch_fill_f64(double* p, double v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, st, p, v, n);
    return hipGetLastError();
}

}  // namespace tx
