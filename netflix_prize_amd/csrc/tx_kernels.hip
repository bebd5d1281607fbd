// HIP kernels of the MI355X tsvdx4 (timeSVD++) path, gfx950.
//
//   k_sgd_serial     one wavefront replays TxModel::fit_epoch
//                    (src/tx.rs:734-1115) user by user in the reference's
//                    order with the reference's float operation order.
//   k_sgd_plain      one wave per user, no prefetch or queue.
//   k_sgd_par        all users in parallel (template in tx_device.hpp,
//                    instantiated in tx_sgd_lm/day/user/mw.hip): one team of
//                    TEAM lanes per user, VEC factor elements per lane, user
//                    state in registers, item updates as device-scope atomic
//                    deltas; run once per stratum of the rotating user-block
//                    x item-block schedule.
//   k_nsvd_gather, k_item_pass  the tsvdx4p schedule's NSVD1 apply and item
//                    pass (src/tsvdx4p.rs:341-385, 644-762).
//   k_rebuild_ycache[_day]  epoch-end NSVD1 caches (src/tx.rs:391-454).
//   k_predict        predict / predict_subscores / calc_rmse
//                    (src/tx.rs:303-339, 663-724; src/lib.rs:659-673).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "tx_launch.hpp"
#include "tx_types.hpp"
#include "tx_device.hpp"

namespace tx {

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        y[e] = to_bf16(x[e]);
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
        // the epoch's global neighbours of u (src/tx.rs:762-796): rb sampled
        // R(u) positions then ib sampled N(u)\R(u) positions, or every one
        uint32_t nb_rb = 0, nb_ib = 0;
        float nb_rs = 1.0f, nb_is = 1.0f, nb_nr = 0.0f, nb_nn = 0.0f;
        bool nb_smp = false;
        const uint32_t imp0 = m.nb_on ? m.imp_start[u] : 0u;
        if (m.nb_on) {
            const uint32_t nr_full = end - start, ni_full = m.imp_start[u + 1] - imp0;
            const uint32_t cw = m.smp_cnt[u];
            nb_smp = cw != 0xffffffffu;
            nb_rb = nb_smp ? (cw & 0xffffu) : nr_full;
            nb_ib = nb_smp ? (cw >> 16) : ni_full;
            if (nb_smp) {
                nb_rs = static_cast<float>(nr_full) / static_cast<float>(nb_rb);
                nb_is = nb_ib > 0 ? static_cast<float>(ni_full) / static_cast<float>(nb_ib) : 1.0f;
            }
            nb_nr = m.norm_ru[u];
            nb_nn = m.norm_nu[u];
        }
        const uint16_t* nb_pos = m.nb_on && nb_smp ? m.smp + static_cast<size_t>(u) * m.nb_max : nullptr;
        // k-th global neighbour: item, r - b~ (rated) and whether it is rated
        auto nb_at = [&](uint32_t k, uint32_t& j, float& rmb) {
            if (k < nb_rb) {
                const uint32_t t2 = start + (nb_smp ? nb_pos[k] : k);
                j = m.t_item[t2];
                rmb = m.t_rmb[t2];
                return true;
            }
            j = m.imp_item[imp0 + (nb_smp ? nb_pos[k] : k - nb_rb)];
            rmb = 0.0f;
            return false;
        };

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
            // neighbourhood terms (src/tx.rs:903-956), each sum in list order;
            // every lane evaluates them from the same loads
            float nbr_part = 0.0f, nn_day = 0.0f, nr_day = 0.0f;
            const size_t irow = static_cast<size_t>(i) * m.I;
            if (m.nb_on) {
                float c_day_sum = 0.0f, w_day_sum = 0.0f;
                if (m.nbd_on) {
                    uint32_t n_same_rated = 0, n_same_total = 0;
                    for (uint32_t e = m.sd_start[ud]; e < m.sd_start[ud + 1]; ++e) {
                        const uint32_t j = m.sd_item[e] & 0x7fffffffu;
                        const bool rated = (m.sd_item[e] >> 31) != 0u;
                        if (j == i) continue;
                        n_same_rated += rated ? 1u : 0u;
                        n_same_total += 1u;
                        c_day_sum += m.nc_day[irow + j];
                        if (rated) w_day_sum += m.sd_rmb[e] * m.nw_day[irow + j];
                    }
                    nn_day = m.inv_sqrt_n[n_same_total];
                    nr_day = m.inv_sqrt_n[n_same_rated];
                }
                float c_sum = 0.0f, w_sum = 0.0f;
                for (uint32_t k = 0; k < nb_rb + nb_ib; ++k) {
                    uint32_t j;
                    float rmb;
                    const bool rated = nb_at(k, j, rmb);
                    if (j == i) continue;
                    if (rated) {
                        c_sum += m.nc[irow + j] * nb_rs;
                        w_sum += rmb * m.nw[irow + j] * nb_rs;
                    } else {
                        c_sum += m.nc[irow + j] * nb_is;
                    }
                }
                nbr_part = nb_nn * c_sum + nb_nr * w_sum + nn_day * c_day_sum + nr_day * w_day_sum;
            }
            const float score = bias_part + dot + nbr_part;

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
            const float err_n = err * h.w_nbr;

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

            // neighbourhood steps in list order (src/tx.rs:1038-1095), lane 0
            if (m.nb_on && lane == 0) {
                for (uint32_t k = 0; k < nb_rb + nb_ib; ++k) {
                    uint32_t j;
                    float rmb;
                    const bool rated = nb_at(k, j, rmb);
                    if (j == i) continue;
                    const size_t idx = irow + j;
                    if (rated) {
                        const float w_ij = m.nw[idx];
                        m.nw[idx] = w_ij - h.lr_w * (err_n * nb_nr * rmb * nb_rs + h.reg_w * w_ij);
                        const float c_ij = m.nc[idx];
                        m.nc[idx] = c_ij - h.lr_c * (err_n * nb_nn * nb_rs + h.reg_c * c_ij);
                    } else {
                        const float c_ij = m.nc[idx];
                        m.nc[idx] = c_ij - h.lr_c * (err_n * nb_nn * nb_is + h.reg_c * c_ij);
                    }
                }
                if (m.nbd_on) {
                    for (uint32_t e = m.sd_start[ud]; e < m.sd_start[ud + 1]; ++e) {
                        const uint32_t j = m.sd_item[e] & 0x7fffffffu;
                        if (j == i) continue;
                        const size_t idx = irow + j;
                        if ((m.sd_item[e] >> 31) != 0u) {
                            const float wd = m.nw_day[idx];
                            m.nw_day[idx] = wd - h.lr_w_day * (err_n * nr_day * m.sd_rmb[e] + h.reg_w_day * wd);
                        }
                        const float cd = m.nc_day[idx];
                        m.nc_day[idx] = cd - h.lr_c_day * (err_n * nn_day + h.reg_c_day * cd);
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

// nbr_score (src/tx.rs:341-389) of one query, the team's lanes striding over
// the user's lists: every R(u) and N(u)\R(u) neighbour unscaled (no sampling
// at prediction time), and the same-day neighbours from the list of the
// query's SparseUD day (ud < 0: the user has no rating that day, no terms).
// Every lane returns the team's sum.
template <int TEAM>
__device__ float nbr_score_team(const DevModel& m, int u, int i, int64_t ud, int lane) {
    const size_t irow = static_cast<size_t>(i) * m.I;
    const uint32_t ui = static_cast<uint32_t>(i);
    float cs = 0.0f, ws = 0.0f, cd = 0.0f, wd = 0.0f, nsr = 0.0f, nst = 0.0f;
    for (uint32_t t = m.t_off[u] + lane; t < m.t_off[u + 1]; t += TEAM) {
        const uint32_t j = m.t_item[t];
        if (j == ui) continue;
        cs += m.nc[irow + j];
        ws += m.t_rmb[t] * m.nw[irow + j];
    }
    for (uint32_t k = m.imp_start[u] + lane; k < m.imp_start[u + 1]; k += TEAM) {
        const uint32_t j = m.imp_item[k];
        if (j != ui) cs += m.nc[irow + j];
    }
    if (m.nbd_on && ud >= 0) {
        for (uint32_t e = m.sd_start[ud] + lane; e < m.sd_start[ud + 1]; e += TEAM) {
            const uint32_t it = m.sd_item[e], j = it & 0x7fffffffu;
            if (j == ui) continue;
            nst += 1.0f;
            cd += m.nc_day[irow + j];
            if (it >> 31) {
                nsr += 1.0f;
                wd += m.sd_rmb[e] * m.nw_day[irow + j];
            }
        }
    }
    cs = team_sum<TEAM>(cs);
    ws = team_sum<TEAM>(ws);
    cd = team_sum<TEAM>(cd);
    wd = team_sum<TEAM>(wd);
    const int n_tot = static_cast<int>(team_sum<TEAM>(nst)), n_rat = static_cast<int>(team_sum<TEAM>(nsr));
    const float nn_day = m.inv_sqrt_n[n_tot < m.n_isq ? n_tot : m.n_isq - 1];
    const float nr_day = m.inv_sqrt_n[n_rat < m.n_isq ? n_rat : m.n_isq - 1];
    return m.norm_nu[u] * cs + m.norm_ru[u] * ws + nn_day * cd + nr_day * wd;
}

template <int TEAM, int VEC, int MODE>
__global__ __launch_bounds__(256) void k_predict(DevModel m, int64_t n, const int32_t* __restrict__ us,
                                                 const int32_t* __restrict__ is, const int16_t* __restrict__ ds,
                                                 float* __restrict__ out0, float* __restrict__ out1,
                                                 float* __restrict__ out2, float* __restrict__ out3,
                                                 const float* __restrict__ target,
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
        const float nbr = m.nb_on ? nbr_score_team<TEAM>(m, u, i, ud, lane) : 0.0f;
        if (MODE == 1) {
            mf = team_sum<TEAM>(mf);
            ns = team_sum<TEAM>(ns);
            if (lane == 0 && out3) out3[qn] = nbr;
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
        score = score + nbr;
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

hipError_t launch_sgd_serial(const DevModel& m, const int64_t* order, float* scratch, hipStream_t st) {
    const size_t shm = sizeof(float) * (static_cast<size_t>(m.fs) + 1);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_sgd_serial, dim3(1), dim3(64), shm, st, m, order, scratch);
    return hipGetLastError();
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
    if (m.nb_on && pass == 0) {
        if (sv.n_strata != 1 || (day && !sv.day_acc) || m.nb_max < 1 || m.nb_max > 128) return hipErrorInvalidValue;
        return m.nbd_on ? launch_sgd_par_nb2(m, sv, stratum, team, vec, blocks, ord, day, st)
                        : launch_sgd_par_nb1(m, sv, stratum, team, vec, blocks, ord, day, st);
    }
    if (day) {
        if (pass != 0 || sv.n_strata != 1 || !sv.day_acc) return hipErrorInvalidValue;
        return launch_sgd_par_day(m, sv, stratum, team, vec, blocks, qf, p2, ord, st);
    }
    if (pass == 0) return launch_sgd_par_lm(m, sv, stratum, team, vec, blocks, qf, p2, ord, st);
    if (pass == 1 || pass == 2) {
        if (ord) return hipErrorInvalidValue;
        return launch_sgd_par_user(m, sv, stratum, team, vec, blocks, qf, p2, pass, st);
    }
    return hipErrorInvalidValue;
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
                          const int16_t* ds, float* o0, float* o1, float* o2, float* o3, const float* target,
                          const int8_t* is_test, double* out_d, int team, int vec, hipStream_t st) {
    if (n <= 0) return hipSuccess;
#define X(T, V)                                                                                            \
    if (team == T && vec == V) {                                                                            \
        const int g = grid_for(n, T);                                                                       \
        (void)hipGetLastError();                                                                            \
        if (mode == 0)                                                                                      \
            hipLaunchKernelGGL((k_predict<T, V, 0>), dim3(g), dim3(256), 0, st, m, n, us, is, ds, o0, o1, o2, \
                               o3, target, is_test, out_d);                                                     \
        else if (mode == 1)                                                                                 \
            hipLaunchKernelGGL((k_predict<T, V, 1>), dim3(g), dim3(256), 0, st, m, n, us, is, ds, o0, o1, o2, \
                               o3, target, is_test, out_d);                                                     \
        else                                                                                                \
            hipLaunchKernelGGL((k_predict<T, V, 2>), dim3(g), dim3(256), 0, st, m, n, us, is, ds, o0, o1, o2, \
                               o3, target, is_test, out_d);                                                     \
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
