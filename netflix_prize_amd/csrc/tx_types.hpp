// Device-side view of one TxModel: parameter tables, the static per-rating
// records derived from the training set, and the hyper-parameters, passed by
// value to every kernel.
#pragma once

#include <cstdint>

namespace tx {

// Hyper-parameters used inside kernels (a subset of TxConfig).
struct Hyper {
    float lr_u, lr_ud, lr_u2, lr_ub, lr_ubd, lr_i, lr_ib, lr_y, lr_yb, lr_yd;
    float lr_tu, lr_ti, lr_ta, lr_ibf, lr_iqf, lr_cu;
    float reg_iqf, reg_cu, reg_u, reg_u2, reg_ud, reg_i, reg_y, reg_yd;
    float w_bias, w_factor;
    float beta;
    float lr_t, reg_t;
    // neighbourhood terms (tx5/tx6 modes)
    float lr_w, lr_c, reg_w, reg_c, lr_w_day, lr_c_day, reg_w_day, reg_c_day, w_nbr;
};

struct DevModel {
    // sizes
    int32_t U, I, f, fs, F, B;  // f = n_feat, fs = padded row stride
    int32_t day_range;
    int32_t low_mem, full_su, has_ord, sum_err_bug;
    // tables that are all zero and never trained: the parallel kernel skips them
    int32_t skip_cut, skip_ibf, skip_but, skip_yb;
    float gbias;
    Hyper h;

    // parameter tables (row stride fs for the factor tables).  b_i and its
    // time bins share one row of ibs floats per item (ibias[i * ibs],
    // bit_bin[i * ibs + b], bit_bin = ibias + 1; ibs a multiple of 32), so the
    // two item-bias updates of a rating touch one 128-byte line.
    float *ubias, *ibias, *ufeat, *ufeat2, *ifeat, *yfeat, *ycache, *yfeat_day;
    float *tu_mean, *alpha_u, *but_bin, *bit_bin, *ybias, *ycache_bias;
    int32_t ibs;
    float *ubias_day, *ufeat_day, *ycache_day, *ibias_freq, *ifeat_freq, *cu, *cut;
    float* thresholds;  // [4]

    // SparseUD (src/tx.rs:21-84): per-user sorted unique days
    const uint32_t* sud_start;  // [U+1]
    const int16_t* sud_date;    // [D]
    const uint8_t* sud_fbin;    // [D] freq_bin(day_cnt)
    const float* sud_norm;      // [D] sqrt(#train+#heldout ratings on that day)

    // training ratings, reference (user-sorted file) order
    const uint32_t* t_off;   // [U+1]
    const uint32_t* t_item;  // [N]
    const uint32_t* t_ud;    // [N] SparseUD index of (user, date)
    const float* t_dev;      // [N] dev_u(t), evaluated once on the host
    const float* t_r;        // [N] residual target
    const uint8_t* t_bin;    // [N] time bin
    const int8_t* t_raw;     // [N] raw rating (ordinal head)
    const uint32_t* t_run_stop;  // [N] last index of the same-day run (day terms)
    const float* t_run_norm;     // [N] sqrt(#ratings of the run + #heldout that day)

    // held-out rows grouped by user / by (user, day), held-out file order
    const uint32_t* h_start;  // [U+1]
    const uint32_t* h_item;
    const uint32_t* hd_start; // [D+1]
    const uint32_t* hd_item;
    // training items grouped by (user, day), training file order
    const uint32_t* td_start; // [D+1]
    const uint32_t* td_item;
    // per-user day streams of the parallel day-term kernel (date-ordered
    // ratings only); entries with bit 31 set are day headers (SparseUD index):
    //   dg: for each training day d: its training items, (full_su) its
    //       held-out items, then header d  -- the day sums
    //   dy: for each training day d: header d, then its training items  -- the
    //       y_day steps
    const uint32_t* dg_off;   // [U+1]
    const uint32_t* dg_item;
    const uint32_t* dy_off;   // [U+1]
    const uint32_t* dy_item;

    const float* norm_train;    // [U] sqrt(|R(u)| + |P(u)| if full_su)
    const float* norm_rebuild;  // [U] sqrt(train_cnt + heldout_cnt)

    // Neighbourhood terms (tx5 / tx6 modes, src/tx.rs:253-262, 483-585).
    // nb_on: w and c built (lr_w > 0 or lr_c > 0), with the per-user lists
    // R(u) (the training ratings, file order, r - b~ in t_rmb) and N(u)\R(u)
    // (imp_*: held-out items not rated in training, sorted by item [, day],
    // deduplicated); nbd_on: w_day and c_day built (their terms need the
    // lists, so they act only when nb_on is set too).  The dense tables are
    // I x I row-major, element (i, j) at i * I + j.
    int32_t nb_on, nbd_on, nb_max;
    float *nw, *nc, *nw_day, *nc_day;
    const float* t_rmb;          // [N]
    const uint32_t* imp_start;   // [U+1]
    const uint32_t* imp_item;    // [n_imp]
    const int16_t* imp_day;      // [n_imp]
    const float* norm_nu;        // [U] |N(u)|^-0.5 (0 for an empty set), host powf
    const float* norm_ru;        // [U] |R(u)|^-0.5
    const float* inv_sqrt_n;     // [n_isq] k^-0.5 by host powf (k = 0: 0), for same-day counts
    int32_t n_isq;
    // same-day lists per SparseUD day (nb_on && nbd_on): the user's training
    // items of that day in file order (bit 31 set), then its N(u)\R(u) items
    // of that day in list order
    const uint32_t* sd_start;    // [D+1]
    const uint32_t* sd_item;
    const float* sd_rmb;         // r - b~ of the rated entries, 0 for the others
    // the epoch's global neighbour samples (src/tx.rs:750-793), drawn on the
    // host from the reference's stream: smp_cnt[u] = rb | ib << 16 with the
    // rb sorted R(u) positions then the ib sorted N(u)\R(u) positions in
    // smp[u * nb_max ..]; 0xffffffff = no sampling (every neighbour, scale 1)
    const uint32_t* smp_cnt;     // [U]
    const uint16_t* smp;         // [U * nb_max]

    // BF16 read copies of the factor tables (row stride fs; FP32 tables stay
    // the masters); predictions read them when bf16_pred is set, and the
    // passes that hold a side frozen (the user passes: q, q_f; the item pass:
    // p, p2, ycache) read that side from them when bf16_frozen is set
    int32_t bf16_pred, bf16_frozen;
    const uint16_t *b_ifeat, *b_ifeat_freq, *b_ufeat, *b_ufeat2, *b_ycache;
};

// Parallel-schedule state (stratified user x item-block rotation).
struct StrataView {
    int32_t n_strata;
    int32_t n_cells;            // user blocks = item blocks = n_strata (1 => one cell)
    const uint32_t* seg_off;    // [U * (n_strata + 1)] offsets into the permuted records
    // packed training records in segment order:
    //   x = item | (raw & 15) << 20 | bin << 24, y = SparseUD index | freq bin << 28,
    //   z = dev, w = target (freq bin 0 for low-memory training)
    const uint4* p_rec;
    const uint32_t* q_users;    // per (stratum, cell) queues of users, longest segment first
    const uint32_t* q_start;    // [n_strata * n_cells + 1]
    uint32_t* q_head;           // [n_strata * n_cells] work counters (reset per epoch)
    float* gacc;                // [U * fs] per-user sum_err_q across strata; user pass: sum_err_q / norm
    float* gacc_b;              // [U]      per-user sum_err across strata; user pass: sum_err / norm
    int32_t q_mode;             // item-row update: 0 device atomic, 1 XCD-local atomic, 2 plain
    int32_t y_mode;             // NSVD1-row update, same codes
    float* day_acc;             // [D * fs] day terms: a day's y_day sum (user start), then its gradient sum
    // user passes: the NSVD1 sum of a user is ycache[u] (current, and built
    // over the same rows: full_su), not gathered again from y
    int32_t su_cached;
    int32_t nb_mode;            // neighbourhood-table update: 0 device atomic, 2 plain
    int32_t block_threads;      // SGD workgroup size of the one-wave teams: 64 x teams per block (0 = 256)
    // length buckets in one launch (one stratum, one-wave teams, one team per
    // workgroup): workgroups [bq_first[k], bq_first[k + 1]) serve queue k of
    // q_start / q_head, k < n_bq = 4; bq_first is nondecreasing (an empty
    // bucket has an empty range).  n_bq = 0: every workgroup serves the
    // (stratum, cell) queue.
    int32_t n_bq;
    int32_t bq_first[5];
};

// Item pass of the tsvdx4p schedule: ratings grouped by item.
struct ItemView {
    const uint32_t* q_items;   // items with ratings, most ratings first
    uint32_t n_items;          // length of q_items
    uint32_t* q_head;          // work counter (reset per pass)
    const uint32_t* bi_start;  // [I+1] offsets into bi_rec
    // per rating, training-file order within an item (ByItemIndex, src/tsvdx4p.rs:29-57):
    //   x = user | bin << 20 | freq_bin << 28, y = SparseUD index, z = dev, w = target
    const uint4* bi_rec;
    const uint32_t* bi_user;   // [N] the user of each rating, same order
};

}  // namespace tx
