/*
 * tsvd_tx.h — C ABI of the MI355X timeSVD++ (tsvdx4 / TxModel) trainer and
 * predictor.
 *
 * The library libtsvd_tx.so replaces the reference's `impl Regressor for
 * TxModel` (src/tx.rs:457-1116) behind the `Regressor` trait
 * (src/lib.rs:616-651).  One tx_ctx is one model: it owns every parameter
 * table in GPU memory.  All host buffers belong to the caller; the library
 * copies what it needs during the call and never keeps a host pointer.
 *
 * Conventions:
 *   - every entry point returns 0 on success and a negative tx_status on
 *     failure; tx_last_error() then describes the failure (per thread);
 *   - tx_predict, tx_predict_subscores, tx_calc_rmse, tx_param_len and
 *     tx_get_param may run concurrently from several host threads on one
 *     tx_ctx (each call stages through device memory of its own and is
 *     ordered after the training work queued before it); every other entry
 *     point needs the ctx to itself, with no call in flight on it;
 *   - integer/byte inputs use the reference's on-disk dtypes (user i32,
 *     item i32 after widening, date i16, rating i8, is_test i8).
 */
#ifndef TSVD_TX_H
#define TSVD_TX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tx_ctx tx_ctx;

enum tx_status {
    TX_OK = 0,
    TX_EINVAL = -1,     /* bad argument or configuration */
    TX_EDEVICE = -2,    /* HIP runtime / kernel launch failure */
    TX_ENOMEM = -3,     /* allocation failed */
    TX_EUNSUPPORTED = -4
};

/* Mirror of TxConfig (src/tx.rs:128-204) plus OrdinalHeadConfig
 * (src/lib.rs:502-507).  Booleans are int32 0/1.  The neighbourhood knobs
 * (lr_w .. reg_c_day, w_nbr, max_neighbors, lambda1/2) switch on the tsvdx5
 * (global w/c) and tsvdx6 (plus same-day w_day/c_day) terms as the reference
 * does: w and c exist when lr_w > 0 or lr_c > 0, w_day and c_day when
 * lr_w_day > 0 or lr_c_day > 0 (src/tx.rs:483-585). */
typedef struct tx_config {
    uint64_t seed;
    int32_t n_feat, n_epochs, n_time_bins, n_freq_bins;
    int32_t reset_u_epoch, max_neighbors;
    int32_t shuffle_users, sum_err_bug, low_memory, full_su, save_ifeat, has_ordinal;
    float beta;
    float lr_u, lr_ud, lr_u2, lr_ub, lr_ubd, lr_i, lr_ib, lr_y, lr_yb, lr_yd;
    float lr_tu, lr_ti, lr_ta, lr_ibf, lr_iqf, lr_cu;
    float reg_iqf, reg_cu, reg_u, reg_u2, reg_ud, reg_i, reg_y, reg_yd;
    float sigma_iqf, sigma_u, sigma_i, sigma_y, sigma_yd;
    float lr_w, lr_c, reg_w, reg_c, lr_w_day, lr_c_day, reg_w_day, reg_c_day;
    float w_bias, w_factor, w_nbr;
    float lambda1, lambda2;
    float th_init[4];
    float th_gap, lr_t, reg_t;
    /* Tsvdx4pConfig extras (src/tsvdx4p.rs:128-187); unused by TxModel. */
    float nsvd_norm_exp;
    int32_t n_threads;
    int32_t seq_epochs[5];
} tx_config;

/* A rating set: Dataset (src/lib.rs:227-243).  For the held-out set passed
 * to tx_create (the reference's MaskedDataset, src/lib.rs:421-435)
 * residuals and raw_ratings are NULL. is_test may be NULL (all zero). */
typedef struct tx_ratings {
    int64_t n_ratings;
    int32_t n_users, n_items;
    const int32_t* user_idxs;
    const int32_t* item_idxs;
    const int16_t* dates;
    const float* residuals;
    const int8_t* raw_ratings;
    const int8_t* is_test;
    const int32_t* user_cnts;   /* [n_users] */
} tx_ratings;

/* Training schedule on the GPU.
 *   TX_SCHED_SERIAL: one wavefront walks users in the reference's shuffled
 *     order with the reference's per-rating arithmetic; results are
 *     identical to the serial reference trainer (src/tx.rs:734-1115).
 *   TX_SCHED_STRATA: all users in parallel, one team of lanes per user, items
 *     split into n_strata blocks processed as a rotating user-block x
 *     item-block schedule (n_strata = 1: every item in one stratum, each
 *     user's NSVD1 step at its end; n_strata > 1: the y rows stay fixed
 *     through the epoch's strata and the NSVD1 steps are applied per item after
 *     the last one, divided by n_contrib^nsvd_norm_exp, exponent 1 when
 *     nsvd_norm_exp is 0, as apply_nsvd_deltas, src/tsvdx4p.rs:341-385).
 *   TX_SCHED_TSVDX4P: the epoch of Tsvdx4pModel::fit_epoch
 *     (src/tsvdx4p.rs:798-805): a user pass with item tables frozen and the
 *     NSVD1 gradients accumulated per item, the delta apply divided by
 *     n_contrib^nsvd_norm_exp, the ycache rebuild, then an item pass with user
 *     tables frozen; epochs listed in cfg.seq_epochs run the TX_SCHED_STRATA
 *     epoch instead.  Low-memory models without ordinal head; the q_f stream
 *     reads bin 0 in the user pass and the day's real freq bin in the item
 *     pass, as the reference does. */
/*   TX_SCHED_PLAIN: the plain per-user wavefront kernel: `blocks` one-wave
 *     blocks (0 = 1,024) take the epoch's users in the reference's order, one
 *     after another per wave; each loads and waits for what a rating needs,
 *     atomic item updates, the y step at the user's end.  No prefetch, queue,
 *     length buckets or graph: the reference point the batched
 *     TX_SCHED_STRATA kernel is measured against.  Low-memory models without
 *     ordinal head, f <= 1024. */
enum tx_sched_mode { TX_SCHED_SERIAL = 0, TX_SCHED_STRATA = 1, TX_SCHED_TSVDX4P = 2, TX_SCHED_PLAIN = 3 };

typedef struct tx_schedule {
    int32_t mode;
    int32_t n_strata;        /* STRATA only: 1 or a multiple of 8 */
    int32_t teams_per_block; /* STRATA, f <= 1024: one-wave teams per 256-lane workgroup, 1..4
                                (0 = 4); blocks = 1 with 1 runs the kernel as a single team in
                                the user order (serial-debug order) */
    int32_t blocks;          /* STRATA only: persistent grid size, 0 = auto */
    int32_t use_graph;       /* capture each epoch's launches in a hipGraph */
    int32_t update_flags;    /* STRATA only: how shared rows are updated, see below */
    int32_t user_order;      /* STRATA only: 0 = the reference's per-epoch shuffled user
                                order (get_users), 1 = longest user first, 2 = length
                                buckets: the reference's order with the users of at
                                least heavy_min ratings spread evenly over the first
                                half of the epoch (so no long user starts near its end),
                                3 = length buckets launched separately (below) */
    int32_t heavy_min;       /* user_order 2: bound of the long-user bucket, 0 = 1024 */
    int32_t split_waves;     /* STRATA, one stratum, user_order 0-2, plain NSVD1 rows, low-memory
                                model without ordinal head or neighbourhood terms, 64-lane teams:
                                1 = two waves per team (128-thread workgroups, `blocks` = teams):
                                a rating wave and a streaming wave that takes the team's users
                                from the queue, gathers each one's NSVD1 sum before its rating
                                loop and applies its y step during the next user's loop; the
                                two meet once per user.  0 = one wave per team. */
} tx_schedule;
/* user_order 3 (one-stratum TX_SCHED_STRATA, f <= 1024): the users in length
 * buckets by training-rating count, [1, 16], [17, 256], [257, 4096] and
 * [4097, ...), each bucket its own queue (the reference's per-epoch order
 * within it), team shape (lanes per user: 16-lane teams, four users per
 * wave, or one 64-lane wave; waves per user: 1 or 2, each wave half of every
 * row) and number of teams in flight.  With one-wave teams in every bucket
 * (the default) one launch serves all buckets, its workgroups split between
 * the bucket queues; any other shape (16 lanes: padded row <= 256; two
 * waves: 128 <= padded row <= 1024; low-memory models without ordinal head
 * or neighbourhood terms only) gives each bucket a launch of its own on a
 * stream of its own, and kernels on streams that share a hardware queue run
 * one after the other.  Default teams: the fidelity bound's, split by each
 * bucket's ratings.  tx_autotune chooses shape and teams per bucket for this
 * order; tx_get_buckets reports them. */
#define TX_N_BUCKETS 4

/* update_flags: bits 0-1 item rows/biases (q, qf, b_i, b_i,bin, b_i,f),
 * bits 2-3 NSVD1 rows (y, y-bias):
 *   0 = device-scope atomic add of the SGD delta,
 *   1 = XCD-local atomic add (coherent only when one XCD owns the rows,
 *       i.e. item rows with n_strata >= 8),
 *   2 = plain read-modify-write (lock-free "Hogwild" updates).
 * Default (tx_create): atomic item rows, plain NSVD1 rows.
 * Bit 4 (TX_UPD_STRATA_EXPERIMENTAL) is accepted and ignored (it was required
 * for n_strata > 1 while that schedule applied the NSVD1 steps inside the last
 * stratum; it now gathers them per item after it).
 * Bits 5-6: the neighbourhood tables (w, c, w_day, c_day), 0 (default) or 2;
 * set as (mode << TX_UPD_NBR_SHIFT). */
#define TX_UPD_NBR_SHIFT 5
#define TX_UPD_ATOMIC 0
#define TX_UPD_XCD_ATOMIC 1
#define TX_UPD_PLAIN 2
#define TX_UPD_STRATA_EXPERIMENTAL 16

/* Per-epoch device timing (hipEvents on the library's stream). */
typedef struct tx_epoch_stats {
    double ms_total;        /* whole fit_epoch on the GPU (train + ycache rebuild) */
    double ms_sgd;          /* SGD kernels */
    double ms_rebuild;      /* ycache (+ ycache_day) rebuild */
    double ms_other;        /* merges / queue resets */
    int64_t n_ratings;      /* training ratings visited */
    int32_t n_launches;
    int32_t graph;          /* 1 when the epoch ran as a hipGraph */
} tx_epoch_stats;

/* Parameter ids for tx_get_param / tx_set_param, the tables of TxModel
 * (src/tx.rs:210-272).  Shapes: U = users, I = items, f = n_feat,
 * F = n_freq_bins, B = n_time_bins, D = (user, day) pairs. */
enum tx_param {
    TX_P_UBIAS = 0,      /* [U]          ubias        */
    TX_P_IBIAS,          /* [I]          ibias        */
    TX_P_UFEAT,          /* [U, f]       ufeat  (p)   */
    TX_P_UFEAT2,         /* [U, f]       ufeat2 (p2)  */
    TX_P_IFEAT,          /* [I, f]       ifeat  (q)   */
    TX_P_YFEAT,          /* [I, f]       yfeat  (y)   */
    TX_P_YCACHE,         /* [U, f]       ycache (s)   */
    TX_P_YFEAT_DAY,      /* [I, f]       (0 elements when low_memory) */
    TX_P_TU_MEAN,        /* [U]          */
    TX_P_ALPHA_U,        /* [U]          */
    TX_P_BUT_BIN,        /* [U, B]       */
    TX_P_BIT_BIN,        /* [I, B]       */
    TX_P_YBIAS,          /* [I]          */
    TX_P_YCACHE_BIAS,    /* [U]          */
    TX_P_UBIAS_DAY,      /* [D]          */
    TX_P_UFEAT_DAY,      /* [D, f]       (0 elements when low_memory) */
    TX_P_YCACHE_DAY,     /* [D, f]       (0 elements when low_memory) */
    TX_P_IBIAS_FREQ,     /* [I, F]       */
    TX_P_IFEAT_FREQ,     /* [I*F, f]     */
    TX_P_CU,             /* [U]          */
    TX_P_CUT,            /* [D]          */
    TX_P_GBIAS,          /* [1]          */
    TX_P_THRESHOLDS,     /* [4]          ordinal head thresholds */
    TX_P_NB_W,           /* [I, I]       w     (0 elements unless lr_w > 0 or lr_c > 0) */
    TX_P_NB_C,           /* [I, I]       c     (same) */
    TX_P_NB_W_DAY,       /* [I, I]       w_day (0 elements unless lr_w_day > 0 or lr_c_day > 0) */
    TX_P_NB_C_DAY,       /* [I, I]       c_day (same) */
    TX_P_COUNT
};

/* TxModel::new (src/tx.rs:460-649): builds SparseUD, tu_mean, day_range,
 * gbias, draws the initial factors from the reference's RNG stream
 * (StdRng seeded with cfg->seed) and uploads everything to the GPU. */
int tx_create(const tx_config* cfg, const tx_ratings* train, const tx_ratings* heldout, tx_ctx** out);

/* Drop the model and its GPU memory. */
int tx_destroy(tx_ctx* ctx);

/* Choose the training schedule (default: TX_SCHED_STRATA, n_strata = 1,
 * blocks = 0, update_flags = TX_UPD_ATOMIC | TX_UPD_PLAIN << 2, user_order
 * = 2, heavy_min = 1024).  Day terms (low_memory = 0) train in TX_SCHED_STRATA
 * with n_strata = 1 when every user's training ratings are in date order;
 * otherwise the default is TX_SCHED_SERIAL and TX_SCHED_STRATA returns
 * TX_EUNSUPPORTED. */
int tx_set_schedule(tx_ctx* ctx, const tx_schedule* sched);

/* The model's current schedule (after tx_create: its default schedule). */
int tx_get_schedule(tx_ctx* ctx, tx_schedule* out);

/* Autotuner of the parallel schedule (no reference counterpart): times
 * trial_epochs epochs for each candidate of teams in flight (1/2, 3/4 and all
 * of the measured fidelity bound for this f) x user order (the reference's,
 * or length buckets with heavy_min 256 / 1024 / 4096) x team form (one wave
 * per team, and split teams where the model takes them), then the fastest of
 * those over the stratum grids 8 x 8 and 64 x 64 (low-memory models without
 * ordinal head or neighbourhood terms), each trial from the same snapshot of
 * every parameter table, then restores the tables and applies the fastest
 * candidate (use_graph kept).  With user_order 3 it tunes each bucket's team
 * shape and teams instead (tx_get_buckets).  The model's parameters are
 * unchanged; *best (may be NULL) receives the chosen schedule.  Needs device
 * memory for one copy of the parameter tables.  Called on a one-stratum
 * TX_SCHED_STRATA schedule. */
int tx_autotune(tx_ctx* ctx, int32_t trial_epochs, tx_schedule* best);

/* One trial of the last tx_autotune (no reference counterpart). */
typedef struct tx_autotune_trial {
    int32_t bucket;          /* -1: a whole-schedule trial; 0..3: a length bucket's shape (user_order 3) */
    int32_t n_strata;        /* stratum grid (n_strata x n_strata user x item blocks) */
    int32_t blocks;          /* persistent grid of the trial */
    int32_t user_order, heavy_min;
    int32_t lanes_per_user, waves_per_user;  /* bucket trials: the bucket's team shape */
    int32_t teams;           /* teams in flight (of the bucket, for bucket trials) */
    int32_t chosen;          /* 1: the candidate tx_autotune applied */
    int32_t split_waves;     /* the trial's team form (tx_schedule.split_waves) */
    double ms;               /* epoch time (median of the trial's epochs) */
} tx_autotune_trial;

/* The trials of the last tx_autotune: *n = their count, the first max of them
 * copied to out. */
int tx_get_autotune_trials(tx_ctx* ctx, tx_autotune_trial* out, int32_t max, int32_t* n);

/* One length bucket of user_order 3 (no reference counterpart). */
typedef struct tx_bucket_info {
    int32_t lo, hi;            /* bounds of the bucket's training-rating counts (inclusive) */
    int64_t users, ratings;    /* users of the bucket and their training ratings */
    int32_t lanes_per_user;    /* 16 (four users per wave) or 64 */
    int32_t waves_per_user;    /* 1 or 2 */
    int32_t teams;             /* teams in flight */
    int32_t one_launch;        /* 1: the last bucketed epoch ran every bucket in one launch */
    double ms;                 /* the bucket's launch in the last epoch run with stream launches
                                  (with one launch: that launch), else 0 */
} tx_bucket_info;

/* The buckets of user_order 3 (out[TX_N_BUCKETS]); valid once a schedule
 * with user_order 3 has been set. */
int tx_get_buckets(tx_ctx* ctx, tx_bucket_info* out);

/* Set one bucket's team shape and teams in flight (user_order 3; the shapes
 * other than 64 lanes x 1 wave need a low-memory model without ordinal head
 * or neighbourhood terms: 16 lanes with a padded row <= 256, 2 waves with a
 * padded row of 128 to 1024, not both). */
int tx_set_bucket_shape(tx_ctx* ctx, int32_t bucket, int32_t lanes_per_user, int32_t waves_per_user, int32_t teams);

/* Regressor::fit_epoch (src/lib.rs:630, src/tx.rs:734-1115); epoch is
 * 1-based.  Includes the epoch-end ycache rebuild (src/tx.rs:1109-1110). */
int tx_fit_epoch(tx_ctx* ctx, int32_t epoch);

/* One user-only epoch: every user's parameters (p, p2, b_u, b_u,bin, alpha_u,
 * b_u,day, c_u) trained by SGD over its training ratings with all item
 * tables and y frozen, then the ycache rebuild.  This is the user pass of
 * Tsvdx4pModel (fit_epoch_user, src/tsvdx4p.rs:390-639) without its NSVD1
 * accumulators: folding users into fixed item tables. */
int tx_fit_users_epoch(tx_ctx* ctx, int32_t epoch);

/* Precision of the factor tables that predictions read (tx_predict*,
 * tx_calc_rmse): TX_PREC_F32 (default) or TX_PREC_BF16, BF16 copies of q,
 * q_f, p, p2 and ycache (round to nearest even) refreshed from the FP32
 * master tables after any change; training always updates the FP32 masters.
 * No reference counterpart (the reference computes in FP32 throughout). */
enum tx_precision { TX_PREC_F32 = 0, TX_PREC_BF16 = 1 };
int tx_set_table_precision(tx_ctx* ctx, int32_t precision);

/* Precision of the side a pass holds frozen (no reference counterpart; the
 * reference computes in FP32 throughout): with TX_PREC_BF16 the user passes of
 * TX_SCHED_TSVDX4P and tx_fit_users_epoch read q and q_f, and the item pass of
 * TX_SCHED_TSVDX4P reads p, p2 and ycache, from BF16 copies (round to nearest
 * even) made from the FP32 masters once before each pass; every update still
 * goes to the FP32 masters, in FP32.  TX_SCHED_STRATA epochs, which update
 * both sides per rating, are unaffected.  f <= 1024. */
int tx_set_frozen_precision(tx_ctx* ctx, int32_t precision);

/* rebuild_ycache (and rebuild_ycache_day when not low_memory) from the
 * current y tables (src/tx.rs:391-454): after tx_set_param of yfeat / ybias,
 * predictions use the new NSVD1 sums only once this has run. */
int tx_rebuild_ycache(tx_ctx* ctx);

/* Timing of the last tx_fit_epoch / tx_fit_users_epoch. */
int tx_last_epoch_stats(tx_ctx* ctx, tx_epoch_stats* out);

/* Regressor::predict (src/tx.rs:713-724) for n (user, item, day) triples
 * held in host memory. */
int tx_predict(tx_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items,
               const int16_t* days, float* out);

/* Same, with every buffer already in GPU memory; runs on `stream` (a
 * hipStream_t, NULL = the library's stream) and returns without waiting. */
int tx_predict_device(tx_ctx* ctx, int64_t n, const int32_t* d_users, const int32_t* d_items,
                      const int16_t* d_days, float* d_out, void* stream);

/* No counterpart in the reference.  tx_predict for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same predictions, staged through a
 * fixed scratch array of the library (27 MiB per device, held until the
 * process ends) instead of memory allocated on first use.  At most
 * TX_NOALLOC_MAX_ROWS queries per call (more: TX_EINVAL).  Every call on one
 * device shares that array, whatever its ctx: calls that overlap on one
 * device race. */
#define TX_NOALLOC_MAX_ROWS 1048576
int tx_predict_noalloc(tx_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days,
                       float* out);

/* Regressor::predict_subscores (src/tx.rs:663-711): bias, mf, nsvd1, and nbr
 * (nbr_score, src/tx.rs:341-389) when the model has the neighbourhood terms
 * (tx_n_subscores = 4); nbr may be NULL otherwise. */
int tx_predict_subscores(tx_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items,
                         const int16_t* days, float* bias, float* mf, float* nsvd1, float* nbr);

/* Regressor::n_subscores (src/tx.rs:653): 4 with the neighbourhood terms
 * ("bias", "mf", "nsvd1", "nbr"), else 3. */
int tx_n_subscores(tx_ctx* ctx, int32_t* out);

/* calc_rmse (src/lib.rs:659-673): FP64 RMSE of predict against residuals
 * over rows with is_test == 0. */
int tx_calc_rmse(tx_ctx* ctx, const tx_ratings* ds, double* out);

/* Element count of a parameter table. */
int tx_param_len(tx_ctx* ctx, int32_t param_id, int64_t* n);

/* Copy a parameter table to / from host memory (n must equal tx_param_len).
 * Used for the .ifeat artifact (save_artifacts, src/tx.rs:726-732) and to
 * load identical parameters into two implementations. */
int tx_get_param(tx_ctx* ctx, int32_t param_id, float* out, int64_t n);
int tx_set_param(tx_ctx* ctx, int32_t param_id, const float* in, int64_t n);

/* tx_set_param from GPU memory: `rows` rows of the table's column count,
 * read from `in` with a row pitch of `ld` floats, after the work queued on
 * `stream` (a hipStream_t, NULL: the default stream) has finished; returns
 * when the copy is done.  No counterpart in the reference: it is for callers
 * whose values are already on the device (the fold-in encoder's warm start of
 * p_u and b_u).  gbias and the tables whose all-zero state lets the trainer
 * skip work (cut, ibias_freq, but_bin, ybias) are refused (TX_EINVAL): set
 * them with tx_set_param. */
int tx_set_param_device(tx_ctx* ctx, int32_t param_id, const float* in, int64_t rows, int64_t ld, void* stream);

/* get_users (src/lib.rs:486-494): the per-epoch user order. Host only. */
int tx_get_users(int64_t n, int32_t shuffle, uint64_t seed, int64_t epoch, int64_t* out);

/* rand_array2 (src/lib.rs:64-67): n draws of Normal(0, sigma) from
 * StdRng::seed_from_u64(seed), row-major. Host only. */
int tx_rand_normal(uint64_t seed, float sigma, int64_t n, float* out);

/* Description of the last failure on this thread ("" if none). */
const char* tx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* TSVD_TX_H */
