/* C ABI of the GPU KnnfModel, KnnsModel and Knn3Model (libtsvd_tx.so):
 * item-item kNN with temporal decay at predict time, over the cosine of a
 * factor table (the `.ifeat` a tsvdx4 job saves; KnnfModel, src/knnf.rs:1-144)
 * or over the shrunk lift of item co-occurrence (KnnsModel, src/knns.rs:1-149);
 * and item-item regression weights by NNLS (Knn3Model, src/knn3.rs:1-396).
 * Drop-ins behind the reference's Regressor trait for the `*__knnf`,
 * `*__knns` and `*__knn3` jobs (src/bin/tsvdx4-new.rs:43-57,
 * src/bin/tsvdx4p-new.rs:24).
 *
 * Conventions as include/tsvd_tx.h: calls return 0 or a negative TX_E* code
 * with tx_last_error() describing it; host buffers are copied, never kept.
 * The *_predict calls stage through device memory allocated per call, so
 * they may run concurrently from several threads on one ctx; create and
 * destroy need the ctx to themselves. */
#ifndef TSVD_KNN_H
#define TSVD_KNN_H

#include <stdint.h>

#include "tsvd_tx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* KnnfConfig (src/knnf.rs:8-27) without the factor path: the caller passes
 * the factor table itself (KnnfModel::new_with_factors, src/knnf.rs:60-90).
 * Defaults (KnnfConfig::with_factors): k = 25, scaling = 1, tau = 0.025. */
typedef struct knnf_config {
    int32_t k;
    float scaling;
    float tau;
} knnf_config;

typedef struct knnf_ctx knnf_ctx;

/* KnnfModel::new_with_factors (src/knnf.rs:60-90): factors is n_items x
 * n_feat row-major FP32 (the `.ifeat` layout); train gives each user's
 * training ratings (user-sorted, residual targets).  Builds the item-item
 * similarity max(0, cos) on the GPU (FP32 MFMA). */
int knnf_create(const knnf_config* cfg, const float* factors, int32_t n_items, int32_t n_feat,
                const tx_ratings* train, knnf_ctx** out);

/* KnnfModel::predict (src/knnf.rs:101-141) for n queries, out[n] on the host. */
int knnf_predict(knnf_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days,
                 float* out);

/* KnnsConfig (src/knns.rs:19-43) with SuppSource::Compute, the source the
 * reference's jobs use (src/bin/tsvdx4-new.rs:563-574). */
typedef struct knns_config {
    int32_t k;
    float shrinkage;
    float scaling;
    float tau;
} knns_config;

/* KnnsModel::new with SuppSource::Compute (src/knns.rs:78-100): the support
 * matrix of build_bin_nbstats(train, heldout, "supp") (src/nbstats.rs:73-117,
 * every ordered pair of each user's training and held-out items) and the
 * similarity phi n / (n + shrinkage), phi = n N_USERS / (supp_ii supp_jj)
 * (src/knns.rs:64-70), built on the GPU.  heldout needs only users and items
 * (MaskedDataset). */
int knns_create(const knns_config* cfg, const tx_ratings* train, const tx_ratings* heldout, knnf_ctx** out);

/* KnnsModel::predict (src/knns.rs:106-147): the same rule as knnf_predict. */
int knns_predict(knnf_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days,
                 float* out);
int knns_destroy(knnf_ctx* ctx);

/* KnnConfig (src/knn.rs:19-57) in the form the `*__knn-d` job uses
 * (src/bin/tsvdx4-new.rs:590-610): SimType::Support over a loaded support
 * stat, use_stat1 on, no threshold, no regression. */
typedef struct knnd_config {
    int32_t k;
    float shrinkage;
    float scaling;
    float tau;
} knnd_config;

/* KnnModel::new (src/knn.rs:191-256) with stat0 and stat1 given as
 * n_items x n_items row-major FP32 tables (the `{sim_dir}/{name}.{set}.npy`
 * files it loads): similarity phi n / (n + shrinkage), phi = n 480,189 /
 * (stat0_ii stat0_jj) (support_sim, :105-113), and the averaged value
 * stat1_ij / (stat0_ij + shrinkage) in place of the residual (:321-334). */
int knnd_create(const knnd_config* cfg, const float* stat0, const float* stat1, int32_t n_items,
                const tx_ratings* train, knnf_ctx** out);

/* KnnModel::predict (src/knn.rs:263-336) on the use_stat1 path; the same
 * top-k rule as knnf_predict. */
int knnd_predict(knnf_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days,
                 float* out);
int knnd_destroy(knnf_ctx* ctx);

/* Knn3Config (src/knn3.rs:7-31).  Defaults (Knn3Config::default): threshold
 * 0.25, k_min 10, k_max 60, shrinkage 20000, reg 0.01, x 0.8, bl_reg_m 25,
 * bl_reg_u 10.  k_max above 64 is TX_EUNSUPPORTED. */
typedef struct knn3_config {
    float threshold;
    int32_t k_min, k_max;
    float shrinkage, reg, x, bl_reg_m, bl_reg_u;
} knn3_config;

typedef struct knn3_ctx knn3_ctx;

/* Knn3Model::new (src/knn3.rs:350-362): over the non-test training ratings
 * (train needs residuals and raw_ratings; is_test NULL means none), the
 * regularised baseline (compute_baseline, :123-165), per item the candidates
 * by shrunk lift of co-rating counts (:195-216), the mixed FP64 normal
 * equations of the residual and baseline-corrected signals (:222-286) and
 * their NNLS solution (solve_nnls, :79-120), built on the GPU.  One rating per
 * user and item (TX_EINVAL otherwise); at most 32,768 items. */
int knn3_create(const knn3_config* cfg, const tx_ratings* train, knn3_ctx** out);

/* Knn3Model::predict (src/knn3.rs:379-389): sum of weight x residual over the
 * item's weighted neighbours the user rated; days may be NULL (unused). */
int knn3_predict(knn3_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days,
                 float* out);

/* The model's neighbour table: counts[n_items], items and weights
 * [n_items x 64], row m holding item m's positive weights in candidate order
 * (tests and tools). */
int knn3_weights(knn3_ctx* ctx, int32_t* counts, int32_t* items, float* weights);

/* The candidate lists before the regression: counts[n_items], items
 * [n_items x 64] by prox, largest first (tests and tools). */
int knn3_candidates(knn3_ctx* ctx, int32_t* counts, int32_t* items);

/* The regularised baseline: *mu, bm[n_items], bu[n_users] (tests and tools). */
int knn3_baseline(knn3_ctx* ctx, float* mu, float* bm, float* bu);

int knn3_destroy(knn3_ctx* ctx);

/* The similarity of items i and j as the model uses it (tests and tools). */
int knnf_similarity(knnf_ctx* ctx, int64_t n, const int32_t* i, const int32_t* j, float* out);

int knnf_destroy(knnf_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* TSVD_KNN_H */
