/* C ABI of the GPU KnnfModel (libtsvd_tx.so): item-item kNN over a factor
 * table produced by an MF-family model (the `.ifeat` a tsvdx4 job saves),
 * with temporal decay at predict time.  Drop-in for KnnfModel
 * (src/knnf.rs:1-144) behind the reference's Regressor trait; used by the
 * `*__knnf` jobs (src/bin/tsvdx4-new.rs:55-57, src/bin/tsvdx4p-new.rs:24).
 *
 * Conventions as include/tsvd_tx.h: calls return 0 or a negative TX_E* code
 * with tx_last_error() describing it; host buffers are copied, never kept. */
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

/* The similarity of items i and j as the model uses it (tests and tools). */
int knnf_similarity(knnf_ctx* ctx, int64_t n, const int32_t* i, const int32_t* j, float* out);

int knnf_destroy(knnf_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* TSVD_KNN_H */
