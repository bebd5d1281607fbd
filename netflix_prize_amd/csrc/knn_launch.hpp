// Device view and launchers of the KnnfModel kernels (knn_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace knn {

struct KnnView {
    const float* S;          // [ip x ip] item-item similarity max(0, cos), zero rows past the items
    int32_t ip;              // padded item count (multiple of 64)
    const uint32_t* t_off;   // [U+1] per-user offsets into the training ratings
    const int32_t* t_item;   // [N] rated item
    const float* t_res;      // [N] residual target
    const float* t_day;      // [N] day as f32
    int32_t k;               // neighbours kept
    float scaling, tau;      // similarity exponent, temporal decay
    const float* D = nullptr;  // [ip x ip] per-pair target replacing the residual (KnnModel use_stat1), or none
};

hipError_t launch_norms(const float* F, int fp, int rows, float* norm, hipStream_t st);
hipError_t launch_cosine(const float* F, int fp, const float* norm, float* S, int ip, hipStream_t st);
// support matrix of the per-user item lists (off[U+1], items sorted within
// each list), then the KnnsModel similarity in place; S is ip x ip (every
// element written); scratch holds support_scratch_words(n_rows, ip) words
size_t support_scratch_words(int64_t n_rows, int ip);
hipError_t launch_knns_sim(const uint32_t* off, const int32_t* items, int32_t n_users, int64_t n_rows, float* S,
                           float* diag, int ip, float shrinkage, uint32_t* scratch, hipStream_t st);
// the support matrix alone (the first two steps of launch_knns_sim) and its diagonal
hipError_t launch_support(const uint32_t* off, const int32_t* items, int32_t n_users, int64_t n_rows, float* S,
                          float* diag, int ip, uint32_t* scratch, hipStream_t st);
// KnnModel with SimType::Support and use_stat1 over loaded stats: S (stat0 in)
// becomes the support similarity, D (stat1 in) becomes stat1 / (stat0 + shrinkage)
hipError_t launch_stat_sim(float* S, float* D, float* diag, int ip, float shrinkage, hipStream_t st);
hipError_t launch_knnf_predict(const KnnView& v, int64_t n, const int32_t* us, const int32_t* is, const int16_t* ds,
                               float* out, hipStream_t st);

// Knn3Model (src/knn3.rs) device state.  Ratings are the non-test training
// ratings only; each user's list is sorted by item, each item's list of users
// ascending with the rating's own residual and r_prime beside it.
struct Knn3View {
    int32_t n_items, ip, n_users;
    const float* cnt;        // [ip x ip] co-rating counts, diagonal = users of the item
    const float* diag;       // [ip] users of each item
    const uint32_t* u_off;   // [U+1]
    const int32_t* u_item;   // user lists, items ascending
    const float* u_err;      // residual target
    const float* u_rp;       // r_prime (raw minus the regularised baseline)
    const uint32_t* i_off;   // [I+1]
    const int32_t* i_user;   // item lists, users ascending
    const float* i_err;
    const float* i_rp;
    const int32_t* heavy_first;  // [I] items by user count, largest first (k3_cov's block order)
    int32_t* nb_cnt;         // [I] candidates kept
    int32_t* nb_item;        // [I x 64] in candidate order
    double* sys;             // [I x 64 x 64] the mixed, regularised normal equations
    double* rhs;             // [I x 64]
    int32_t* w_cnt;          // [I] positive weights
    int32_t* w_item;         // [I x 64]
    float* w_val;            // [I x 64]
    float threshold, shrinkage;
    int32_t k_min, k_max;
    double x, reg;
};

// candidates, normal equations and NNLS weights for items [0, n_items)
hipError_t launch_knn3_build(const Knn3View& v, hipStream_t st);
hipError_t launch_knn3_predict(const Knn3View& v, int64_t n, const int32_t* us, const int32_t* is, float* out,
                               hipStream_t st);

}  // namespace knn
