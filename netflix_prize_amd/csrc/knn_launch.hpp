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
};

hipError_t launch_norms(const float* F, int fp, int rows, float* norm, hipStream_t st);
hipError_t launch_cosine(const float* F, int fp, const float* norm, float* S, int ip, hipStream_t st);
// support matrix of the per-user item lists (off[U+1], items), then the
// KnnsModel similarity in place; S is ip x ip and zeroed by the caller
hipError_t launch_knns_sim(const uint32_t* off, const int32_t* items, int32_t n_users, int64_t n_rows, float* S,
                           float* diag, int ip, float shrinkage, hipStream_t st);
hipError_t launch_knnf_predict(const KnnView& v, int64_t n, const int32_t* us, const int32_t* is, const int16_t* ds,
                               float* out, hipStream_t st);

}  // namespace knn
