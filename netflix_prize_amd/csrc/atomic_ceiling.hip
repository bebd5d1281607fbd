// atomic-ceiling: the device's throughput for the update pattern of the SGD
// kernel's item rows, which bounds that kernel at large f.
//
// Each wave picks pseudo-random rows of a 17,770 x 256 FP32 table (the q
// table at f=256, 18.2 MB) and updates the whole 1 KB row, in one of these
// forms:
//   f32-atomic       four 64-lane FP32 atomic adds (the SGD kernel's q update)
//   u64-atomic       two 64-lane 64-bit integer atomic adds (same bytes)
//   plain-rmw        four 64-lane loads, then four stores
//   loads            four 64-lane loads
//   bias-atomic      one FP32 atomic add with two active lanes on the row's
//                    128-byte bias row (b_i and one time bin of a 17,770 x 32
//                    table: the SGD kernel's item-bias update)
//   row+bias-atomic  f32-atomic and bias-atomic on the same row: the SGD
//                    kernel's whole per-rating item update at f=256
//   scatter-atomic   one 64-lane FP32 atomic add per "row", each lane on a
//                    pseudo-random element of a 17,770 x 17,770 table (1.26 GB,
//                    the neighbourhood tables' shape): 64 lines per instruction
// and prints rows/s, lanes/s and atomic instructions/s per form and wave
// count, one line each.
// Usage: atomic-ceiling [iterations per wave]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                     \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

namespace {

constexpr int kRows = 17770;
constexpr int kRowFloats = 256;

__device__ __forceinline__ uint32_t mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// one wave per block; row r of the table is q[r * kRowFloats, (r + 1) * kRowFloats)
template <int FORM>
__global__ __launch_bounds__(64) void k_rows(float* q, float* bias, float* big, int iters) {
    const int lane = threadIdx.x;
    float acc = 0.0f;
    for (int it = 0; it < iters; ++it) {
        const uint32_t r = mix(blockIdx.x * 1000003u + static_cast<uint32_t>(it)) % kRows;
        float* p = q + static_cast<size_t>(r) * kRowFloats + lane;
        if constexpr (FORM == 0 || FORM == 4 || FORM == 5) {
            if constexpr (FORM != 4) {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    __hip_atomic_fetch_add(p + c * 64, 1e-7f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if constexpr (FORM != 0) {
                // lane 0: b_i at offset 0; lane 1: time bin 1 + it mod 30 (< 32)
                if (lane < 2)
                    __hip_atomic_fetch_add(bias + static_cast<size_t>(r) * 32 + (lane == 0 ? 0 : 1 + it % 30), 1e-7f,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if constexpr (FORM == 6) {
            // (i, j) from the lane's own hash: both < kRows
            const uint32_t h = mix(r * 64u + static_cast<uint32_t>(lane) + 0x9e3779b9u * static_cast<uint32_t>(it));
            const uint32_t j = h % kRows;
            __hip_atomic_fetch_add(big + static_cast<size_t>(r) * kRows + j, 1e-7f, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else if constexpr (FORM == 1) {
            unsigned long long* p8 = reinterpret_cast<unsigned long long*>(q + static_cast<size_t>(r) * kRowFloats) + lane;
#pragma unroll
            for (int c = 0; c < 2; ++c) __hip_atomic_fetch_add(p8 + c * 64, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if constexpr (FORM == 2) {
            float v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = p[c * 64];
#pragma unroll
            for (int c = 0; c < 4; ++c) p[c * 64] = v[c] + 1e-7f;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += p[c * 64];
        }
    }
    if constexpr (FORM == 3) {
        if (acc == 12345.0f) q[lane] = acc;  // keeps the loads
    }
}

}  // namespace

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 2000;
    if (iters < 1 || iters > 100000) {
        std::fprintf(stderr, "iterations per wave: 1..100000\n");
        return 2;
    }
    float* q = nullptr;
    float* bias = nullptr;
    CK(hipMalloc(&q, sizeof(float) * kRows * kRowFloats));
    CK(hipMemset(q, 0, sizeof(float) * kRows * kRowFloats));
    CK(hipMalloc(&bias, sizeof(float) * kRows * 32));
    CK(hipMemset(bias, 0, sizeof(float) * kRows * 32));
    float* big = nullptr;
    const size_t n_big = static_cast<size_t>(kRows) * kRows;
    CK(hipMalloc(&big, sizeof(float) * n_big));
    CK(hipMemset(big, 0, sizeof(float) * n_big));
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    const char* names[7] = {"f32-atomic",  "u64-atomic",      "plain-rmw",     "loads",
                            "bias-atomic", "row+bias-atomic", "scatter-atomic"};
    const int lanes_per_row[7] = {256, 128, 256, 256, 2, 258, 64};
    const int atomics_per_row[7] = {4, 2, 0, 0, 1, 5, 1};  // atomic instructions per row and wave
    const int waves_list[3] = {1536, 4096, 16384};
    for (int form = 0; form < 7; ++form) {
        for (int w : waves_list) {
            float ms = 0.0f;
            for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
                CK(hipEventRecord(a));
                switch (form) {
                    case 0: hipLaunchKernelGGL(k_rows<0>, dim3(w), dim3(64), 0, 0, q, bias, big, iters); break;
                    case 1: hipLaunchKernelGGL(k_rows<1>, dim3(w), dim3(64), 0, 0, q, bias, big, iters); break;
                    case 2: hipLaunchKernelGGL(k_rows<2>, dim3(w), dim3(64), 0, 0, q, bias, big, iters); break;
                    case 3: hipLaunchKernelGGL(k_rows<3>, dim3(w), dim3(64), 0, 0, q, bias, big, iters); break;
                    case 4: hipLaunchKernelGGL(k_rows<4>, dim3(w), dim3(64), 0, 0, q, bias, big, iters); break;
                    case 5: hipLaunchKernelGGL(k_rows<5>, dim3(w), dim3(64), 0, 0, q, bias, big, iters); break;
                    default: hipLaunchKernelGGL(k_rows<6>, dim3(w), dim3(64), 0, 0, q, bias, big, iters); break;
                }
                CK(hipGetLastError());
                CK(hipEventRecord(b));
                CK(hipEventSynchronize(b));
                CK(hipEventElapsedTime(&ms, a, b));
            }
            const double rows = static_cast<double>(w) * iters;
            std::printf("{\"form\": \"%s\", \"waves\": %d, \"ms\": %.3f, \"rows_per_s\": %.4e, \"lanes_per_s\": %.4e, "
                        "\"atomic_instr_per_s\": %.4e}\n",
                        names[form], w, ms, rows / (ms * 1e-3), rows * lanes_per_row[form] / (ms * 1e-3),
                        rows * atomics_per_row[form] / (ms * 1e-3));
            std::fflush(stdout);
        }
    }
    CK(hipFree(q));
    CK(hipFree(bias));
    CK(hipFree(big));
    return 0;
}
