// Kernel launchers (defined in tx_kernels.hip and tx_sgd_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tx_types.hpp"

namespace tx {

hipError_t launch_sgd_serial(const DevModel& m, const int64_t* order, float* scratch, hipStream_t st);
// the plain per-user wavefront kernel: `blocks` one-wave blocks over the epoch's user order
hipError_t launch_sgd_plain(const DevModel& m, const int64_t* order, int vec, int blocks, bool qf, hipStream_t st);
// pass: 0 SGD epoch, 1 tsvdx4p user pass (NSVD1 accumulators), 2 user-only fold-in pass;
// day: day terms (pass 0, one stratum, sv.day_acc allocated)
hipError_t launch_sgd_par(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                          bool qf, bool p2, bool ord, int pass, bool day, hipStream_t st);
// multi-wave teams (f > 1024): `teams` workgroups of 64*waves threads, one user
// each, fs = 1024*waves, one stratum; pass as for launch_sgd_par
hipError_t launch_sgd_par_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool qf, bool p2, bool ord,
                             bool day, hipStream_t st, int pass = 0);
// the variant families behind launch_sgd_par (one translation unit each)
hipError_t launch_sgd_par_lm(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                             bool qf, bool p2, bool ord, hipStream_t st);
hipError_t launch_sgd_par_day(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                              bool qf, bool p2, bool ord, hipStream_t st);
hipError_t launch_sgd_par_user(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                               bool qf, bool p2, int pass, hipStream_t st);
// neighbourhood models (m.nb_on; launch_sgd_par and launch_sgd_par_mw route
// PASS 0 here): 1 = global terms, 2 = global and same-day terms
hipError_t launch_sgd_par_nb1(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                              bool ord, bool day, hipStream_t st);
hipError_t launch_sgd_par_nb2(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                              bool ord, bool day, hipStream_t st);
hipError_t launch_sgd_par_nb1_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool day,
                                 hipStream_t st);
hipError_t launch_sgd_par_nb2_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool day,
                                 hipStream_t st);
// length-bucket shapes (low-memory models, no ordinal head, one stratum):
// 16-lane teams, 16 per 256-thread block (fs 64..256); two-wave teams, one
// per 128-thread block (fs 128..1024)
hipError_t launch_sgd_par_t16(const DevModel& m, const StrataView& sv, int blocks, bool qf, bool p2, hipStream_t st);
hipError_t launch_sgd_par_w2(const DevModel& m, const StrataView& sv, int teams, bool qf, bool p2, hipStream_t st);
// split teams (SPLIT in k_sgd_par): `teams` workgroups of 128 threads, a
// rating wave and a streaming wave per user stream (one stratum, plain y rows,
// low-memory models without ordinal head or neighbourhood terms, 64-lane teams)
hipError_t launch_sgd_par_split(const DevModel& m, const StrataView& sv, int vec, int teams, bool qf, bool p2,
                                hipStream_t st);
hipError_t launch_to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t st);
// the NSVD1 step of the tsvdx4p user pass from the stored per-user gradients
// (gacc [U * fs], gacc_b [U]) gathered per item; vec = the shape's VEC
hipError_t launch_nsvd_gather(const DevModel& m, const ItemView& iv, const float* gacc, const float* gacc_b,
                              const float* scale, int vec, hipStream_t st);
// waves > 1: multi-wave teams (f > 1024), one item per block of 64*waves threads
hipError_t launch_item_pass(const DevModel& m, const ItemView& iv, int team, int vec, bool p2, bool qf,
                            int blocks, hipStream_t st, int waves = 1);
hipError_t launch_rebuild(const DevModel& m, int64_t n_ud, int team, int vec, hipStream_t st);
// mode 0 predict (o0), 1 subscores (o0..o2, and o3 = nbr when m.nb_on and o3 != NULL), 2 FP64 SSE
hipError_t launch_predict(const DevModel& m, int mode, int64_t n, const int32_t* us, const int32_t* is,
                          const int16_t* ds, float* o0, float* o1, float* o2, float* o3, const float* target,
                          const int8_t* is_test, double* out_d, int team, int vec, hipStream_t st);
hipError_t launch_fill_f32(float* p, float v, int64_t n, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);
hipError_t launch_fill_f64(double* p, double v, int64_t n, hipStream_t st);

// tx_predict_noalloc's scratch array: staging for one batch of kNoAllocRows
// host-buffer queries (user, item i32; day i16; four float outputs; is_test
// i8; each part 256-byte aligned) followed by two FP64 sums.
constexpr int64_t kNoAllocRows = int64_t{1} << 20;
constexpr int64_t kNoAllocBytes = kNoAllocRows * (4 + 4 + 2 + 4 * 4 + 1) + 256;
// Device address of that array on `device` (the current device), looked up
// once per device and kept.
hipError_t noalloc_scratch(int device, void** out);

}  // namespace tx
