// k_sgd_par instantiations for the length-bucket shapes of the bucketed
// epoch (user_order 3), SGD epoch of low-memory models without ordinal head:
//   short users (1-16 ratings): 16-lane teams, four users per wave;
//   long users (4,097 ratings and more): two-wave teams, each wave a half of
//   every row.
#include "tx_device.hpp"
#include "tx_launch.hpp"

namespace tx {

template <int TEAM, int VEC, int MW>
static hipError_t sgd_par_shape(const DevModel& m, const StrataView& sv, int blocks, int threads, bool qf, bool p2,
                                hipStream_t st) {
    (void)hipGetLastError();
#define TX_B(QF, P2) \
    hipLaunchKernelGGL((k_sgd_par<TEAM, VEC, QF, P2, false, 0, false, MW>), dim3(blocks), dim3(threads), 0, st, m, sv, 0)
    if (qf) { if (p2) TX_B(true, true); else TX_B(true, false); }
    else { if (p2) TX_B(false, true); else TX_B(false, false); }
#undef TX_B
    return hipGetLastError();
}

hipError_t launch_sgd_par_t16(const DevModel& m, const StrataView& sv, int blocks, bool qf, bool p2, hipStream_t st) {
    if (sv.n_strata != 1 || m.low_mem == 0 || m.has_ord || m.nb_on) return hipErrorInvalidValue;
    switch (m.fs) {
        case 64: return sgd_par_shape<16, 4, 0>(m, sv, blocks, 256, qf, p2, st);
        case 128: return sgd_par_shape<16, 8, 0>(m, sv, blocks, 256, qf, p2, st);
        case 256: return sgd_par_shape<16, 16, 0>(m, sv, blocks, 256, qf, p2, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sgd_par_w2(const DevModel& m, const StrataView& sv, int teams, bool qf, bool p2, hipStream_t st) {
    if (sv.n_strata != 1 || m.low_mem == 0 || m.has_ord || m.nb_on) return hipErrorInvalidValue;
    switch (m.fs) {
        case 128: return sgd_par_shape<64, 1, 4>(m, sv, teams, 128, qf, p2, st);
        case 256: return sgd_par_shape<64, 2, 4>(m, sv, teams, 128, qf, p2, st);
        case 512: return sgd_par_shape<64, 4, 4>(m, sv, teams, 128, qf, p2, st);
        case 1024: return sgd_par_shape<64, 8, 4>(m, sv, teams, 128, qf, p2, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace tx
