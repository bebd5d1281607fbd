// k_sgd_par instantiations: split teams (a rating wave and a streaming wave
// per user stream) for the SGD epoch of low-memory models in one stratum.
#include "tx_device.hpp"
#include "tx_launch.hpp"

namespace tx {

template <int V>
hipError_t sgd_par_split_v(const DevModel& m, const StrataView& sv, int teams, bool qf, bool p2, hipStream_t st) {
    (void)hipGetLastError();
#define TX_SPLIT(QF, P2) \
    hipLaunchKernelGGL((k_sgd_par<64, V, QF, P2, false, 0, false, 0, 0, true>), dim3(teams), dim3(128), 0, st, m, sv, 0)
    if (qf) { if (p2) TX_SPLIT(true, true); else TX_SPLIT(true, false); }
    else { if (p2) TX_SPLIT(false, true); else TX_SPLIT(false, false); }
#undef TX_SPLIT
    return hipGetLastError();
}

hipError_t launch_sgd_par_split(const DevModel& m, const StrataView& sv, int vec, int teams, bool qf, bool p2,
                                hipStream_t st) {
    if (teams < 1) return hipErrorInvalidValue;
#define X(T, V) \
    if (vec == V) return sgd_par_split_v<V>(m, sv, teams, qf, p2, st);
    TX_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace tx
