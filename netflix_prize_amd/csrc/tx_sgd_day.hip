// k_sgd_par instantiations: the SGD epoch (PASS 0) with day terms.
#include "tx_device.hpp"
#include "tx_launch.hpp"

namespace tx {

hipError_t launch_sgd_par_day(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                              bool qf, bool p2, bool ord, hipStream_t st) {
#define X(T, V) \
    if (team == T && vec == V) return sgd_par_tvp<T, V, 0, true>(m, sv, stratum, blocks, qf, p2, ord, st);
    TX_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace tx
