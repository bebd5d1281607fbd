// k_sgd_par instantiations: the user passes (PASS 1: tsvdx4p user pass with
// NSVD1 accumulators, PASS 2: fold-in).
#include "tx_device.hpp"
#include "tx_launch.hpp"

namespace tx {

hipError_t launch_sgd_par_user(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                               bool qf, bool p2, int pass, hipStream_t st) {
#define X(T, V)                                                                                          \
    if (team == T && vec == V)                                                                           \
        return pass == 1 ? sgd_par_tvp<T, V, 1, false>(m, sv, stratum, blocks, qf, p2, false, st)         \
                         : sgd_par_tvp<T, V, 2, false>(m, sv, stratum, blocks, qf, p2, false, st);
    TX_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace tx
