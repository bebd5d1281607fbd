// k_sgd_par instantiations: the SGD epoch with the neighbourhood global w/c terms (tsvdx5).
#include "tx_device.hpp"
#include "tx_launch.hpp"

namespace tx {

hipError_t launch_sgd_par_nb1(const DevModel& m, const StrataView& sv, int stratum, int team, int vec, int blocks,
                               bool ord, bool day, hipStream_t st) {
#define X(T, V)                                                                                       \
    if (team == T && vec == V)                                                                        \
        return day ? sgd_par_nb_tvp<T, V, true, 1>(m, sv, stratum, blocks, ord, st)                 \
                   : sgd_par_nb_tvp<T, V, false, 1>(m, sv, stratum, blocks, ord, st);
    TX_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_sgd_par_nb1_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool day,
                                  hipStream_t st) {
    return day ? sgd_par_nb_mw<true, 1>(m, sv, teams, waves, st) : sgd_par_nb_mw<false, 1>(m, sv, teams, waves, st);
}

}  // namespace tx
