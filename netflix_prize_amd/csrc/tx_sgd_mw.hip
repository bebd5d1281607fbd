// k_sgd_par instantiations: multi-wave teams (f > 1024).
#include "tx_device.hpp"
#include "tx_launch.hpp"

namespace tx {

template <bool DAY, int MW, int PASS = 0>
static hipError_t sgd_par_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool qf, bool p2, bool ord,
                             hipStream_t st) {
    (void)hipGetLastError();
#define TX_MW(QF, P2, ORD)                                                                                       \
    hipLaunchKernelGGL((k_sgd_par<64, 16, QF, P2, ORD, PASS, DAY, MW>), dim3(teams), dim3(64 * waves), 0, st, m, \
                       sv, 0)
    if constexpr (PASS != 0) {
        if (ord || DAY) return hipErrorInvalidValue;
        if (qf) { if (p2) TX_MW(true, true, false); else TX_MW(true, false, false); }
        else { if (p2) TX_MW(false, true, false); else TX_MW(false, false, false); }
        return hipGetLastError();
    } else if (ord) {
        if (qf) { if (p2) TX_MW(true, true, true); else TX_MW(true, false, true); }
        else { if (p2) TX_MW(false, true, true); else TX_MW(false, false, true); }
    } else {
        if (qf) { if (p2) TX_MW(true, true, false); else TX_MW(true, false, false); }
        else { if (p2) TX_MW(false, true, false); else TX_MW(false, false, false); }
    }
#undef TX_MW
    return hipGetLastError();
}

hipError_t launch_sgd_par_mw(const DevModel& m, const StrataView& sv, int teams, int waves, bool qf, bool p2, bool ord,
                             bool day, hipStream_t st, int pass) {
    if (waves < 2 || waves > 8 || m.fs != 1024 * waves || sv.n_strata != 1 || (day && !sv.day_acc))
        return hipErrorInvalidValue;
    if (m.nb_on && pass == 0) {
        if (ord || m.nb_max < 1 || m.nb_max > 128) return hipErrorInvalidValue;
        return m.nbd_on ? launch_sgd_par_nb2_mw(m, sv, teams, waves, day, st)
                        : launch_sgd_par_nb1_mw(m, sv, teams, waves, day, st);
    }
    if (pass == 1 || pass == 2) {
        if (day || ord) return hipErrorInvalidValue;
        if (waves <= 4)
            return pass == 1 ? sgd_par_mw<false, 4, 1>(m, sv, teams, waves, qf, p2, false, st)
                             : sgd_par_mw<false, 4, 2>(m, sv, teams, waves, qf, p2, false, st);
        return pass == 1 ? sgd_par_mw<false, 8, 1>(m, sv, teams, waves, qf, p2, false, st)
                         : sgd_par_mw<false, 8, 2>(m, sv, teams, waves, qf, p2, false, st);
    }
    if (pass != 0) return hipErrorInvalidValue;
    if (waves <= 4)
        return day ? sgd_par_mw<true, 4>(m, sv, teams, waves, qf, p2, ord, st)
                   : sgd_par_mw<false, 4>(m, sv, teams, waves, qf, p2, ord, st);
    return day ? sgd_par_mw<true, 8>(m, sv, teams, waves, qf, p2, ord, st)
               : sgd_par_mw<false, 8>(m, sv, teams, waves, qf, p2, ord, st);
}

}  // namespace tx
