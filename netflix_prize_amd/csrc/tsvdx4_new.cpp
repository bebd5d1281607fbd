// tsvdx4-new: job-level drop-in for the reference's `tsvdx4-new <job>`
// (src/bin/tsvdx4-new.rs), running TxModel on the GPU through libtsvd_tx.so;
// the same binary (also installed as tsvdx4p-new) runs the `tsvdx4p-new <job>`
// jobs (src/bin/tsvdx4p-new.rs) as Tsvdx4pModel epochs.
//
// It restates the reference driver: Dataset::load with target specs
// (src/lib.rs:252-290), fit2_inner phase 1 split.tr -> split.pr and phase 2
// fulltrain -> qual through fit (src/lib.rs:764-815, 881-996), report_rmse
// (src/lib.rs:678-686), gen_preds_parallel + write_npy outputs, subscores,
// train predictions, save_artifacts (.ifeat), the `.out` log, and the
// `__epochs` blend jobs (epoch_blend_apply, src/lib.rs:1181-1235).
//
// Environment knobs (the reference has none; they select GPU schedules and
// let synthetic data of another shape through Dataset::load):
//   TSVD_SCHEDULE = strata | serial   (default: the library's, one-stratum
//                                      parallel; serial for day terms on
//                                      ratings not sorted by date per user;
//                                      Tsvdx4pModel epochs for tsvdx4p jobs)
//   TSVD_STRATA   = 1 | 8 | ...       (strata of the parallel schedule)
//   TSVD_N_USERS / TSVD_N_ITEMS       (default 480189 / 17770, as the reference)
//   TSVD_DATA_DIR                     (default "data")
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "npy.hpp"
#include "tsvd_knn.h"
#include "tsvd_tx.h"

namespace {

// ---------------------------------------------------------------------------
// tee logging (src/lib.rs:93-135)
// ---------------------------------------------------------------------------
FILE* g_log = nullptr;

void tee(const char* fmt, ...) {
    char buf[4096];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::fputs(buf, stdout);
    std::fflush(stdout);
    if (g_log) {
        std::fputs(buf, g_log);
        std::fflush(g_log);
    }
}

[[noreturn]] void panic(const std::string& msg) {
    std::fprintf(stderr, "panic: %s\n", msg.c_str());
    std::exit(101);
}

void check(int rc, const char* what) {
    if (rc != 0) panic(std::string(what) + ": " + tx_last_error());
}

std::string env(const char* k, const char* dflt) {
    const char* v = std::getenv(k);
    return v && *v ? v : dflt;
}

// ---------------------------------------------------------------------------
// TxConfig fields by name, for the job table and the {:?} print
// ---------------------------------------------------------------------------
struct Field {
    const char* name;
    char kind;  // 'i' int32, 'b' bool(int32), 'f' float, 'u' uint64
    size_t off;
};

#define FI(n) {#n, 'i', offsetof(tx_config, n)}
#define FB(n) {#n, 'b', offsetof(tx_config, n)}
#define FF(n) {#n, 'f', offsetof(tx_config, n)}
const Field kFields[] = {
    FI(n_feat), FI(n_epochs), {"seed", 'u', offsetof(tx_config, seed)}, FB(shuffle_users), FI(n_time_bins),
    FF(beta), FI(n_freq_bins), FF(lr_u), FF(lr_ud), FF(lr_u2), FF(lr_ub), FF(lr_ubd), FF(lr_i), FF(lr_ib),
    FF(lr_y), FF(lr_yb), FF(lr_yd), FF(lr_tu), FF(lr_ti), FF(lr_ta), FF(lr_ibf), FF(lr_iqf), FF(lr_cu),
    FF(reg_iqf), FF(reg_cu), FF(reg_u), FF(reg_u2), FF(reg_ud), FF(reg_i), FF(reg_y), FF(reg_yd),
    FF(sigma_iqf), FF(sigma_u), FF(sigma_i), FF(sigma_y), FF(sigma_yd), FI(reset_u_epoch), FI(max_neighbors),
    FF(lr_w), FF(lr_c), FF(reg_w), FF(reg_c), FF(lr_w_day), FF(lr_c_day), FF(reg_w_day), FF(reg_c_day),
    FF(w_bias), FF(w_factor), FF(w_nbr), FB(sum_err_bug), FF(lambda1), FF(lambda2),
};
#undef FI
#undef FB
#undef FF

const Field* field(const std::string& n) {
    for (const Field& f : kFields)
        if (n == f.name) return &f;
    panic("unknown TxConfig field " + n);
}

// Rust's shortest round-trip float formatting as {:?} prints it.
std::string rust_f32(float v) {
    if (v == 0.0f) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    for (int p = 1; p <= 9; ++p) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, static_cast<double>(v));
        if (std::strtof(buf, nullptr) == v) break;
    }
    // buf = d.ddde[+-]XX
    std::string s(buf);
    const size_t e = s.find('e');
    std::string mant = s.substr(0, e);
    const int ex = std::atoi(s.c_str() + e + 1);
    const bool neg = mant[0] == '-';
    if (neg) mant = mant.substr(1);
    std::string digits;
    for (char ch : mant)
        if (ch != '.') digits += ch;
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    std::string out;
    if (ex < -4 || ex >= 16) {
        out = digits.substr(0, 1) + (digits.size() > 1 ? "." + digits.substr(1) : "") + "e" + std::to_string(ex);
    } else if (ex < 0) {
        out = "0." + std::string(-ex - 1, '0') + digits;
    } else {
        if (static_cast<int>(digits.size()) <= ex + 1) out = digits + std::string(ex + 1 - digits.size(), '0') + ".0";
        else out = digits.substr(0, ex + 1) + "." + digits.substr(ex + 1);
    }
    return (neg ? "-" : "") + out;
}

std::string debug_config(const tx_config& c) {
    std::string s = "TxConfig { ";
    const char* base = reinterpret_cast<const char*>(&c);
    for (const Field& f : kFields) {
        s += f.name;
        s += ": ";
        if (f.kind == 'i') s += std::to_string(*reinterpret_cast<const int32_t*>(base + f.off));
        else if (f.kind == 'b') s += *reinterpret_cast<const int32_t*>(base + f.off) ? "true" : "false";
        else if (f.kind == 'u') s += std::to_string(*reinterpret_cast<const uint64_t*>(base + f.off));
        else s += rust_f32(*reinterpret_cast<const float*>(base + f.off));
        s += ", ";
    }
    if (c.has_ordinal)
        s += "ordinal_head: Some(OrdinalHeadConfig { th_init: [" + rust_f32(c.th_init[0]) + ", " + rust_f32(c.th_init[1]) +
             ", " + rust_f32(c.th_init[2]) + ", " + rust_f32(c.th_init[3]) + "], th_gap: " + rust_f32(c.th_gap) +
             ", lr_t: " + rust_f32(c.lr_t) + ", reg_t: " + rust_f32(c.reg_t) + " }), ";
    else
        s += "ordinal_head: None, ";
    s += std::string("save_ifeat: ") + (c.save_ifeat ? "true" : "false") + ", low_memory: " +
         (c.low_memory ? "true" : "false") + ", full_su: " + (c.full_su ? "true" : "false") + " }";
    return s;
}

// Debug rendering of Tsvdx4pConfig (src/tsvdx4p.rs:128-186): its own field
// order, the TxConfig fields it shares, then the three tsvdx4p extras.
bool g_tsvdx4p = false;  // the job trains a Tsvdx4pModel (tsvdx4p-new jobs)

std::string debug_config_p(const tx_config& c) {
    static const char* kOrder[] = {"n_feat", "n_epochs", "seed", "shuffle_users", "n_time_bins", "beta",
        "n_freq_bins", "lr_u", "lr_ud", "lr_u2", "lr_ub", "lr_ubd", "lr_i", "lr_ib", "lr_y", "lr_yb", "lr_yd",
        "lr_tu", "lr_ti", "lr_ta", "lr_ibf", "lr_iqf", "reg_iqf", "sigma_iqf", "lr_cu", "reg_cu", "reg_u",
        "reg_u2", "reg_ud", "reg_i", "reg_y", "reg_yd", "sigma_u", "sigma_i", "sigma_y", "sigma_yd",
        "reset_u_epoch"};
    std::string s = "Tsvdx4pConfig { ";
    const char* base = reinterpret_cast<const char*>(&c);
    for (const char* name : kOrder) {
        const Field& f = *field(name);
        s += f.name;
        s += ": ";
        if (f.kind == 'i') s += std::to_string(*reinterpret_cast<const int32_t*>(base + f.off));
        else if (f.kind == 'b') s += *reinterpret_cast<const int32_t*>(base + f.off) ? "true" : "false";
        else if (f.kind == 'u') s += std::to_string(*reinterpret_cast<const uint64_t*>(base + f.off));
        else s += rust_f32(*reinterpret_cast<const float*>(base + f.off));
        s += ", ";
    }
    s += std::string("save_ifeat: ") + (c.save_ifeat ? "true" : "false") + ", low_memory: " +
         (c.low_memory ? "true" : "false") + ", full_su: " + (c.full_su ? "true" : "false") +
         ", nsvd_norm_exp: " + rust_f32(c.nsvd_norm_exp) + ", n_threads: " + std::to_string(c.n_threads) +
         ", seq_epochs: [";
    for (int k = 0; k < 5; ++k) s += std::to_string(c.seq_epochs[k]) + (k < 4 ? ", " : "");
    return s + "] }";
}

tx_config make_config(std::initializer_list<std::pair<const char*, double>> kv) {
    tx_config c{};
    c.w_bias = c.w_factor = c.w_nbr = 1.0f;
    c.shuffle_users = 1;
    c.full_su = 1;
    char* base = reinterpret_cast<char*>(&c);
    for (auto& p : kv) {
        const Field* f = field(p.first);
        if (f->kind == 'i' || f->kind == 'b') *reinterpret_cast<int32_t*>(base + f->off) = static_cast<int32_t>(p.second);
        else if (f->kind == 'u') *reinterpret_cast<uint64_t*>(base + f->off) = static_cast<uint64_t>(p.second);
        else *reinterpret_cast<float*>(base + f->off) = static_cast<float>(p.second);
    }
    return c;
}

void ordinal_std(tx_config& c) {
    c.has_ordinal = 1;
    const float th[4] = {0.5f, 1.25f, 3.75f, 5.5f};
    std::memcpy(c.th_init, th, sizeof th);
    c.th_gap = 0.001f;
    c.lr_t = 0.0f;
    c.reg_t = 0.0f;
}

// ---------------------------------------------------------------------------
// Dataset (src/lib.rs:227-402)
// ---------------------------------------------------------------------------
struct Dataset {
    std::vector<int32_t> user_idxs, user_cnts, item_idxs, item_cnts, item_years;
    std::vector<int8_t> raw_ratings, is_test;
    std::vector<float> residuals;
    std::vector<int16_t> dates;
    std::string name;
    int32_t n_users = 0, n_items = 0;
    bool transposed = false;

    int64_t n_ratings() const { return static_cast<int64_t>(user_idxs.size()); }
    bool contains_test_examples() const {
        return std::any_of(is_test.begin(), is_test.end(), [](int8_t v) { return v != 0; });
    }
    tx_ratings view(bool masked) const {
        tx_ratings r{};
        r.n_ratings = n_ratings();
        r.n_users = n_users;
        r.n_items = n_items;
        r.user_idxs = user_idxs.data();
        r.item_idxs = item_idxs.data();
        r.dates = dates.data();
        r.residuals = masked ? nullptr : residuals.data();
        r.raw_ratings = masked ? nullptr : raw_ratings.data();
        r.is_test = is_test.data();
        r.user_cnts = user_cnts.data();
        return r;
    }
};

std::string trim(std::string s) {
    while (!s.empty() && s.front() == ' ') s.erase(s.begin());
    while (!s.empty() && s.back() == ' ') s.pop_back();
    return s;
}

Dataset load_dataset(const std::string& name, const std::string& target, const std::string& preds_dir) {
    const std::string dir = env("TSVD_DATA_DIR", "data") + "/" + name + "/";
    auto rd = [&](const char* f) { return dir + f; };
    Dataset d;
    try {
        d.user_idxs = npy::load<int32_t>(rd("user_idxs.npy"));
        d.user_cnts = npy::load<int32_t>(rd("user_cnts.npy"));
        d.item_idxs = npy::load<int32_t>(rd("item_idxs.npy"));
        d.item_cnts = npy::load<int32_t>(rd("item_cnts.npy"));
        d.item_years = npy::load<int32_t>(rd("item_years.npy"));
        d.raw_ratings = npy::load<int8_t>(rd("ratings.npy"));
        d.dates = npy::load<int16_t>(rd("dates.npy"));
        d.is_test = npy::load<int8_t>(rd("is_test.npy"));
    } catch (const std::exception& e) {
        panic(e.what());
    }
    d.residuals.resize(d.raw_ratings.size());
    for (size_t k = 0; k < d.raw_ratings.size(); ++k) d.residuals[k] = static_cast<float>(d.raw_ratings[k]);
    if (target != "rtg") {
        // "w*model + w*model": residuals = ratings - sum w * preds (src/lib.rs:262-274)
        size_t p = 0;
        while (p <= target.size()) {
            size_t q = target.find(" + ", p);
            const std::string term = trim(target.substr(p, q == std::string::npos ? std::string::npos : q - p));
            float w = 1.0f;
            std::string model = term;
            const size_t star = term.find('*');
            if (star != std::string::npos) {
                w = std::stof(trim(term.substr(0, star)));
                model = trim(term.substr(star + 1));
            }
            const std::string path = preds_dir + "/" + model + "." + name + ".npy";
            std::vector<float> preds;
            try {
                preds = npy::load<float>(path);
            } catch (const std::exception& e) {
                panic(e.what());
            }
            if (preds.size() != d.residuals.size()) panic(path + ": length differs from the dataset");
            for (size_t k = 0; k < preds.size(); ++k) d.residuals[k] = d.residuals[k] - preds[k] * w;
            if (q == std::string::npos) break;
            p = q + 3;
        }
    }
    d.name = name;
    d.n_users = std::atoi(env("TSVD_N_USERS", "480189").c_str());
    d.n_items = std::atoi(env("TSVD_N_ITEMS", "17770").c_str());
    return d;
}

// Dataset::transpose (src/lib.rs:295-402): swap user and item roles; with
// reorder the rows are stably sorted by (new user, date).
Dataset transpose(const Dataset& s, bool reorder) {
    Dataset t;
    const int64_t n = s.n_ratings();
    std::vector<int64_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    if (reorder) {
        int32_t max_day = 0;
        for (int16_t d : s.dates) max_day = std::max<int32_t>(max_day, d);
        const uint32_t n_days = static_cast<uint32_t>(max_day) + 1;
        std::vector<uint32_t> key(n);
        for (int64_t k = 0; k < n; ++k) key[k] = static_cast<uint32_t>(s.item_idxs[k]) * n_days + static_cast<uint32_t>(s.dates[k]);
        std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    }
    t.user_idxs.resize(n); t.item_idxs.resize(n); t.raw_ratings.resize(n); t.residuals.resize(n);
    t.dates.resize(n); t.is_test.resize(n);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t o = ord[k];
        t.user_idxs[k] = s.item_idxs[o];
        t.item_idxs[k] = s.user_idxs[o];
        t.raw_ratings[k] = s.raw_ratings[o];
        t.residuals[k] = s.residuals[o];
        t.dates[k] = s.dates[o];
        t.is_test[k] = s.is_test[o];
    }
    t.user_cnts = s.item_cnts;
    t.item_cnts = s.user_cnts;
    t.item_years.assign(s.n_users, 0);
    t.name = s.name;
    t.n_users = s.n_items;
    t.n_items = s.n_users;
    t.transposed = !s.transposed;
    return t;
}

// ---------------------------------------------------------------------------
// model wrapper and evaluation helpers
// ---------------------------------------------------------------------------
struct Model {
    tx_ctx* ctx = nullptr;
    tx_config cfg;
    int32_t n_items = 0;
    Model(const Dataset& tr, const Dataset& pr, const tx_config& c) : cfg(c), n_items(tr.n_items) {
        const tx_ratings t = tr.view(false), p = pr.view(true);
        check(tx_create(&cfg, &t, &p, &ctx), "TxModel::new");
        // the library's default schedule (Tsvdx4pModel epochs for tsvdx4p
        // jobs), epochs captured in a graph, mode and strata from the environment
        const std::string mode = env("TSVD_SCHEDULE", "");
        tx_schedule s{TX_SCHED_STRATA, 1, 0, 0, 1, TX_UPD_ATOMIC | (TX_UPD_PLAIN << 2), 2, 1024};
        if (g_tsvdx4p) s.mode = TX_SCHED_TSVDX4P;
        if (mode == "serial") s.mode = TX_SCHED_SERIAL;
        else if (mode == "strata") s.mode = TX_SCHED_STRATA;
        s.n_strata = std::atoi(env("TSVD_STRATA", "1").c_str());
        if (s.n_strata > 1) s.update_flags |= TX_UPD_STRATA_EXPERIMENTAL;  // asked for by name
        int rc = tx_set_schedule(ctx, &s);
        if (rc == TX_EUNSUPPORTED && mode.empty() && !cfg.low_memory) {
            s.mode = TX_SCHED_SERIAL;  // day terms on ratings not sorted by date
            rc = tx_set_schedule(ctx, &s);
        }
        check(rc, "tx_set_schedule");
    }
    ~Model() { tx_destroy(ctx); }
    void fit_epoch(int epoch) { check(tx_fit_epoch(ctx, epoch), "fit_epoch"); }
};

std::vector<float> predict_all(Model& m, const Dataset& ds) {
    std::vector<float> out(ds.n_ratings());
    check(tx_predict(m.ctx, ds.n_ratings(), ds.user_idxs.data(), ds.item_idxs.data(), ds.dates.data(), out.data()),
          "predict");
    return out;
}

// gen_preds_parallel (src/lib.rs:710-722): predict + (raw - residual)
std::vector<float> gen_preds(Model& m, const Dataset& ds) {
    std::vector<float> p = predict_all(m, ds);
    for (int64_t k = 0; k < ds.n_ratings(); ++k)
        p[k] = p[k] + (static_cast<float>(ds.raw_ratings[k]) - ds.residuals[k]);
    return p;
}

double calc_rmse(Model& m, const Dataset& ds) {
    const tx_ratings r = ds.view(false);
    double out = 0.0;
    check(tx_calc_rmse(m.ctx, &r, &out), "calc_rmse");
    return out;
}

// rmse_from_preds (src/lib.rs:725-735)
double rmse_from_preds(const std::vector<float>& preds, const Dataset& ds) {
    double sse = 0.0, cnt = 0.0;
    for (int64_t k = 0; k < ds.n_ratings(); ++k) {
        if (ds.is_test[k] != 0) continue;
        const double e = static_cast<double>(preds[k] - static_cast<float>(ds.raw_ratings[k]));
        sse += e * e;
        cnt += 1.0;
    }
    return std::sqrt(sse / cnt);
}

void report_rmse(Model& m, int epoch, int n_epochs, const Dataset& ds) {
    tee("Epoch %02d", epoch);
    if (ds.contains_test_examples() && epoch < n_epochs) {
        tee("\n");
        return;
    }
    tee(" | RMSE %.5f\n", calc_rmse(m, ds));
}

void write_preds(const std::string& path, const std::vector<float>& p) {
    try {
        npy::write_f32(path, p.data(), {static_cast<int64_t>(p.size())});
    } catch (const std::exception& e) {
        panic(e.what());
    }
}

void save_subscores(Model& m, const Dataset& ds, const std::string& preds_dir, const std::string& job,
                    const std::string& set) {
    const int64_t n = ds.n_ratings();
    int32_t k = 0;
    check(tx_n_subscores(m.ctx, &k), "n_subscores");
    std::vector<float> b(n), mf(n), ns(n), nb(k > 3 ? n : 0);
    check(tx_predict_subscores(m.ctx, n, ds.user_idxs.data(), ds.item_idxs.data(), ds.dates.data(), b.data(), mf.data(),
                               ns.data(), k > 3 ? nb.data() : nullptr),
          "predict_subscores");
    // subscore_names (src/tx.rs:655-661): bias, mf, nsvd1[, nbr]
    write_preds(preds_dir + "/" + job + "-bias." + set + ".npy", b);
    write_preds(preds_dir + "/" + job + "-mf." + set + ".npy", mf);
    write_preds(preds_dir + "/" + job + "-nsvd1." + set + ".npy", ns);
    if (k > 3) write_preds(preds_dir + "/" + job + "-nbr." + set + ".npy", nb);
}

// save_artifacts (src/tx.rs:726-732)
void save_artifacts(Model& m, const std::string& job, const std::string& tr_set, const std::string& preds_dir) {
    if (!m.cfg.save_ifeat) return;
    int64_t n = 0;
    check(tx_param_len(m.ctx, TX_P_IFEAT, &n), "param_len");
    std::vector<float> q(n);
    check(tx_get_param(m.ctx, TX_P_IFEAT, q.data(), n), "get_param");
    const std::string path = preds_dir + "/" + job + ".ifeat." + tr_set + ".npy";
    try {
        npy::write_f32(path, q.data(), {m.n_items, m.cfg.n_feat});
    } catch (const std::exception& e) {
        panic(e.what());
    }
    tee("Saved item factors to %s\n", path.c_str());
}

struct Split {
    const char *tr, *pr, *fulltrain_tr, *fulltrain_pr, *preds_dir;
};
const Split SPLIT_NEW = {"trainx", "probex", "fulltrain", "qual", "preds_new"};

struct Opts {
    bool save_train = false, save_probe_each_epoch = false, save_subscores = false, no_fulltrain = false,
         transpose = false;
};

// fit (src/lib.rs:764-815)
void fit(const tx_config& cfg, const std::string& target, const std::string& tr_set, const std::string& pr_set,
         const std::string& job, const Opts& o, const std::string& preds_dir) {
    tee("%s => %s%s\n", tr_set.c_str(), pr_set.c_str(), o.transpose ? " (transposed)" : "");
    Dataset tr = load_dataset(tr_set, target, preds_dir);
    Dataset pr = load_dataset(pr_set, target, preds_dir);
    if (o.transpose) {
        tr = transpose(tr, true);
        pr = transpose(pr, false);
    }
    Model m(tr, pr, cfg);
    const int n_epochs = cfg.n_epochs;
    for (int e = 1; e <= n_epochs; ++e) {
        m.fit_epoch(e);
        report_rmse(m, e, n_epochs, pr);
        if (o.save_probe_each_epoch) {
            char p[1024];
            std::snprintf(p, sizeof p, "%s/%s_ep%02d.%s.npy", preds_dir.c_str(), job.c_str(), e, pr_set.c_str());
            write_preds(p, gen_preds(m, pr));
        }
    }
    write_preds(preds_dir + "/" + job + "." + pr_set + ".npy", gen_preds(m, pr));
    if (o.save_subscores) save_subscores(m, pr, preds_dir, job, pr_set);
    if (o.save_train) write_preds(preds_dir + "/" + job + "." + tr_set + ".npy", gen_preds(m, tr));
    save_artifacts(m, job, tr_set, preds_dir);
}

// fit2_inner (src/lib.rs:881-996)
void fit2(const tx_config& cfg, const std::string& target, const std::string& job, const Split& sp, const Opts& o) {
    const std::string preds_dir = sp.preds_dir;
    mkdir(preds_dir.c_str(), 0755);
    const bool owns_log = g_log == nullptr;
    if (owns_log) {
        g_log = std::fopen((preds_dir + "/" + job + ".out").c_str(), "w");
        if (!g_log) panic(preds_dir + "/" + job + ".out: cannot create");
    }
    tee("[%s]\n", job.c_str());
    tee("target = \"%s\"\n", target.c_str());
    tee("%s\n", (g_tsvdx4p ? debug_config_p(cfg) : debug_config(cfg)).c_str());
    {
        tee("%s => %s%s\n", sp.tr, sp.pr, o.transpose ? " (transposed)" : "");
        Dataset tr = load_dataset(sp.tr, target, preds_dir);
        Dataset pr = load_dataset(sp.pr, target, preds_dir);
        std::unique_ptr<Dataset> qual;
        if (o.no_fulltrain) qual = std::make_unique<Dataset>(load_dataset(sp.fulltrain_pr, target, preds_dir));
        if (o.transpose) {
            tr = transpose(tr, true);
            pr = transpose(pr, false);
            if (qual) *qual = transpose(*qual, false);
        }
        Model m(tr, pr, cfg);
        const int n_epochs = cfg.n_epochs;
        for (int e = 1; e <= n_epochs; ++e) {
            m.fit_epoch(e);
            report_rmse(m, e, n_epochs, pr);
            if (o.save_probe_each_epoch) {
                char p[1024];
                std::snprintf(p, sizeof p, "%s/%s_ep%02d.%s.npy", preds_dir.c_str(), job.c_str(), e, sp.pr);
                write_preds(p, gen_preds(m, pr));
                if (o.no_fulltrain) {
                    std::snprintf(p, sizeof p, "%s/%s_ep%02d.%s.npy", preds_dir.c_str(), job.c_str(), e, sp.fulltrain_pr);
                    write_preds(p, gen_preds(m, *qual));
                }
            }
        }
        const std::vector<float> pr_preds = gen_preds(m, pr);
        tee("%s RMSE %.5f\n", sp.pr, rmse_from_preds(pr_preds, pr));
        write_preds(preds_dir + "/" + job + "." + sp.pr + ".npy", pr_preds);
        if (o.save_subscores) save_subscores(m, pr, preds_dir, job, sp.pr);
        if (o.save_train) write_preds(preds_dir + "/" + job + "." + sp.tr + ".npy", gen_preds(m, tr));
        save_artifacts(m, job, sp.tr, preds_dir);
        if (o.no_fulltrain) {
            tee("%s => %s (no_fulltrain)\n", sp.tr, sp.fulltrain_pr);
            write_preds(preds_dir + "/" + job + "." + sp.fulltrain_pr + ".npy", gen_preds(m, *qual));
            if (o.save_subscores) save_subscores(m, *qual, preds_dir, job, sp.fulltrain_pr);
        }
    }
    if (!o.no_fulltrain) fit(cfg, target, sp.fulltrain_tr, sp.fulltrain_pr, job, o, preds_dir);
    tee("\n");
    if (owns_log) {
        std::fclose(g_log);
        g_log = nullptr;
    }
}

// kNN chain jobs: fit2 of a model without epochs (Regressor::n_epochs() = 0),
// subscores or artifacts, on the base job's residuals.
//   `<base>__knnf` (run_knnf, src/bin/tsvdx4-new.rs:55-57, 576-588 and
//   src/bin/tsvdx4p-new.rs:24, 90-105): KnnfModel with the default KnnfConfig
//   over the base job's saved item factors `{preds}/{base}.ifeat.{train set}.npy`,
//   target `1.0*{base}`.
//   `<base>__knns` (run_knns, src/bin/tsvdx4-new.rs:49-52, 563-574): KnnsModel
//   with SuppSource::Compute and shrinkage 100, target `{w}*{base}`.
//   `<base>__knn3` (run_knn3, src/bin/tsvdx4-new.rs:43-46, 557-561): Knn3Model
//   with Knn3Config::default, target `1.0*{base}`.
//   `tsvdx4-4800lm__knn-d` (run_knn_d, src/bin/tsvdx4-new.rs:60, 590-610):
//   KnnModel, support similarity over sim/rtg_supp.{set}.npy, averaged value
//   from sim/{base}_diff1.{set}.npy, target `1.0*{base}`.
// One fit phase: knn-d when kd is given, knn3 when k3 is given, knns when ks
// is given, else knnf.
void knn_fit(const std::string& base, const std::string& target, const std::string& tr_set,
             const std::string& pr_set, const std::string& job, const std::string& preds_dir, bool phase1,
             const knns_config* ks, const knn3_config* k3 = nullptr, const knnd_config* kd = nullptr) {
    tee("%s => %s\n", tr_set.c_str(), pr_set.c_str());
    Dataset tr = load_dataset(tr_set, target, preds_dir);
    Dataset pr = load_dataset(pr_set, target, preds_dir);
    knnf_ctx* ctx = nullptr;
    const tx_ratings t = tr.view(false);
    if (kd) {
        // KnnModel::new (src/knn.rs:191-199): stat0 and stat1 from {sim_dir}/{name}.{train set}.npy
        auto stat = [&](const std::string& name) {
            const std::string path = std::string("sim/") + name + "." + tr_set + ".npy";
            npy::Array a;
            try {
                a = npy::read(path);
            } catch (const std::exception& e) {
                panic(e.what());
            }
            if (a.shape.size() != 2 || a.shape[0] != a.shape[1]) panic(path + ": expected a square item x item table");
            return std::make_pair(npy::as<float>(a, path), static_cast<int32_t>(a.shape[0]));
        };
        const auto s0 = stat("rtg_supp");
        const auto s1 = stat(base + "_diff1");
        if (s0.second != s1.second) panic("stat tables of different sizes");
        check(knnd_create(kd, s0.first.data(), s1.first.data(), s0.second, &t, &ctx), "KnnModel::new");
    } else if (k3) {
        // Knn3Model::new (src/knn3.rs:350-362) and its progress lines (:292-340)
        tee("  KNN3: computing baseline...\n");
        tee("  KNN3: building inverted index...\n");
        tee("  KNN3: computing per-item weights (%d items, parallel)...\n", tr.n_items);
        knn3_ctx* c3 = nullptr;
        check(knn3_create(k3, &t, &c3), "Knn3Model::new");
        std::vector<int32_t> cnt(tr.n_items), it(static_cast<size_t>(tr.n_items) * 64);
        std::vector<float> w(it.size());
        check(knn3_weights(c3, cnt.data(), it.data(), w.data()), "Knn3Model::new");
        int64_t total = 0, nonzero = 0;
        for (int32_t c : cnt) {
            total += c;
            nonzero += c > 0;
        }
        tee("  KNN3: %lld items with neighbors, %lld total weights\n", static_cast<long long>(nonzero),
            static_cast<long long>(total));
        std::vector<float> p(pr.n_ratings());
        check(knn3_predict(c3, pr.n_ratings(), pr.user_idxs.data(), pr.item_idxs.data(), pr.dates.data(), p.data()),
              "predict");
        knn3_destroy(c3);
        for (int64_t k = 0; k < pr.n_ratings(); ++k)
            p[k] = p[k] + (static_cast<float>(pr.raw_ratings[k]) - pr.residuals[k]);
        if (phase1) tee("%s RMSE %.5f\n", pr_set.c_str(), rmse_from_preds(p, pr));
        write_preds(preds_dir + "/" + job + "." + pr_set + ".npy", p);
        return;
    } else if (ks) {
        const tx_ratings h = pr.view(true);
        check(knns_create(ks, &t, &h, &ctx), "KnnsModel::new");
    } else {
        const std::string fpath = preds_dir + "/" + base + ".ifeat." + tr_set + ".npy";
        npy::Array fa;
        try {
            fa = npy::read(fpath);
        } catch (const std::exception& e) {
            panic(e.what());
        }
        if (fa.shape.size() != 2) panic(fpath + ": expected a 2-D factor table");
        const std::vector<float> F = npy::as<float>(fa, fpath);
        const knnf_config kc{25, 1.0f, 0.025f};  // KnnfConfig::with_factors (src/knnf.rs:23-26)
        check(knnf_create(&kc, F.data(), static_cast<int32_t>(fa.shape[0]), static_cast<int32_t>(fa.shape[1]), &t,
                          &ctx),
              "KnnfModel::new");
    }
    std::vector<float> p(pr.n_ratings());
    check(knnf_predict(ctx, pr.n_ratings(), pr.user_idxs.data(), pr.item_idxs.data(), pr.dates.data(), p.data()),
          "predict");
    knnf_destroy(ctx);
    for (int64_t k = 0; k < pr.n_ratings(); ++k) p[k] = p[k] + (static_cast<float>(pr.raw_ratings[k]) - pr.residuals[k]);
    if (phase1) tee("%s RMSE %.5f\n", pr_set.c_str(), rmse_from_preds(p, pr));
    write_preds(preds_dir + "/" + job + "." + pr_set + ".npy", p);
}

void fit2_knn(const std::string& job, const std::string& base, const std::string& target, const Split& sp,
              const knns_config* ks, const knn3_config* k3 = nullptr, const knnd_config* kd = nullptr) {
    const std::string preds_dir = sp.preds_dir;
    mkdir(preds_dir.c_str(), 0755);
    g_log = std::fopen((preds_dir + "/" + job + ".out").c_str(), "w");
    if (!g_log) panic(preds_dir + "/" + job + ".out: cannot create");
    tee("[%s]\n", job.c_str());
    tee("target = \"%s\"\n", target.c_str());
    if (kd)
        tee("KnnConfig { stat0: Some(\"rtg_supp\"), stat1: Some(\"%s_diff1\"), stat2: None, factors: None, "
            "sim_type: Support, k: %d, sim_threshold: None, shrinkage: %s, scaling: %s, tau: %s, use_stat1: true, "
            "regression: false, regression_lambda: 0.1, sim_dir: \"sim\", preds_dir: \"%s\" }\n",
            base.c_str(), kd->k, rust_f32(kd->shrinkage).c_str(), rust_f32(kd->scaling).c_str(),
            rust_f32(kd->tau).c_str(), preds_dir.c_str());
    else if (k3)
        tee("Knn3Config { threshold: %s, k_min: %d, k_max: %d, shrinkage: %s, reg: %s, x: %s, bl_reg_m: %s, "
            "bl_reg_u: %s }\n",
            rust_f32(k3->threshold).c_str(), k3->k_min, k3->k_max, rust_f32(k3->shrinkage).c_str(),
            rust_f32(k3->reg).c_str(), rust_f32(k3->x).c_str(), rust_f32(k3->bl_reg_m).c_str(),
            rust_f32(k3->bl_reg_u).c_str());
    else if (ks)
        tee("KnnsConfig { k: %d, shrinkage: %s, scaling: %s, tau: %s, supp_source: Compute }\n", ks->k,
            rust_f32(ks->shrinkage).c_str(), rust_f32(ks->scaling).c_str(), rust_f32(ks->tau).c_str());
    else
        tee("KnnfConfig { factors: \"%s/%s.ifeat\", k: 25, scaling: 1.0, tau: 0.025 }\n", preds_dir.c_str(),
            base.c_str());
    knn_fit(base, target, sp.tr, sp.pr, job, preds_dir, true, ks, k3, kd);
    knn_fit(base, target, sp.fulltrain_tr, sp.fulltrain_pr, job, preds_dir, false, ks, k3, kd);
    tee("\n");
    std::fclose(g_log);
    g_log = nullptr;
}

// epoch_blend_apply (src/lib.rs:1181-1235)
void epoch_blend_apply(const std::string& model, const std::vector<double>& w, double offset, const Split& sp) {
    const std::string preds_dir = sp.preds_dir;
    mkdir(preds_dir.c_str(), 0755);
    const std::string name = model + "__epochs";
    const bool owns_log = g_log == nullptr;
    if (owns_log) g_log = std::fopen((preds_dir + "/" + name + ".out").c_str(), "w");
    tee("[%s]\n", name.c_str());
    tee("eblend legacy mode: %zu epoch weights, constant offset=%+.6f\n", w.size(), offset);
    for (size_t i = 0; i < w.size(); ++i) tee("  epoch %2zu weight = %+.6f\n", i + 1, w[i]);
    for (const char* ds : {sp.pr, sp.fulltrain_pr}) {
        std::vector<std::vector<float>> ep;
        for (size_t e = 1; e <= w.size(); ++e) {
            char p[1024];
            std::snprintf(p, sizeof p, "%s/%s_ep%02zu.%s.npy", preds_dir.c_str(), model.c_str(), e, ds);
            try {
                ep.push_back(npy::load<float>(p));
            } catch (const std::exception& ex) {
                panic(ex.what());
            }
        }
        const size_t n = ep[0].size();
        std::vector<float> blend(n);
        for (size_t k = 0; k < n; ++k) {
            double pred = offset;
            for (size_t i = 0; i < w.size(); ++i) pred += w[i] * static_cast<double>(ep[i][k]);
            blend[k] = static_cast<float>(pred);
        }
        const std::string out = preds_dir + "/" + name + "." + ds + ".npy";
        write_preds(out, blend);
        tee("eblend saved %s\n", out.c_str());
    }
    tee("\n");
    if (owns_log && g_log) {
        std::fclose(g_log);
        g_log = nullptr;
    }
}

// ---------------------------------------------------------------------------
// job table (src/bin/tsvdx4-new.rs:16-555): hyper-parameters per job
// ---------------------------------------------------------------------------
tx_config lm(int n_feat, int n_epochs, uint64_t seed, int n_time_bins, bool save_ifeat, bool low_memory) {
    tx_config c = make_config({
        {"n_feat", n_feat}, {"n_epochs", n_epochs}, {"seed", static_cast<double>(seed)}, {"n_time_bins", n_time_bins},
        {"beta", 0.3}, {"n_freq_bins", 16}, {"lr_u", 0.003}, {"lr_ud", 0.00125}, {"lr_u2", 7e-6}, {"lr_ub", 0.0031},
        {"lr_ubd", 0.003}, {"lr_i", 0.0036}, {"lr_ib", 0.0036}, {"lr_y", 0.0005}, {"lr_yb", 2.5e-5},
        {"lr_yd", 0.000267}, {"lr_tu", 0.0}, {"lr_ti", 0.000225}, {"lr_ta", 2.25e-5}, {"lr_ibf", 5e-5},
        {"lr_iqf", 5e-6}, {"lr_cu", 0.002}, {"reg_iqf", 0.007}, {"reg_cu", 0.01}, {"reg_u", 0.0504},
        {"reg_u2", 0.4}, {"reg_ud", 0.04}, {"reg_i", 0.00735}, {"reg_y", 0.04}, {"reg_yd", 0.02667},
        {"sigma_iqf", 0.005}, {"sigma_u", 0.0015}, {"sigma_i", 0.005}, {"sigma_y", 0.00333}, {"sigma_yd", 0.009},
        {"reset_u_epoch", 13},
    });
    c.save_ifeat = save_ifeat;
    c.low_memory = low_memory;
    return c;
}

tx_config small_o(int n_feat) {
    tx_config c = make_config({
        {"n_feat", n_feat}, {"n_epochs", 28}, {"seed", 64}, {"n_time_bins", 20}, {"beta", 0.3}, {"n_freq_bins", 16},
        {"lr_u", 0.003}, {"lr_ud", 0.0015}, {"lr_u2", 1e-5}, {"lr_ub", 0.006}, {"lr_ubd", 0.006}, {"lr_i", 0.0072},
        {"lr_ib", 0.0072}, {"lr_y", 0.0012}, {"lr_yb", 2e-6}, {"lr_yd", 0.000267}, {"lr_tu", 0.0},
        {"lr_ti", 0.0001}, {"lr_ta", 1.5e-5}, {"lr_ibf", 5e-5}, {"lr_iqf", 5e-6}, {"lr_cu", 0.002},
        {"reg_iqf", 0.007}, {"reg_cu", 0.01}, {"reg_u", 0.04}, {"reg_u2", 0.4}, {"reg_ud", 0.04}, {"reg_i", 0.007},
        {"reg_y", 0.04}, {"reg_yd", 0.04}, {"sigma_iqf", 0.005}, {"sigma_u", 0.0}, {"sigma_i", 0.005},
        {"sigma_y", 0.005}, {"sigma_yd", 0.005}, {"reset_u_epoch", 1024},
    });
    ordinal_std(c);
    c.low_memory = 0;
    return c;
}

struct Job {
    tx_config cfg;
    Opts opts;
};

bool find_job(const std::string& job, Job* out) {
    Opts o;
    auto O = [](bool train, bool sub, bool each, bool tp = false) {
        Opts x;
        x.save_train = train;
        x.save_subscores = sub;
        x.save_probe_each_epoch = each;
        x.transpose = tp;
        return x;
    };
    if (job == "tsvdx4-60") { *out = {lm(60, 20, 42, 29, false, false), O(true, false, true)}; return true; }
    if (job == "tsvdx4-400lm") { *out = {lm(400, 20, 42, 29, true, true), O(true, true, true)}; return true; }
    if (job == "tsvdx4-1600lm") { *out = {lm(1600, 2, 1600, 29, false, true), O(false, false, true)}; return true; }
    if (job == "tsvdx4-2400lm") { *out = {lm(2400, 25, 2400, 32, false, true), O(true, true, false)}; return true; }
    if (job == "tsvdx4-4800lm") { *out = {lm(4800, 27, 2400, 32, true, true), O(true, false, true)}; return true; }
    if (job == "tsvdx4-60o") { *out = {small_o(60), O(true, false, true)}; return true; }
    if (job == "tsvdx4-80o") { *out = {small_o(80), O(false, true, false)}; return true; }
    if (job == "tsvdx4-160o") { *out = {small_o(160), O(false, true, false)}; return true; }
    const double Z = 0.0;
    if (job == "tsvdx4-d0tb") {
        tx_config c = make_config({{"n_feat", 0}, {"n_epochs", 5}, {"seed", 8003}, {"n_time_bins", 29}, {"beta", 0.3},
            {"n_freq_bins", 16}, {"lr_ub", 0.012}, {"lr_ubd", 0.012}, {"lr_ib", 0.012}, {"lr_tu", 0.003},
            {"lr_ti", 0.001}, {"lr_ta", 0.0001}, {"lr_ibf", 0.0003}, {"lr_cu", 0.006}, {"reg_iqf", 0.007},
            {"reg_cu", 0.01}, {"reg_u", 0.0504}, {"reg_u2", 0.4}, {"reg_ud", 0.04}, {"reg_i", 0.00735},
            {"reg_y", 0.04}, {"reg_yd", 0.02667}, {"reset_u_epoch", 1024}, {"lr_u", Z}});
        c.low_memory = 1;
        *out = {c, O(true, false, true)};
        return true;
    }
    if (job == "tsvdx4-f1") {
        tx_config c = make_config({{"n_feat", 1}, {"n_epochs", 5}, {"seed", 8005}, {"n_time_bins", 29}, {"beta", 0.3},
            {"n_freq_bins", 16}, {"lr_u", 0.009}, {"lr_u2", 2.1e-5}, {"lr_ub", 0.0093}, {"lr_ubd", 0.009},
            {"lr_i", 0.0108}, {"lr_ib", 0.0108}, {"lr_y", 0.0015}, {"lr_yb", 7.5e-5}, {"lr_ti", 0.000675},
            {"lr_ta", 6.75e-5}, {"lr_ibf", 0.00015}, {"lr_iqf", 1.5e-5}, {"lr_cu", 0.006}, {"reg_iqf", 0.007},
            {"reg_cu", 0.01}, {"reg_u", 0.0504}, {"reg_u2", 0.4}, {"reg_ud", 0.04}, {"reg_i", 0.00735},
            {"reg_y", 0.04}, {"reg_yd", 0.02667}, {"sigma_iqf", 0.005}, {"sigma_u", 0.0015}, {"sigma_i", 0.005},
            {"sigma_y", 0.00333}, {"reset_u_epoch", 1024}});
        c.low_memory = 1;
        *out = {c, O(false, false, true)};
        return true;
    }
    if (job == "tsvdx4-h8nb" || job == "tsvdx4-l32nb" || job == "tsvdx4-n8nbo") {
        const bool l32 = job == "tsvdx4-l32nb", n8 = job == "tsvdx4-n8nbo";
        tx_config c = make_config({{"n_feat", l32 ? 32 : 8}, {"n_epochs", 5}, {"seed", l32 ? 8011 : (n8 ? 8013 : 8007)},
            {"n_time_bins", n8 ? 20 : 29}, {"beta", 0.3}, {"n_freq_bins", 16},
            {"lr_u", l32 ? 0.01 : (n8 ? 0.015 : 0.012)}, {"lr_u2", l32 ? 2.5e-5 : (n8 ? 0.0 : 3e-5)},
            {"lr_i", l32 ? 0.02 : (n8 ? 0.018 : 0.015)}, {"lr_y", l32 ? 0.002 : 0.003}, {"reg_iqf", 0.007},
            {"reg_cu", 0.01}, {"reg_u", 0.03}, {"reg_u2", 0.4}, {"reg_ud", 0.04},
            {"reg_i", l32 ? 0.004 : 0.005}, {"reg_y", 0.03}, {"reg_yd", 0.02667}, {"sigma_u", 0.003},
            {"sigma_i", 0.008}, {"sigma_y", 0.005}, {"reset_u_epoch", 1024}});
        c.low_memory = 1;
        if (n8) ordinal_std(c);
        *out = {c, n8 ? O(true, false, false) : (l32 ? O(false, false, true) : O(true, false, true))};
        return true;
    }
    if (job == "tsvdx4-m1n") {
        tx_config c = make_config({{"n_feat", 1}, {"n_epochs", 5}, {"seed", 8012}, {"n_time_bins", 29}, {"beta", 0.3},
            {"n_freq_bins", 16}, {"lr_ub", 0.015}, {"lr_ubd", 0.015}, {"lr_i", 0.0108}, {"lr_ib", 0.0108},
            {"lr_y", 0.004}, {"lr_yb", 0.0002}, {"lr_ti", 0.000675}, {"lr_ta", 6.75e-5}, {"lr_ibf", 0.00015},
            {"lr_iqf", 1.5e-5}, {"lr_cu", 0.006}, {"reg_iqf", 0.007}, {"reg_cu", 0.01}, {"reg_u", 0.0504},
            {"reg_u2", 0.4}, {"reg_ud", 0.04}, {"reg_i", 0.00735}, {"reg_y", 0.03}, {"reg_yd", 0.02667},
            {"sigma_iqf", 0.005}, {"sigma_i", 0.005}, {"sigma_y", 0.008}, {"reset_u_epoch", 1024}});
        c.low_memory = 1;
        *out = {c, O(false, false, false)};
        return true;
    }
    if (job == "tsvdx4-q16tb") {
        tx_config c = make_config({{"n_feat", 16}, {"n_epochs", 8}, {"seed", 9001}, {"n_time_bins", 48}, {"beta", 0.3},
            {"n_freq_bins", 16}, {"lr_u", 0.005}, {"lr_ud", 0.002}, {"lr_u2", 1e-5}, {"lr_ub", 0.005},
            {"lr_ubd", 0.005}, {"lr_i", 0.006}, {"lr_ib", 0.006}, {"lr_y", 0.001}, {"lr_yb", 5e-5},
            {"lr_yd", 0.0004}, {"lr_tu", 0.0003}, {"lr_ti", 0.0004}, {"lr_ta", 4e-5}, {"lr_ibf", 0.0001},
            {"lr_iqf", 1e-5}, {"lr_cu", 0.004}, {"reg_iqf", 0.007}, {"reg_cu", 0.01}, {"reg_u", 0.04},
            {"reg_u2", 0.4}, {"reg_ud", 0.03}, {"reg_i", 0.006}, {"reg_y", 0.03}, {"reg_yd", 0.02},
            {"sigma_iqf", 0.005}, {"sigma_u", 0.002}, {"sigma_i", 0.006}, {"sigma_y", 0.004}, {"sigma_yd", 0.01},
            {"reset_u_epoch", 1024}});
        c.low_memory = 1;
        *out = {c, O(true, false, true)};
        return true;
    }
    if (job == "tsvdx4-s20fo") {
        tx_config c = make_config({{"n_feat", 20}, {"n_epochs", 10}, {"seed", 9003}, {"n_time_bins", 25}, {"beta", 0.3},
            {"n_freq_bins", 24}, {"lr_u", 0.004}, {"lr_ud", 0.002}, {"lr_u2", 1.5e-5}, {"lr_ub", 0.007},
            {"lr_ubd", 0.007}, {"lr_i", 0.008}, {"lr_ib", 0.008}, {"lr_y", 0.0015}, {"lr_yb", 5e-6},
            {"lr_yd", 0.0004}, {"lr_ti", 0.00015}, {"lr_ta", 2e-5}, {"lr_ibf", 8e-5}, {"lr_iqf", 8e-6},
            {"lr_cu", 0.003}, {"reg_iqf", 0.005}, {"reg_cu", 0.008}, {"reg_u", 0.02}, {"reg_u2", 0.3},
            {"reg_ud", 0.02}, {"reg_i", 0.004}, {"reg_y", 0.02}, {"reg_yd", 0.02}, {"sigma_iqf", 0.005},
            {"sigma_i", 0.005}, {"sigma_y", 0.005}, {"sigma_yd", 0.005}, {"reset_u_epoch", 1024}});
        ordinal_std(c);
        c.low_memory = 1;
        *out = {c, O(true, false, true)};
        return true;
    }
    if (job == "tsvdx4-t8cu") {
        tx_config c = make_config({{"n_feat", 8}, {"n_epochs", 8}, {"seed", 9004}, {"n_time_bins", 29}, {"beta", 0.3},
            {"n_freq_bins", 16}, {"lr_u", 0.006}, {"lr_u2", 1.5e-5}, {"lr_ub", 0.006}, {"lr_ubd", 0.006},
            {"lr_i", 0.008}, {"lr_ib", 0.008}, {"lr_y", 0.005}, {"lr_yb", 0.0005}, {"lr_ti", 0.0003},
            {"lr_ta", 3e-5}, {"lr_ibf", 0.0001}, {"lr_iqf", 1e-5}, {"lr_cu", 0.05}, {"reg_iqf", 0.007},
            {"reg_cu", 0.003}, {"reg_u", 0.04}, {"reg_u2", 0.4}, {"reg_ud", 0.04}, {"reg_i", 0.006},
            {"reg_y", 0.02}, {"reg_yd", 0.02667}, {"sigma_iqf", 0.005}, {"sigma_u", 0.003}, {"sigma_i", 0.008},
            {"sigma_y", 0.008}, {"reset_u_epoch", 3}});
        c.low_memory = 1;
        *out = {c, O(true, false, true)};
        return true;
    }
    if (job == "tsvdx4-16-t") {
        tx_config c = make_config({{"n_feat", 16}, {"n_epochs", 4}, {"seed", 9001}, {"n_time_bins", 22}, {"beta", 0.35},
            {"n_freq_bins", 12}, {"lr_u", 0.004}, {"lr_ud", 0.0015}, {"lr_u2", 1e-5}, {"lr_ub", 0.004},
            {"lr_ubd", 0.0035}, {"lr_i", 0.0025}, {"lr_ib", 0.0028}, {"lr_y", 0.0004}, {"lr_yb", 2e-5},
            {"lr_yd", 0.0002}, {"lr_ti", 0.0002}, {"lr_ta", 2e-5}, {"lr_ibf", 4e-5}, {"lr_cu", 0.0015},
            {"reg_iqf", 0.007}, {"reg_cu", 0.008}, {"reg_u", 0.008}, {"reg_u2", 0.3}, {"reg_ud", 0.035},
            {"reg_i", 0.045}, {"reg_y", 0.035}, {"reg_yd", 0.025}, {"sigma_u", 0.006}, {"sigma_i", 0.0012},
            {"sigma_y", 0.003}, {"sigma_yd", 0.008}, {"reset_u_epoch", 1024}});
        c.save_ifeat = 1;
        c.low_memory = 1;
        *out = {c, O(true, false, true, true)};
        return true;
    }
    if (job == "tsvdx4-p0tbo") {
        tx_config c = make_config({{"n_feat", 0}, {"n_epochs", 5}, {"seed", 8014}, {"n_time_bins", 29}, {"beta", 0.3},
            {"n_freq_bins", 16}, {"lr_ub", 0.018}, {"lr_ubd", 0.018}, {"lr_ib", 0.015}, {"lr_tu", 0.003},
            {"lr_ti", 0.001}, {"lr_ta", 0.0001}, {"lr_ibf", 0.0003}, {"lr_cu", 0.006}, {"reg_iqf", 0.007},
            {"reg_cu", 0.01}, {"reg_u", 0.0504}, {"reg_u2", 0.4}, {"reg_ud", 0.04}, {"reg_i", 0.00735},
            {"reg_y", 0.04}, {"reg_yd", 0.02667}, {"reset_u_epoch", 1024}});
        ordinal_std(c);
        c.low_memory = 1;
        *out = {c, O(true, false, false)};
        return true;
    }
    (void)o;
    return false;
}

// tsvdx5-new base jobs (src/bin/tsvdx5-new.rs:21-33, run_base :130-183 and
// run_ord :185-225) and the two tsvdx6 binaries (src/bin/tsvdx6-300-new.rs,
// src/bin/tsvdx6-300-v2-new.rs): the lm set with global neighbourhood terms
// (and same-day terms for tsvdx6), sub-model weights w_bias = w_factor = w_nbr.
bool find_job_nbr(const std::string& job, Job* out) {
    struct B { const char* name; int f, epochs, max_nbr; float lr_wc; bool save_ifeat; int reset; bool lm, sub, each; };
    static const B kBase[] = {
        {"tsvdx5-20", 20, 20, 25, 0.002f, false, 1024, false, true, false},
        {"tsvdx5-40", 40, 25, 25, 0.002f, false, 1024, false, true, false},
        {"tsvdx5-60", 60, 23, 50, 0.002f, false, 1024, false, false, false},
        {"tsvdx5-120", 120, 23, 100, 0.0015f, true, 1024, false, false, false},
        {"tsvdx5-150", 150, 23, 100, 0.0015f, false, 1024, false, true, false},
        {"tsvdx5-300", 300, 23, 100, 0.0015f, false, 10, false, false, true},
        {"tsvdx5-400", 400, 26, 100, 0.0015f, false, 10, false, false, false},
        {"tsvdx5-1000", 1000, 30, 100, 0.0015f, false, 10, false, false, false},
        {"tsvdx5-1200lm", 1200, 23, 100, 0.0015f, false, 1024, true, true, true},
    };
    auto nbr = [](tx_config& c, int max_nbr, float lr_wc, float w) {
        c.max_neighbors = max_nbr;
        c.lr_w = c.lr_c = lr_wc;
        c.reg_w = c.reg_c = 0.002f;
        c.w_bias = c.w_factor = c.w_nbr = w;
        c.lambda1 = 25.0f;
        c.lambda2 = 10.0f;
    };
    for (const B& b : kBase) {
        if (job != b.name) continue;
        tx_config c = lm(b.f, b.epochs, 42, 32, b.save_ifeat, b.lm);
        c.reset_u_epoch = b.reset;
        nbr(c, b.max_nbr, b.lr_wc, 0.8f);
        Opts o;
        o.save_train = true;
        o.save_subscores = b.sub;
        o.save_probe_each_epoch = b.each;
        *out = {c, o};
        return true;
    }
    if (job == "tsvdx5-120o" || job == "tsvdx5-150o") {
        const bool f120 = job == "tsvdx5-120o";
        tx_config c = small_o(f120 ? 120 : 150);
        c.n_epochs = f120 ? 28 : 25;
        nbr(c, f120 ? 100 : 120, 0.0015f, f120 ? 0.8f : 1.2f);
        Opts o;
        o.save_train = o.save_subscores = o.save_probe_each_epoch = true;
        *out = {c, o};
        return true;
    }
    if (job == "tsvdx6-300" || job == "tsvdx6-300-v2") {
        const bool v2 = job == "tsvdx6-300-v2";
        tx_config c = lm(300, 23, 42, 32, false, false);
        c.reset_u_epoch = 10;
        nbr(c, 100, 0.0015f, 0.8f);
        c.lr_w_day = c.lr_c_day = v2 ? 0.0008f : 0.001f;
        c.reg_w_day = c.reg_c_day = 0.002f;
        if (v2) c.w_nbr = 0.7f;
        Opts o;
        o.save_train = o.save_probe_each_epoch = !v2;
        *out = {c, o};
        return true;
    }
    return false;
}

// tsvdx4p-new job table (src/bin/tsvdx4p-new.rs:12-26, run_lm :30-85): the lm
// set with lr_yb 2e-4, reset at epoch 1024, n^1.0 NSVD1 normalisation, 16
// threads, low-memory; save_train jobs write train predictions, the others
// every epoch's probe predictions.
bool find_job_p(const std::string& job, Job* out) {
    struct P { const char* name; int f; bool save_ifeat; int seq[5]; bool save_train; };
    static const P kJobs[] = {
        {"tsvdx4p-1000lm", 1000, true, {1, 10, 20, 0, 0}, true},
        {"tsvdx4p-2000lm", 2000, true, {1, 10, 20, 0, 0}, true},
        {"tsvdx4p-4000lm", 4000, false, {1, 10, 20, 0, 0}, false},
        {"tsvdx4p-6000lm", 6000, true, {0, 0, 0, 0, 0}, true},
        {"tsvdx4p-8000lm", 8000, true, {1, 10, 20, 0, 0}, true},
        {"tsvdx4p-16000lm", 16000, false, {1, 10, 20, 0, 0}, false},
    };
    for (const P& p : kJobs) {
        if (job != p.name) continue;
        tx_config c = lm(p.f, 20, 42, 29, p.save_ifeat, true);
        c.lr_yb = 0.0002f;
        c.reset_u_epoch = 1024;
        c.nsvd_norm_exp = 1.0f;
        c.n_threads = 16;
        for (int k = 0; k < 5; ++k) c.seq_epochs[k] = p.seq[k];
        Opts o;
        o.save_train = p.save_train;
        o.save_probe_each_epoch = !p.save_train;
        *out = {c, o};
        return true;
    }
    return false;
}

bool find_eblend(const std::string& job, std::vector<double>* w, double* off) {
    static const std::map<std::string, std::pair<std::vector<double>, double>> table = {
        {"tsvdx4-60__epochs",
         {{0.037902, 0.038349, 0.038870, 0.039168, 0.039484, 0.039743, 0.040036, 0.040299, 0.040508, 0.040723,
           0.040934, 0.041124, 0.039802, 0.040261, 0.040536, 0.040834, 0.041041, 0.041260, 0.041453, 0.041609},
          0.628000}},
        {"tsvdx4-60o__epochs",
         {{0.027400, 0.028007, 0.028270, 0.028537, 0.028736, 0.029028, 0.029214, 0.029395, 0.029596, 0.029754,
           0.029904, 0.030050, 0.030198, 0.030342, 0.030463, 0.030626, 0.030730, 0.030845, 0.030969, 0.031096,
           0.031185, 0.031290, 0.031392, 0.031510, 0.031593, 0.031697, 0.031757, 0.031838},
          0.487734}},
        {"tsvdx4-d0tb__epochs", {{0.127633, 0.128381, 0.128978, 0.129290, 0.129347}, 0.978328}},
        {"tsvdx4-f1__epochs", {{0.127902, 0.128805, 0.129345, 0.129645, 0.129651}, 0.972784}},
        {"tsvdx4-q16tb__epochs",
         {{0.091397, 0.092004, 0.092560, 0.093097, 0.093596, 0.094026, 0.094293, 0.094378}, 0.702795}},
        // tsvdx5-new run_eblend jobs (src/bin/tsvdx5-new.rs:104-128)
        {"tsvdx5-120o__epochs",
         {{0.027859, 0.028633, 0.029052, 0.029406, 0.029674, 0.030008, 0.030212, 0.030423, 0.030624, 0.030802,
           0.030952, 0.031114, 0.031269, 0.031435, 0.031551, 0.031699, 0.031830, 0.031948, 0.032081, 0.032203,
           0.032305, 0.032421, 0.032529, 0.032643, 0.032748, 0.032844, 0.032927, 0.033024},
          0.381012}},
        {"tsvdx5-300__epochs",
         {{0.034800, 0.035376, 0.035884, 0.036167, 0.036416, 0.036645, 0.036883, 0.037075, 0.037255, 0.036797,
           0.037004, 0.037221, 0.037410, 0.037604, 0.037767, 0.037958, 0.038098, 0.038248, 0.038387, 0.038531,
           0.038679, 0.038773, 0.038903},
          0.440369}},
        {"tsvdx4-400lm__epochs",
         {{0.039362, 0.039928, 0.040560, 0.040957, 0.041354, 0.041709, 0.042063, 0.042378, 0.042626, 0.042909,
           0.043160, 0.043400, 0.041114, 0.041567, 0.041924, 0.042314, 0.042636, 0.042944, 0.043212, 0.043466},
          0.498104}},
    };
    auto it = table.find(job);
    if (it == table.end()) return false;
    *w = it->second.first;
    *off = it->second.second;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string self = std::string(argv[0]).substr(std::string(argv[0]).find_last_of('/') + 1);
    if (argc < 2 && self != "tsvdx6-300-new" && self != "tsvdx6-300-v2-new") {
        std::fprintf(stderr, "usage: %s <job_name> [--print-config]\n", argv[0]);
        return 2;
    }
    // the tsvdx6 binaries take no job argument (src/bin/tsvdx6-300-new.rs:3)
    std::string prog = argv[0];
    prog = prog.substr(prog.find_last_of('/') + 1);
    std::string job = argc > 1 ? argv[1] : "";
    int arg0 = 1;
    if (prog == "tsvdx6-300-new" || prog == "tsvdx6-300-v2-new") {
        job = prog == "tsvdx6-300-new" ? "tsvdx6-300" : "tsvdx6-300-v2";
        arg0 = 0;
    }
    const bool print_only = argc > arg0 + 1 && std::string(argv[arg0 + 1]) == "--print-config";
    Job j;
    g_tsvdx4p = find_job_p(job, &j);
    if (g_tsvdx4p || find_job(job, &j) || find_job_nbr(job, &j)) {
        if (print_only) {
            std::printf("%s\n", (g_tsvdx4p ? debug_config_p(j.cfg) : debug_config(j.cfg)).c_str());
            return 0;
        }
        fit2(j.cfg, "rtg", job, SPLIT_NEW, j.opts);
        return 0;
    }
    std::vector<double> w;
    double off = 0.0;
    if (find_eblend(job, &w, &off)) {
        if (print_only) return 0;
        epoch_blend_apply(job.substr(0, job.size() - std::string("__epochs").size()), w, off, SPLIT_NEW);
        return 0;
    }
    // run_knnf jobs of tsvdx4-new (src/bin/tsvdx4-new.rs:55-57) and tsvdx4p-new
    // (src/bin/tsvdx4p-new.rs:21-24)
    static const char* kKnnf[] = {"tsvdx4-16-t__knnf",    "tsvdx4-400lm__knnf",   "tsvdx4-4800lm__knnf",
                                  "tsvdx4p-1000lm__knnf", "tsvdx4p-2000lm__knnf", "tsvdx4p-6000lm__knnf",
                                  "tsvdx4p-8000lm__knnf"};
    for (const char* name : kKnnf) {
        if (job != name) continue;
        if (print_only) return 0;
        const std::string base = job.substr(0, job.size() - std::string("__knnf").size());
        fit2_knn(job, base, "1.0*" + base, SPLIT_NEW, nullptr);
        return 0;
    }
    struct KnnsJob { const char* name; int k; float scaling, tau; const char* w; };
    static const KnnsJob kKnns[] = {
        {"tsvdx4-60__knns", 25, 1.0f, 0.025f, "1.0"}, {"tsvdx4-60o__knns", 50, 0.5f, 0.0f, "1.0"},
        {"tsvdx4-2400lm__knns", 15, 1.0f, 0.5f, "1.2"}, {"tsvdx4-4800lm__knns", 40, 3.0f, 0.05f, "0.9"},
    };
    for (const KnnsJob& kj : kKnns) {
        if (job != kj.name) continue;
        if (print_only) return 0;
        const std::string base = job.substr(0, job.size() - std::string("__knns").size());
        const knns_config ks{kj.k, 100.0f, kj.scaling, kj.tau};
        fit2_knn(job, base, std::string(kj.w) + "*" + base, SPLIT_NEW, &ks);
        return 0;
    }
    static const char* kKnn3[] = {"tsvdx4-d0tb__knn3", "tsvdx4-h8nb__knn3", "tsvdx4-n8nbo__knn3",
                                  "tsvdx4-p0tbo__knn3", "tsvdx5-120o__knn3"};
    for (const char* name : kKnn3) {
        if (job != name) continue;
        if (print_only) return 0;
        const std::string base = job.substr(0, job.size() - std::string("__knn3").size());
        const knn3_config k3{0.25f, 10, 60, 20000.0f, 0.01f, 0.8f, 25.0f, 10.0f};  // Knn3Config::default
        fit2_knn(job, base, "1.0*" + base, SPLIT_NEW, nullptr, &k3);
        return 0;
    }
    if (job == "tsvdx4-4800lm__knn-d") {  // run_knn_d (src/bin/tsvdx4-new.rs:590-610)
        if (print_only) return 0;
        const knnd_config kd{15, 100.0f, 3.5f, 0.0f};
        fit2_knn(job, "tsvdx4-4800lm", "1.0*tsvdx4-4800lm", SPLIT_NEW, nullptr, nullptr, &kd);
        return 0;
    }
    panic("invalid job name: " + job);
}
