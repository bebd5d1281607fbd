// Host side of libtsvd_tx.so: the C ABI declared in include/tsvd_tx.h.
//
// tx_create restates TxModel::new (src/tx.rs:460-649): it derives every static
// per-rating quantity once (SparseUD index, time bin, dev_u(t), run bounds and
// norms), draws the initial factors from the reference's RNG stream, and
// uploads one device-resident model.  tx_fit_epoch then runs the chosen GPU
// schedule plus the epoch-end ycache rebuild; predictions, RMSE and parameter
// access are batched device calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "tsvd_tx.h"
#include "tx_launch.hpp"
#include "tx_rng.hpp"
#include "tx_types.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace

// The library's last-error string, shared with the other C-ABI sources.
namespace tx {
void set_error(const std::string& msg) { g_err = msg; }
void clear_error() { g_err.clear(); }
}  // namespace tx

namespace {

#define HIPC(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(TX_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// a statement like HIPC (safe inside if / else without braces)
#define RC(x)                          \
    do {                               \
        if (int rc_ = (x)) return rc_; \
    } while (0)

// Run fn(lo, hi) over [0, n) on up to hardware_concurrency threads.
template <typename Fn>
void parallel_for(int64_t n, Fn fn) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (n < 100000 || nt == 1) {
        fn(int64_t{0}, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned k = 0; k < nt; ++k) {
        const int64_t lo = k * chunk, hi = std::min<int64_t>(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=] { fn(lo, hi); });
    }
    for (auto& t : th) t.join();
}

// Phase times of tx_create on stderr when TX_TIMING is set in the environment.
struct PhaseClock {
    bool on = std::getenv("TX_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char* what) {
        if (!on) return;
        const auto n = std::chrono::steady_clock::now();
        std::fprintf(stderr, "tx_create: %-30s %7.3f s\n", what, std::chrono::duration<double>(n - t).count());
        t = n;
    }
};

// freq_bin (src/tx.rs:87-90), evaluated with the host libm like the reference.
int freq_bin_host(uint32_t cnt, int nfb) {
    if (cnt <= 1) return 0;
    const float v = std::floor(std::log2(static_cast<float>(cnt)));
    const float cap = static_cast<float>(nfb - 1);
    return static_cast<int>(v < cap ? v : cap);
}

// dev (src/tx.rs:292-300)
float dev_host(float tu_mean, int day, float beta) {
    const float dt = static_cast<float>(day) - tu_mean;
    if (dt == 0.0f) return 0.0f;
    const float s = dt > 0.0f ? 1.0f : -1.0f;
    return s * std::pow(std::fabs(dt), beta);
}

struct Shape {
    int team, vec;
    int waves = 1;  // waves per user in the SGD kernel (f > 1024: chunks of 1024 elements)
};

// One wave (64 lanes) per user at every f: at f=60, four 16-lane teams per
// wave took 95 ms per epoch of the full-shape lm model against 69 ms with one
// element per lane (each wave ran its four users' loops to the longest).
bool pick_shape(int f, Shape* s) {
    if (f <= 64) *s = {64, 1};
    else if (f <= 128) *s = {64, 2};
    else if (f <= 256) *s = {64, 4};
    else if (f <= 512) *s = {64, 8};
    else if (f <= 1024) *s = {64, 16};
    else if (f <= 8192) *s = {64, 16, (f + 1023) / 1024};
    else return false;
    return true;
}

}  // namespace

struct tx_ctx {
    tx_config cfg{};
    std::vector<tx_autotune_trial> trials;  // the last tx_autotune's trials
    int32_t U = 0, I = 0, f = 0, fs = 0, F = 0, B = 0;
    int64_t N = 0, NH = 0, n_ud = 0;
    Shape shape{4, 4};
    hipStream_t st = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    tx::DevModel dm{};
    std::vector<void*> allocs;

    // host copies kept to build the stratified layout
    std::vector<uint32_t> t_off, t_item, t_ud, h_start;
    std::vector<float> t_dev, t_r;
    std::vector<uint8_t> t_bin;
    std::vector<int8_t> t_raw;
    std::vector<uint8_t> sud_fbin;
    uint4* d_rec = nullptr;  // packed records, reference order
    bool days_sorted = true;  // every user's training ratings in date order (day terms in parallel)

    // BF16 read copies of the factor tables for predictions (FP32 masters)
    bool bf16 = false, bf16_stale = true;
    uint16_t *b_ifeat = nullptr, *b_ifeat_freq = nullptr, *b_ufeat = nullptr, *b_ufeat2 = nullptr, *b_ycache = nullptr;

    // tsvdx4p schedule: ratings by item, per-user NSVD1 gradients, per-item scales
    bool iv_built = false;
    tx::ItemView iv{};
    float* d_scale = nullptr;
    // the stratified epoch's NSVD1 apply: n_contrib^e per item, e = nsvd_norm_exp
    // or 1 when that is 0
    float* d_scale_strata = nullptr;
    float *d_gacc = nullptr, *d_gacc_b = nullptr;  // [U * fs], [U]
    // ycache equals the NSVD1 sums of the current y (rebuilt since y last changed)
    bool ycache_current = false;
    // passes that hold a side frozen read it from the BF16 copies
    bool bf16_frozen = false;

    // default: one stratum, atomic item updates, lock-free NSVD1 row updates,
    // the reference's user order in length buckets
    tx_schedule sched{TX_SCHED_STRATA, 1, 0, 0, 0, TX_UPD_ATOMIC | (TX_UPD_PLAIN << 2), 2, 1024};
    int built_S = 0;
    tx::StrataView sv{};
    std::vector<void*> strata_allocs;
    hipGraphExec_t graph = nullptr;
    int graph_S = 0;

    // neighbourhood terms (tx5/tx6): |N(u)\R(u)| per user, and the epoch's
    // global neighbour samples drawn on the host (pinned) for the device
    // (two pinned buffers: the next epoch's samples are drawn while the GPU
    // runs the current epoch)
    bool nb = false, nbd = false;
    std::vector<uint32_t> imp_cnt;
    uint32_t* h_smp_cnt[2] = {nullptr, nullptr};
    uint16_t* h_smp[2] = {nullptr, nullptr};
    int32_t smp_buf_epoch[2] = {-1, -1};  // epoch drawn into each host buffer
    int smp_cur = 0;                       // host buffer last uploaded
    uint32_t* d_smp_cnt = nullptr;
    uint16_t* d_smp = nullptr;
    int32_t smp_epoch = -1;  // epoch whose samples are on the device

    // length-bucketed epoch (user_order 3): users by training-rating count in
    // [1,16], [17,256], [257,4096], [4097,...), each bucket its own queue,
    // team shape and teams in flight; one launch for all of them when every
    // bucket runs one-wave teams, else one concurrent launch per bucket
    struct Bucket {
        int32_t lo, hi;
        int64_t users = 0, ratings = 0;
        int lanes = 64, waves = 1, teams = 0;
        double ms = 0.0;
    };
    Bucket bk[TX_N_BUCKETS] = {{1, 16}, {17, 256}, {257, 4096}, {4097, 0x7fffffff}};
    bool bk_ready = false;
    bool bk_one_launch = false;      // the last bucketed epoch ran as one launch
    uint32_t* d_bq_start = nullptr;  // [5] queue offsets of the buckets
    uint32_t* d_bq_head = nullptr;   // [4] their work counters
    hipStream_t bst[TX_N_BUCKETS - 1] = {};
    hipEvent_t bev_fork = nullptr, bev_join[TX_N_BUCKETS - 1] = {};
    hipEvent_t bev_t[2 * TX_N_BUCKETS] = {};

    std::vector<uint32_t> h_queue;  // pinned order of the current epoch's queue
    // the next epoch's queue, built on the host while the GPU runs this one
    std::vector<uint32_t> h_queue_next;
    int32_t next_epoch = -1, next_order = -1, next_hmin = -1;
    int64_t* d_order = nullptr;
    float* d_scratch = nullptr;
    int device = 0;              // the device current at tx_create
    void* d_noalloc = nullptr;   // that device's tx_predict_noalloc scratch array
    // Host-buffer predict / subscores / rmse calls stage through a stager
    // checked out for the call: its own stream, device buffer and FP64 sums,
    // so calls on one ctx from several threads never share device memory.
    struct Stager {
        hipStream_t st = nullptr;
        hipEvent_t ready = nullptr;  // the model's stream up to the call
        void* buf = nullptr;
        size_t bytes = 0;
        double* acc = nullptr;
    };
    std::mutex stage_mu;  // the stager pool and the BF16 refresh
    std::vector<std::unique_ptr<Stager>> stagers;
    std::vector<Stager*> free_stagers;
    tx_epoch_stats last{};

    template <typename T>
    int alloc(T** p, size_t n, std::vector<void*>* list = nullptr) {
        void* q = nullptr;
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(TX_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
        (list ? list : &allocs)->push_back(q);
        *p = static_cast<T*>(q);
        return 0;
    }

    template <typename T>
    int upload(T** p, const T* src, size_t n, std::vector<void*>* list = nullptr) {
        if (int rc = alloc(p, n, list)) return rc;
        if (n) HIPC(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    }


    void free_strata() {
        for (void* p : strata_allocs) (void)hipFree(p);
        strata_allocs.clear();
        built_S = 0;
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_S = 0;
    }

    ~tx_ctx() {
        free_strata();
        for (auto& g : stagers) {
            if (g->st) (void)hipStreamSynchronize(g->st);
            if (g->buf) (void)hipFree(g->buf);
            if (g->acc) (void)hipFree(g->acc);
            if (g->ready) (void)hipEventDestroy(g->ready);
            if (g->st) (void)hipStreamDestroy(g->st);
        }
        for (void* p : allocs) (void)hipFree(p);
        for (auto& x : bst)
            if (x) (void)hipStreamDestroy(x);
        for (auto& e : bev_join)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : bev_t)
            if (e) (void)hipEventDestroy(e);
        if (bev_fork) (void)hipEventDestroy(bev_fork);
        for (int b = 0; b < 2; ++b) {
            if (h_smp[b]) (void)hipHostFree(h_smp[b]);
            if (h_smp_cnt[b]) (void)hipHostFree(h_smp_cnt[b]);
        }
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

int check_ratings(const tx_ratings* r, const char* what, bool need_targets) {
    if (!r) return fail(TX_EINVAL, "%s: NULL rating set", what);
    if (r->n_ratings < 0 || r->n_users < 0 || r->n_items < 0) return fail(TX_EINVAL, "%s: negative size", what);
    if (r->n_ratings > 0 && (!r->user_idxs || !r->item_idxs || !r->dates))
        return fail(TX_EINVAL, "%s: missing user/item/date arrays", what);
    if (!r->user_cnts && r->n_users > 0) return fail(TX_EINVAL, "%s: missing user_cnts", what);
    if (need_targets && r->n_ratings > 0 && !r->residuals) return fail(TX_EINVAL, "%s: missing residuals", what);
    for (int64_t t = 0; t < r->n_ratings; ++t) {
        if (r->user_idxs[t] < 0 || r->user_idxs[t] >= r->n_users)
            return fail(TX_EINVAL, "%s: user index %d out of range at row %lld", what, r->user_idxs[t], (long long)t);
        if (r->item_idxs[t] < 0 || r->item_idxs[t] >= r->n_items)
            return fail(TX_EINVAL, "%s: item index %d out of range at row %lld", what, r->item_idxs[t], (long long)t);
    }
    return 0;
}

tx::Hyper make_hyper(const tx_config& c) {
    tx::Hyper h{};
    h.lr_u = c.lr_u; h.lr_ud = c.lr_ud; h.lr_u2 = c.lr_u2; h.lr_ub = c.lr_ub; h.lr_ubd = c.lr_ubd;
    h.lr_i = c.lr_i; h.lr_ib = c.lr_ib; h.lr_y = c.lr_y; h.lr_yb = c.lr_yb; h.lr_yd = c.lr_yd;
    h.lr_tu = c.lr_tu; h.lr_ti = c.lr_ti; h.lr_ta = c.lr_ta; h.lr_ibf = c.lr_ibf; h.lr_iqf = c.lr_iqf;
    h.lr_cu = c.lr_cu;
    h.reg_iqf = c.reg_iqf; h.reg_cu = c.reg_cu; h.reg_u = c.reg_u; h.reg_u2 = c.reg_u2; h.reg_ud = c.reg_ud;
    h.reg_i = c.reg_i; h.reg_y = c.reg_y; h.reg_yd = c.reg_yd;
    h.w_bias = c.w_bias; h.w_factor = c.w_factor;
    h.beta = c.beta;
    h.lr_t = c.lr_t; h.reg_t = c.reg_t;
    h.lr_w = c.lr_w; h.lr_c = c.lr_c; h.reg_w = c.reg_w; h.reg_c = c.reg_c;
    h.lr_w_day = c.lr_w_day; h.lr_c_day = c.lr_c_day; h.reg_w_day = c.reg_w_day; h.reg_c_day = c.reg_c_day;
    h.w_nbr = c.w_nbr;
    return h;
}

// Rows of a dense host table [rows, cols] to a device table whose rows are
// `pitch` floats apart, and back.  A pitched host<->device copy goes row by
// row (seconds for 480k rows), so the rows are re-pitched in a host buffer
// and moved in one contiguous copy; columns beyond `cols` of each device row
// (zero padding, or the other table sharing the item-bias rows) are read
// first and written back unchanged, unless `fresh` says the device rows are
// all zero.
int rows_to_device(float* dst, int64_t pitch, const float* src, int64_t cols, int64_t rows, bool fresh = false) {
    if (rows <= 0 || cols <= 0) return 0;
    if (pitch == cols) {
        HIPC(hipMemcpy(dst, src, static_cast<size_t>(rows * cols) * sizeof(float), hipMemcpyHostToDevice));
        return 0;
    }
    const size_t span = static_cast<size_t>((rows - 1) * pitch + cols);
    std::vector<float> st(span, 0.0f);
    if (!fresh) HIPC(hipMemcpy(st.data(), dst, span * sizeof(float), hipMemcpyDeviceToHost));
    parallel_for(rows, [&](int64_t lo, int64_t hi) {
        for (int64_t r = lo; r < hi; ++r)
            std::memcpy(st.data() + r * pitch, src + r * cols, static_cast<size_t>(cols) * sizeof(float));
    });
    HIPC(hipMemcpy(dst, st.data(), span * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int rows_to_host(float* dst, const float* src, int64_t pitch, int64_t cols, int64_t rows) {
    if (rows <= 0 || cols <= 0) return 0;
    if (pitch == cols) {
        HIPC(hipMemcpy(dst, src, static_cast<size_t>(rows * cols) * sizeof(float), hipMemcpyDeviceToHost));
        return 0;
    }
    const size_t span = static_cast<size_t>((rows - 1) * pitch + cols);
    std::vector<float> st(span);
    HIPC(hipMemcpy(st.data(), src, span * sizeof(float), hipMemcpyDeviceToHost));
    parallel_for(rows, [&](int64_t lo, int64_t hi) {
        for (int64_t r = lo; r < hi; ++r)
            std::memcpy(dst + r * cols, st.data() + r * pitch, static_cast<size_t>(cols) * sizeof(float));
    });
    return 0;
}

// Upload a dense [rows, f] host table into a zero-padded [rows, fs] device table.
int upload_padded(tx_ctx* c, float** dst, const float* src, int64_t rows) {
    const int64_t alloc_rows = std::max<int64_t>(rows, 1);
    if (int rc = c->alloc(dst, static_cast<size_t>(alloc_rows) * c->fs)) return rc;
    HIPC(hipMemset(*dst, 0, alloc_rows * c->fs * sizeof(float)));
    if (rows > 0 && c->f > 0 && src) return rows_to_device(*dst, c->fs, src, c->f, rows, true);
    return 0;
}

int zeros(tx_ctx* c, float** dst, int64_t n) {
    if (int rc = c->alloc(dst, static_cast<size_t>(n))) return rc;
    HIPC(hipMemset(*dst, 0, std::max<int64_t>(n, 1) * sizeof(float)));
    return 0;
}

struct ParamDesc {
    float* p;
    int64_t rows, cols;
    int64_t pitch;  // floats between device rows (cols, fs for factor tables, ibs for the item-bias rows)
};

ParamDesc param_desc(tx_ctx* c, int id) {
    const auto& m = c->dm;
    const int64_t U = c->U, I = c->I, D = c->n_ud, f = c->f;
    const bool lm = c->cfg.low_memory != 0;
    switch (id) {
    case TX_P_UBIAS: return {m.ubias, U, 1, 1};
    case TX_P_IBIAS: return {m.ibias, I, 1, m.ibs};
    case TX_P_UFEAT: return {m.ufeat, U, f, c->fs};
    case TX_P_UFEAT2: return {m.ufeat2, U, f, c->fs};
    case TX_P_IFEAT: return {m.ifeat, I, f, c->fs};
    case TX_P_YFEAT: return {m.yfeat, I, f, c->fs};
    case TX_P_YCACHE: return {m.ycache, U, f, c->fs};
    case TX_P_YFEAT_DAY: return {m.yfeat_day, lm ? 0 : I, f, c->fs};
    case TX_P_TU_MEAN: return {m.tu_mean, U, 1, 1};
    case TX_P_ALPHA_U: return {m.alpha_u, U, 1, 1};
    case TX_P_BUT_BIN: return {m.but_bin, U * c->B, 1, 1};
    case TX_P_BIT_BIN: return {m.bit_bin, I, c->B, m.ibs};
    case TX_P_YBIAS: return {m.ybias, I, 1, 1};
    case TX_P_YCACHE_BIAS: return {m.ycache_bias, U, 1, 1};
    case TX_P_UBIAS_DAY: return {m.ubias_day, D, 1, 1};
    case TX_P_UFEAT_DAY: return {m.ufeat_day, lm ? 0 : D, f, c->fs};
    case TX_P_YCACHE_DAY: return {m.ycache_day, lm ? 0 : D, f, c->fs};
    case TX_P_IBIAS_FREQ: return {m.ibias_freq, I * c->F, 1, 1};
    case TX_P_IFEAT_FREQ: return {m.ifeat_freq, I * c->F, f, c->fs};
    case TX_P_CU: return {m.cu, U, 1, 1};
    case TX_P_CUT: return {m.cut, D, 1, 1};
    case TX_P_GBIAS: return {nullptr, 1, 1, 1};
    case TX_P_THRESHOLDS: return {m.thresholds, 4, 1, 1};
    case TX_P_NB_W: return {m.nw, c->nb ? I : 0, I, I};
    case TX_P_NB_C: return {m.nc, c->nb ? I : 0, I, I};
    case TX_P_NB_W_DAY: return {m.nw_day, c->nbd ? I : 0, I, I};
    case TX_P_NB_C_DAY: return {m.nc_day, c->nbd ? I : 0, I, I};
    default: return {nullptr, -1, 0, 0};
    }
}

// Packed training record (see StrataView::p_rec).
inline uint4 pack_rec(const tx_ctx* c, int64_t t) {
    uint4 r;
    const uint32_t raw = c->t_raw.empty() ? 0u : static_cast<uint32_t>(c->t_raw[t]) & 15u;
    r.x = c->t_item[t] | (raw << 20) | (static_cast<uint32_t>(c->t_bin[t]) << 24);
    // training uses freq bin 0 in low-memory mode (src/tx.rs:838-842), the day's bin otherwise
    const uint32_t fb = c->cfg.low_memory ? 0u : static_cast<uint32_t>(c->sud_fbin[c->t_ud[t]]);
    r.y = c->t_ud[t] | (fb << 28);
    std::memcpy(&r.z, &c->t_dev[t], 4);
    std::memcpy(&r.w, &c->t_r[t], 4);
    return r;
}

// Build the stratified layout: item blocks and user blocks balanced by rating
// count (largest first to the lightest block), segments of each user's
// ratings per item block in the rotation order of its block, and per
// (stratum, cell) user queues sorted by segment length (longest first).
int build_item_view(tx_ctx* c);

int build_strata(tx_ctx* c, int S) {
    c->free_strata();
    const int64_t U = c->U, I = c->I, N = c->N;
    const int C = S;  // one cell per user block
    std::vector<int32_t> iblock(I, 0), ublock(U, 0);
    if (S > 1) {
        std::vector<int64_t> icnt(I, 0);
        for (int64_t t = 0; t < N; ++t) icnt[c->t_item[t]]++;
        std::vector<int32_t> ord(I);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return icnt[a] > icnt[b]; });
        std::vector<int64_t> load(S, 0);
        for (int32_t i : ord) {
            const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            iblock[i] = k;
            load[k] += icnt[i];
        }
        std::vector<int32_t> uord(U);
        std::iota(uord.begin(), uord.end(), 0);
        std::stable_sort(uord.begin(), uord.end(), [&](int32_t a, int32_t b) {
            return (c->t_off[a + 1] - c->t_off[a]) > (c->t_off[b + 1] - c->t_off[b]);
        });
        std::fill(load.begin(), load.end(), 0);
        for (int32_t u : uord) {
            const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            ublock[u] = k;
            load[k] += c->t_off[u + 1] - c->t_off[u];
        }
    }
    // segments
    std::vector<uint32_t> seg(static_cast<size_t>(U) * (S + 1));
    std::vector<uint32_t> perm(N);
    parallel_for(U, [&](int64_t lo, int64_t hi) {
        std::vector<uint32_t> cnt(S);
        for (int64_t u = lo; u < hi; ++u) {
            const uint32_t a = c->t_off[u], b = c->t_off[u + 1];
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (uint32_t t = a; t < b; ++t) {
                const int s = S > 1 ? (iblock[c->t_item[t]] - ublock[u] + S) % S : 0;
                cnt[s]++;
            }
            uint32_t* so = &seg[static_cast<size_t>(u) * (S + 1)];
            so[0] = a;
            for (int s = 0; s < S; ++s) so[s + 1] = so[s] + cnt[s];
            std::vector<uint32_t> pos(so, so + S);
            for (uint32_t t = a; t < b; ++t) {
                const int s = S > 1 ? (iblock[c->t_item[t]] - ublock[u] + S) % S : 0;
                perm[pos[s]++] = t;
            }
        }
    });
    // queues
    std::vector<uint32_t> qstart(static_cast<size_t>(S) * C + 1, 0);
    std::vector<std::vector<uint32_t>> qs(static_cast<size_t>(S) * C);
    for (int64_t u = 0; u < U; ++u) {
        const uint32_t* so = &seg[static_cast<size_t>(u) * (S + 1)];
        if (so[S] == so[0]) continue;
        for (int s = 0; s < S; ++s) {
            // the first and last strata visit every user (NSVD1 gather and the y step)
            if (so[s + 1] > so[s] || s == 0 || s == S - 1) qs[static_cast<size_t>(s) * C + ublock[u]].push_back(static_cast<uint32_t>(u));
        }
    }
    std::vector<uint32_t> qusers;
    for (size_t k = 0; k < qs.size(); ++k) {
        const int s = static_cast<int>(k / C);
        auto& q = qs[k];
        std::stable_sort(q.begin(), q.end(), [&](uint32_t a, uint32_t b) {
            const uint32_t* sa = &seg[static_cast<size_t>(a) * (S + 1)];
            const uint32_t* sb = &seg[static_cast<size_t>(b) * (S + 1)];
            // the full-user work of the first/last stratum dominates there
            const uint32_t la = (s == 0 || s == S - 1) ? sa[S] - sa[0] : sa[s + 1] - sa[s];
            const uint32_t lb = (s == 0 || s == S - 1) ? sb[S] - sb[0] : sb[s + 1] - sb[s];
            return la > lb;
        });
        qstart[k + 1] = qstart[k] + static_cast<uint32_t>(q.size());
        qusers.insert(qusers.end(), q.begin(), q.end());
    }
    auto* L = &c->strata_allocs;
    tx::StrataView& sv = c->sv;
    sv = tx::StrataView{};
    sv.q_mode = c->sched.update_flags & 3;
    sv.y_mode = (c->sched.update_flags >> 2) & 3;
    sv.nb_mode = (c->sched.update_flags >> 5) & 3;
    sv.block_threads = (c->sched.teams_per_block > 0 ? c->sched.teams_per_block : 4) * 64;
    sv.n_strata = S;
    sv.n_cells = C;
    uint32_t *d_seg, *d_qu, *d_qs, *d_qh;
    if (int rc = c->upload(&d_seg, seg.data(), seg.size(), L)) return rc;
    if (int rc = c->upload(&d_qu, qusers.data(), qusers.size(), L)) return rc;
    if (int rc = c->upload(&d_qs, qstart.data(), qstart.size(), L)) return rc;
    if (int rc = c->alloc(&d_qh, qstart.size(), L)) return rc;
    sv.seg_off = d_seg;
    sv.q_users = d_qu;
    sv.q_start = d_qs;
    sv.q_head = d_qh;
    if (S == 1) {
        sv.p_rec = c->d_rec;
        sv.gacc = nullptr;
        sv.gacc_b = nullptr;
        if (!c->cfg.low_memory && c->n_ud > 0) {
            float* da;
            if (int rc = c->alloc(&da, static_cast<size_t>(c->n_ud) * c->fs, L)) return rc;
            sv.day_acc = da;
        }
    } else {
        std::vector<uint4> rec(N);
        parallel_for(N, [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k) rec[k] = pack_rec(c, perm[k]);
        });
        uint4* pr;
        float *ga, *gb;
        if (int rc = c->upload(&pr, rec.data(), rec.size(), L)) return rc;
        if (int rc = c->alloc(&ga, static_cast<size_t>(U) * c->fs, L)) return rc;
        if (int rc = c->alloc(&gb, static_cast<size_t>(U), L)) return rc;
        sv.p_rec = pr;
        sv.gacc = ga;
        sv.gacc_b = gb;
        // the epoch's NSVD1 step is gathered per item after the last stratum
        if (int rc = build_item_view(c)) return rc;
    }
    c->built_S = S;
    return 0;
}

// Teams of the parallel schedule in flight at once.  Concurrent users read
// the NSVD1 rows y_j at their start and write them at their end without
// seeing each other's steps; on popular items that delayed feedback grows
// with the number of teams and, past a point, makes the y step overshoot
// (item and user factors tolerate far more concurrency).  The tolerable
// count falls with f.  At the full Netflix shape with the length-bucket order
// (long users started in the first half of the epoch), 3072 teams at f=64 and
// 1536 at f=256 keep the holdout RMSE of 20 epochs within 1e-3 of the serial
// trainer (within 1e-4 measured); 4096 at f=64 and 3072 at f=256 do not.
int target_teams(int f) { return std::min(3072, std::max(768, 393216 / std::max(f, 1))); }

// Teams in flight for a model: target_teams(f), and at most 2,048 for a
// day-term model without neighbourhood terms.  Its y_day steps land once per
// same-day run and are the most staleness-sensitive update: full shape, job
// tsvdx4-60 (f=60), 20 epochs, holdout RMSE 0.90867 with 2,048 teams resident
// against the serial 0.90851, and 0.91548 with 3,072.
int target_teams(const tx_ctx* c) {
    const int t = target_teams(c->cfg.n_feat);
    return (c->cfg.low_memory == 0 && !c->nb) ? std::min(t, 2048) : t;
}

// One-wave teams per SGD workgroup: the schedule's teams_per_block (1-4), 4 by
// default; a multi-wave team is a workgroup of its own.
int teams_per_block(const tx_ctx* c) {
    if (c->shape.waves > 1 || c->sched.split_waves) return 1;
    return c->sched.teams_per_block > 0 ? c->sched.teams_per_block : 256 / c->shape.team;
}

int default_blocks(tx_ctx* c) {
    const int teams_per_block = ::teams_per_block(c);
    const int want = target_teams(c);
    int blocks = (want + teams_per_block - 1) / teams_per_block;
    // a multiple of 8 keeps the cell -> XCD mapping of the round-robin dispatch
    return std::max(8, (blocks + 7) / 8 * 8);
}

// BF16 copies of q, q_f, p, p2 and ycache (row stride fs), allocated once.
int alloc_bf16(tx_ctx* c) {
    if (c->b_ifeat) return 0;
    const size_t U = std::max<int64_t>(c->U, 1), I = std::max<int64_t>(c->I, 1), fs = c->fs, F = c->F;
    RC(c->alloc(&c->b_ifeat, I * fs));
    RC(c->alloc(&c->b_ifeat_freq, I * F * fs));
    RC(c->alloc(&c->b_ufeat, U * fs));
    RC(c->alloc(&c->b_ufeat2, U * fs));
    RC(c->alloc(&c->b_ycache, U * fs));
    c->dm.b_ifeat = c->b_ifeat;
    c->dm.b_ifeat_freq = c->b_ifeat_freq;
    c->dm.b_ufeat = c->b_ufeat;
    c->dm.b_ufeat2 = c->b_ufeat2;
    c->dm.b_ycache = c->b_ycache;
    return 0;
}

// Refresh the BF16 copies when predictions read them and a parameter changed.
// The copies are converted on the model's stream, which is then waited for, so
// a prediction on any stream that starts after this returns reads them whole;
// concurrent callers serialise on the ctx's staging lock.
int ensure_bf16(tx_ctx* c) {
    std::lock_guard<std::mutex> lock(c->stage_mu);
    if (!c->bf16 || !c->bf16_stale) return 0;
    const int64_t U = c->U, I = c->I, fs = c->fs, F = c->F;
    hipStream_t st = c->st;
    HIPC(tx::launch_to_bf16(c->dm.ifeat, c->b_ifeat, I * fs, st));
    HIPC(tx::launch_to_bf16(c->dm.ifeat_freq, c->b_ifeat_freq, I * F * fs, st));
    HIPC(tx::launch_to_bf16(c->dm.ufeat, c->b_ufeat, U * fs, st));
    HIPC(tx::launch_to_bf16(c->dm.ufeat2, c->b_ufeat2, U * fs, st));
    HIPC(tx::launch_to_bf16(c->dm.ycache, c->b_ycache, U * fs, st));
    HIPC(hipStreamSynchronize(st));
    c->bf16_stale = false;
    return 0;
}

// Users of the epoch in the reference's order (get_users, src/lib.rs:486-494)
// into the one-stratum queue; with length buckets (user_order 2) the users of
// at least heavy_min ratings keep their relative order but are spread evenly
// over the first half of the queue.
// The queue of an epoch: the reference's per-epoch user order (get_users),
// users without ratings dropped, with user_order 2 the long users spread over
// the first half.
void build_queue(const tx_ctx* c, int32_t epoch, std::vector<uint32_t>& q) {
    std::vector<int64_t> order(c->U);
    tx::user_order(c->U, c->cfg.shuffle_users != 0, c->cfg.seed, epoch, order.data());
    q.clear();
    if (c->sched.user_order == 3) {
        // the buckets one after another, each in the epoch's order
        for (int b = 0; b < TX_N_BUCKETS; ++b)
            for (int64_t u : order) {
                const int64_t cnt = c->t_off[u + 1] - c->t_off[u];
                if (cnt >= c->bk[b].lo && cnt <= c->bk[b].hi) q.push_back(static_cast<uint32_t>(u));
            }
    } else if (c->sched.user_order == 2) {
        const uint32_t hmin = c->sched.heavy_min > 0 ? static_cast<uint32_t>(c->sched.heavy_min) : 1024u;
        std::vector<uint32_t> heavy, light;
        for (int64_t u : order) {
            const uint32_t cnt = c->t_off[u + 1] - c->t_off[u];
            if (cnt) (cnt >= hmin ? heavy : light).push_back(static_cast<uint32_t>(u));
        }
        const size_t n = heavy.size() + light.size(), half = n / 2;
        size_t k = 0, l = 0;
        for (size_t pos = 0; pos < n; ++pos) {
            const bool take_heavy = k < heavy.size() && (l >= light.size() || pos * heavy.size() >= k * half);
            q.push_back(take_heavy ? heavy[k++] : light[l++]);
        }
    } else {
        for (int64_t u : order)
            if (c->t_off[u + 1] > c->t_off[u]) q.push_back(static_cast<uint32_t>(u));
    }
}

// Build the next epoch's queue ahead (called while the GPU runs an epoch).
void build_next_queue(tx_ctx* c, int32_t epoch) {
    build_queue(c, epoch, c->h_queue_next);
    c->next_epoch = epoch;
    c->next_order = c->sched.user_order;
    c->next_hmin = c->sched.heavy_min;
}

int prepare_user_queue(tx_ctx* c, int32_t epoch, hipStream_t st) {
    if (c->next_epoch == epoch && c->next_order == c->sched.user_order && c->next_hmin == c->sched.heavy_min) {
        std::swap(c->h_queue, c->h_queue_next);
    } else {
        build_queue(c, epoch, c->h_queue);
    }
    c->next_epoch = -1;
    HIPC(hipMemcpyAsync(const_cast<uint32_t*>(c->sv.q_users), c->h_queue.data(), c->h_queue.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice, st));
    return 0;
}

// The epoch's global neighbour samples (src/tx.rs:750-793): one
// StdRng::seed_from_u64(seed + epoch * 1000003) stream, consumed user by user
// in the epoch's order (get_users); a user whose |R(u)| + |N(u)\R(u)|
// exceeds max_neighbors keeps round(max * |R| / n) (at least 1, at most |R|)
// of its R(u) positions and the rest of the budget of its N(u)\R(u)
// positions, each set the prefix of a full shuffle, sorted.  The stream is
// sequential, so its draws are taken in one pass over the users in order
// (no divisions: shuffle_draws) and turned into the shuffles in parallel over
// users (shuffle_apply), into host buffer `buf`.
int draw_nbr_samples(tx_ctx* c, int32_t epoch, int buf) {
    const int64_t U = c->U;
    const int64_t mx = c->cfg.max_neighbors;
    uint32_t* cnt = c->h_smp_cnt[buf];
    for (int64_t u = 0; u < U; ++u) cnt[u] = 0xffffffffu;
    if (mx > 0) {
        std::vector<int64_t> order(U);
        tx::user_order(U, c->cfg.shuffle_users != 0, c->cfg.seed, epoch, order.data());
        tx::ChaChaRng rng(c->cfg.seed + static_cast<uint64_t>(epoch) * 1000003ull);
        rng.set_bulk(true);
        std::vector<uint32_t> pk;
        std::vector<int64_t> su, pk_at;  // sampled users and the offset of their draws
        pk.reserve(static_cast<size_t>(c->N / 2 + 1024));
        for (int64_t k = 0; k < U; ++k) {
            const int64_t u = order[k];
            const int64_t nr = c->t_off[u + 1] - c->t_off[u], ni = c->imp_cnt[u];
            if (nr == 0 || nr + ni <= mx) continue;
            su.push_back(u);
            pk_at.push_back(static_cast<int64_t>(pk.size()));
            tx::shuffle_draws(rng, nr, pk);
            tx::shuffle_draws(rng, ni, pk);
        }
        parallel_for(static_cast<int64_t>(su.size()), [&](int64_t lo, int64_t hi) {
            std::vector<int64_t> ir, ii;
            for (int64_t k = lo; k < hi; ++k) {
                const int64_t u = su[k];
                const int64_t nr = c->t_off[u + 1] - c->t_off[u], ni = c->imp_cnt[u], n = nr + ni;
                int64_t rb = static_cast<int64_t>(std::round(static_cast<double>(mx) * static_cast<double>(nr) /
                                                             static_cast<double>(n)));
                rb = std::min<int64_t>(std::max<int64_t>(rb, 1), nr);
                const int64_t ib = std::min<int64_t>(mx - rb, ni);
                const uint32_t* p = pk.data() + pk_at[k];
                int64_t ng = 0;  // groups of the first shuffle
                tx::shuffle_groups(nr, [&](int64_t, uint32_t, int64_t, uint32_t) { ++ng; });
                ir.resize(nr);
                std::iota(ir.begin(), ir.end(), int64_t{0});
                tx::shuffle_apply(p, ir.data(), nr);
                std::sort(ir.begin(), ir.begin() + rb);
                ii.resize(ni);
                std::iota(ii.begin(), ii.end(), int64_t{0});
                tx::shuffle_apply(p + ng, ii.data(), ni);
                std::sort(ii.begin(), ii.begin() + ib);
                uint16_t* out = c->h_smp[buf] + u * mx;
                for (int64_t q = 0; q < rb; ++q) out[q] = static_cast<uint16_t>(ir[q]);
                for (int64_t q = 0; q < ib; ++q) out[rb + q] = static_cast<uint16_t>(ii[q]);
                cnt[u] = static_cast<uint32_t>(rb) | (static_cast<uint32_t>(ib) << 16);
            }
        });
    }
    c->smp_buf_epoch[buf] = epoch;
    return 0;
}

// The epoch's samples on the device (drawn now unless a host buffer holds
// them already), copied on the ctx's stream ahead of the epoch's kernels.
int upload_nbr_samples(tx_ctx* c, int32_t epoch, hipStream_t st) {
    int b = c->smp_buf_epoch[0] == epoch ? 0 : (c->smp_buf_epoch[1] == epoch ? 1 : -1);
    if (b < 0) {
        b = 1 - c->smp_cur;
        if (int rc = draw_nbr_samples(c, epoch, b)) return rc;
    }
    const int64_t mx = c->cfg.max_neighbors;
    if (mx > 0)
        HIPC(hipMemcpyAsync(c->d_smp, c->h_smp[b], sizeof(uint16_t) * c->U * mx, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(c->d_smp_cnt, c->h_smp_cnt[b], sizeof(uint32_t) * c->U, hipMemcpyHostToDevice, st));
    c->smp_cur = b;
    c->smp_epoch = epoch;
    return 0;
}

// While the GPU runs epoch `epoch`, draw the next epoch's samples into the
// host buffer not being copied.
int draw_next_nbr_samples(tx_ctx* c, int32_t epoch) {
    if (!c->nb || c->smp_buf_epoch[1 - c->smp_cur] == epoch + 1) return 0;
    return draw_nbr_samples(c, epoch + 1, 1 - c->smp_cur);
}

// The tsvdx4p layout: ratings grouped by item in training-file order
// (ByItemIndex, src/tsvdx4p.rs:29-57) with the user, bins, dev and target of
// each; items queued most ratings first; NSVD1 accumulators; per-item scale
// n_contrib^nsvd_norm_exp (src/tsvdx4p.rs:356-365) with the host libm powf.
int build_item_view(tx_ctx* c) {
    if (c->iv_built) return 0;
    const int64_t I = c->I, N = c->N;
    if (c->U > (1 << 20) || c->B > 256 || c->F > 16)
        return fail(TX_EUNSUPPORTED, "tsvdx4p schedule: needs < 2^20 users, <= 256 time bins, <= 16 freq bins");
    std::vector<uint32_t> start(I + 1, 0);
    for (int64_t t = 0; t < N; ++t) start[c->t_item[t] + 1]++;
    for (int64_t i = 0; i < I; ++i) start[i + 1] += start[i];
    std::vector<uint32_t> pos(start.begin(), start.end() - 1);
    std::vector<uint4> rec(std::max<int64_t>(N, 1));
    for (int64_t u = 0; u < c->U; ++u)
        for (uint32_t t = c->t_off[u]; t < c->t_off[u + 1]; ++t) {
            const uint32_t i = c->t_item[t], ud = c->t_ud[t];
            uint4 r;
            r.x = static_cast<uint32_t>(u) | (static_cast<uint32_t>(c->t_bin[t]) << 20) |
                  (static_cast<uint32_t>(c->sud_fbin[ud]) << 28);
            r.y = ud;
            std::memcpy(&r.z, &c->t_dev[t], 4);
            std::memcpy(&r.w, &c->t_r[t], 4);
            rec[pos[i]++] = r;
        }
    std::vector<uint32_t> items;
    for (int64_t i = 0; i < I; ++i)
        if (start[i + 1] > start[i]) items.push_back(static_cast<uint32_t>(i));
    std::stable_sort(items.begin(), items.end(),
                     [&](uint32_t a, uint32_t b) { return start[a + 1] - start[a] > start[b + 1] - start[b]; });
    std::vector<float> scale(std::max<int64_t>(I, 1), 0.0f);
    const float ex = c->cfg.nsvd_norm_exp;
    for (int64_t i = 0; i < I; ++i) {
        const uint32_t nc = start[i + 1] - start[i];
        if (nc) scale[i] = ex == 0.0f ? 1.0f : std::pow(static_cast<float>(nc), ex);
    }
    uint32_t *d_start = nullptr, *d_items = nullptr, *d_head = nullptr, *d_bu = nullptr;
    uint4* d_bi = nullptr;
    RC(c->upload(&d_start, start.data(), start.size()));
    RC(c->upload(&d_bi, rec.data(), rec.size()));
    {
        std::vector<uint32_t> bu(rec.size());
        for (size_t k = 0; k < rec.size(); ++k) bu[k] = rec[k].x & 0xFFFFFu;
        RC(c->upload(&d_bu, bu.data(), bu.size()));
    }
    if (items.empty()) items.push_back(0);
    RC(c->upload(&d_items, items.data(), items.size()));
    RC(c->alloc(&d_head, 1));
    RC(c->upload(&c->d_scale, scale.data(), scale.size()));
    const float exs = ex > 0.0f ? ex : 1.0f;
    for (int64_t i = 0; i < I; ++i) {
        const uint32_t nc = start[i + 1] - start[i];
        scale[i] = nc ? std::pow(static_cast<float>(nc), exs) : 0.0f;
    }
    RC(c->upload(&c->d_scale_strata, scale.data(), scale.size()));
    // each user's NSVD1 gradient over its norm, written by the user pass and
    // gathered per item
    RC(c->alloc(&c->d_gacc, static_cast<size_t>(std::max<int32_t>(c->U, 1)) * c->fs));
    RC(c->alloc(&c->d_gacc_b, static_cast<size_t>(std::max<int32_t>(c->U, 1))));
    c->iv.q_items = d_items;
    c->iv.n_items = start[I] ? static_cast<uint32_t>(items.size()) : 0u;
    c->iv.q_head = d_head;
    c->iv.bi_start = d_start;
    c->iv.bi_rec = d_bi;
    c->iv.bi_user = d_bu;
    c->iv_built = true;
    return 0;
}

int user_pass_blocks(tx_ctx* c) {
    // users are independent in the user passes: fill the device
    const int teams_per_block = c->shape.waves > 1 ? 1 : 256 / c->shape.team;
    const int blocks = (4096 + teams_per_block - 1) / teams_per_block;
    return std::max(8, (blocks + 7) / 8 * 8);
}

// One Tsvdx4pModel epoch (src/tsvdx4p.rs:798-805): user pass with NSVD1
// accumulators, delta apply, ycache rebuild, item pass.
int tsvdx4p_epoch(tx_ctx* c, int32_t epoch, tx_epoch_stats& stats, hipStream_t st) {
    const bool qf = c->cfg.lr_iqf != 0.0f || c->cfg.sigma_iqf != 0.0f;
    const bool p2 = c->cfg.lr_u2 != 0.0f;
    if (!c->cfg.low_memory || c->cfg.has_ordinal)
        return fail(TX_EUNSUPPORTED, "tsvdx4p schedule: low-memory form without ordinal head only");
    if (c->built_S != 1) {
        if (c->built_S != 0) return fail(TX_EUNSUPPORTED, "tsvdx4p schedule: needs the one-stratum layout");
        if (int rc = build_strata(c, 1)) return rc;
    }
    if (int rc = build_item_view(c)) return rc;
    if (int rc = prepare_user_queue(c, epoch, st)) return rc;
    // (split teams count workgroups of one team: the user passes size their own grid then)
    const int sgpr_blocks = c->sched.blocks > 0 && !c->sched.split_waves ? c->sched.blocks : user_pass_blocks(c);
    HIPC(tx::launch_fill_u32(c->sv.q_head, 0u, c->sv.n_cells, st));
    HIPC(tx::launch_fill_u32(c->iv.q_head, 0u, 1, st));
    tx::DevModel dm = c->dm;
    tx::StrataView sv = c->sv;
    sv.gacc = c->d_gacc;
    sv.gacc_b = c->d_gacc_b;
    sv.su_cached = (c->cfg.full_su && c->ycache_current) ? 1 : 0;
    if (c->bf16_frozen) {
        // the item side the user pass holds frozen, once per pass
        dm.bf16_frozen = 1;
        HIPC(tx::launch_to_bf16(c->dm.ifeat, c->b_ifeat, static_cast<int64_t>(c->I) * c->fs, st));
        if (qf) HIPC(tx::launch_to_bf16(c->dm.ifeat_freq, c->b_ifeat_freq, static_cast<int64_t>(c->I) * c->F * c->fs, st));
    }
    HIPC(hipEventRecord(c->ev[1], st));
    if (c->shape.waves > 1)
        HIPC(tx::launch_sgd_par_mw(dm, sv, sgpr_blocks, c->shape.waves, qf, p2, false, false, st, 1));
    else
        HIPC(tx::launch_sgd_par(dm, sv, 0, c->shape.team, c->shape.vec, sgpr_blocks, qf, p2, false, 1, false, st));
    HIPC(tx::launch_nsvd_gather(c->dm, c->iv, c->d_gacc, c->d_gacc_b, c->d_scale, c->shape.vec, st));
    HIPC(hipEventRecord(c->ev[2], st));
    HIPC(tx::launch_rebuild(c->dm, c->n_ud, c->shape.team, c->shape.vec, st));
    if (c->bf16_frozen) {
        // the user side the item pass holds frozen
        const int64_t un = static_cast<int64_t>(c->U) * c->fs;
        HIPC(tx::launch_to_bf16(c->dm.ufeat, c->b_ufeat, un, st));
        HIPC(tx::launch_to_bf16(c->dm.ycache, c->b_ycache, un, st));
        if (p2) HIPC(tx::launch_to_bf16(c->dm.ufeat2, c->b_ufeat2, un, st));
    }
    HIPC(hipEventRecord(c->ev[3], st));
    HIPC(tx::launch_item_pass(dm, c->iv, c->shape.team, c->shape.vec, p2, qf, user_pass_blocks(c), st,
                              c->shape.waves));
    HIPC(hipEventSynchronize(c->ev[3]));
    HIPC(hipStreamSynchronize(st));
    float a = 0, b = 0;
    HIPC(hipEventElapsedTime(&a, c->ev[1], c->ev[2]));
    HIPC(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
    stats.ms_sgd = a;
    stats.ms_rebuild = b;
    stats.n_launches = 4;
    return 0;
}

// Whether the bucket shapes other than one 64-lane wave per user apply:
// low-memory models without ordinal head or neighbourhood terms.
bool bucket_shapes_ok(const tx_ctx* c) {
    return c->cfg.low_memory && !c->cfg.has_ordinal && !c->nb && c->shape.waves == 1;
}

// Bucket sizes, default shapes and teams in flight (once per ctx): one-wave
// teams in every bucket (one launch for all), the teams of the fidelity
// bound split in proportion to each bucket's ratings plus a 16-rating
// per-user overhead, at least 8 (at most the bucket's users).
int init_buckets(tx_ctx* c) {
    if (c->bk_ready) return 0;
    for (auto& b : c->bk) b.users = b.ratings = 0;
    for (int64_t u = 0; u < c->U; ++u) {
        const int64_t cnt = c->t_off[u + 1] - c->t_off[u];
        for (auto& b : c->bk)
            if (cnt >= b.lo && cnt <= b.hi) {
                b.users++;
                b.ratings += cnt;
            }
    }
    double work[TX_N_BUCKETS], tot = 0.0;
    for (int b = 0; b < TX_N_BUCKETS; ++b) {
        work[b] = static_cast<double>(c->bk[b].ratings) + 16.0 * static_cast<double>(c->bk[b].users);
        tot += work[b];
    }
    const int T = target_teams(c);
    for (int b = 0; b < TX_N_BUCKETS; ++b) {
        auto& k = c->bk[b];
        k.lanes = 64;
        k.waves = 1;
        const int64_t want = tot > 0.0 ? static_cast<int64_t>(T * work[b] / tot + 0.5) : 8;
        k.teams = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(want, 8), k.users)));
        k.ms = 0.0;
    }
    uint32_t off[TX_N_BUCKETS + 1] = {0};
    for (int b = 0; b < TX_N_BUCKETS; ++b) off[b + 1] = off[b] + static_cast<uint32_t>(c->bk[b].users);
    RC(c->upload(&c->d_bq_start, off, TX_N_BUCKETS + 1));
    RC(c->alloc(&c->d_bq_head, TX_N_BUCKETS));
    for (auto& x : c->bst) HIPC(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&c->bev_fork, hipEventDisableTiming));
    for (auto& e : c->bev_join) HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : c->bev_t) HIPC(hipEventCreate(&e));
    c->bk_ready = true;
    return 0;
}

// The bucketed epoch.  Every bucket with one-wave teams: one launch on the
// ctx's stream whose workgroup ranges serve the bucket queues, so the buckets
// run together whatever hardware queues the process's streams share.
// Otherwise one launch per bucket, concurrently: bucket 0 on the ctx's
// stream, the others on their own streams forked from it and joined back
// (kernels on streams that share a hardware queue run one after the other).
// `timed` records the launches between events (stream launches only, not
// under graph capture).
int enqueue_buckets(tx_ctx* c, hipStream_t st, bool timed) {
    const bool qf = c->cfg.lr_iqf != 0.0f || c->cfg.sigma_iqf != 0.0f;
    const bool p2 = c->cfg.lr_u2 != 0.0f;
    const bool ord = c->cfg.has_ordinal != 0;
    const bool day = c->cfg.low_memory == 0;
    HIPC(tx::launch_fill_u32(c->d_bq_head, 0u, TX_N_BUCKETS, st));
    bool one = true;
    for (const auto& k : c->bk)
        if (k.users > 0 && (k.lanes != 64 || k.waves != 1)) one = false;
    c->bk_one_launch = one;
    if (one) {
        static_assert(TX_N_BUCKETS == 4, "the kernel names a workgroup's queue from four bucket ranges");
        tx::StrataView sv = c->sv;
        sv.q_start = c->d_bq_start;
        sv.q_head = c->d_bq_head;
        sv.block_threads = 64;  // one team per workgroup
        sv.n_bq = TX_N_BUCKETS;
        int total = 0;
        sv.bq_first[0] = 0;
        for (int b = 0; b < TX_N_BUCKETS; ++b) {
            total += c->bk[b].users > 0 ? c->bk[b].teams : 0;
            sv.bq_first[b + 1] = total;
        }
        if (total < 1) return 0;
        if (timed) HIPC(hipEventRecord(c->bev_t[0], st));
        HIPC(tx::launch_sgd_par(c->dm, sv, 0, c->shape.team, c->shape.vec, total, qf, p2, ord, 0, day, st));
        if (timed) HIPC(hipEventRecord(c->bev_t[1], st));
        return 0;
    }
    HIPC(hipEventRecord(c->bev_fork, st));
    for (int b = 1; b < TX_N_BUCKETS; ++b) HIPC(hipStreamWaitEvent(c->bst[b - 1], c->bev_fork, 0));
    for (int b = 0; b < TX_N_BUCKETS; ++b) {
        const auto& k = c->bk[b];
        if (k.users == 0) continue;
        hipStream_t s = b == 0 ? st : c->bst[b - 1];
        tx::StrataView sv = c->sv;
        sv.q_start = c->d_bq_start + b;
        sv.q_head = c->d_bq_head + b;
        if (timed) HIPC(hipEventRecord(c->bev_t[2 * b], s));
        if (k.lanes == 16) {
            HIPC(tx::launch_sgd_par_t16(c->dm, sv, (k.teams + 15) / 16, qf, p2, s));
        } else if (k.waves == 2) {
            HIPC(tx::launch_sgd_par_w2(c->dm, sv, k.teams, qf, p2, s));
        } else {
            sv.block_threads = 64;  // one team per workgroup: teams = blocks
            HIPC(tx::launch_sgd_par(c->dm, sv, 0, c->shape.team, c->shape.vec, k.teams, qf, p2, ord, 0, day, s));
        }
        if (timed) HIPC(hipEventRecord(c->bev_t[2 * b + 1], s));
    }
    for (int b = 1; b < TX_N_BUCKETS; ++b) {
        HIPC(hipEventRecord(c->bev_join[b - 1], c->bst[b - 1]));
        HIPC(hipStreamWaitEvent(st, c->bev_join[b - 1], 0));
    }
    return 0;
}

// Per-bucket launch times of the last timed epoch (with one launch for all
// buckets, that launch's time for each).
int read_bucket_times(tx_ctx* c) {
    float one_ms = 0.0f;
    if (c->bk_one_launch) HIPC(hipEventElapsedTime(&one_ms, c->bev_t[0], c->bev_t[1]));
    for (int b = 0; b < TX_N_BUCKETS; ++b) {
        float ms = 0.0f;
        if (c->bk[b].users > 0) {
            if (c->bk_one_launch) ms = one_ms;
            else HIPC(hipEventElapsedTime(&ms, c->bev_t[2 * b], c->bev_t[2 * b + 1]));
        }
        c->bk[b].ms = ms;
    }
    return 0;
}

// Enqueue one parallel epoch (queue reset, S strata, ycache rebuild).
int enqueue_strata_epoch(tx_ctx* c, hipStream_t st) {
    const int S = c->built_S;
    const bool qf = c->cfg.lr_iqf != 0.0f || c->cfg.sigma_iqf != 0.0f;
    const bool p2 = c->cfg.lr_u2 != 0.0f;
    const bool ord = c->cfg.has_ordinal != 0;
    const bool day = c->cfg.low_memory == 0;
    const int blocks = c->sched.blocks > 0 ? c->sched.blocks : default_blocks(c);
    HIPC(tx::launch_fill_u32(c->sv.q_head, 0u, static_cast<int64_t>(S) * c->sv.n_cells, st));
    if (c->sched.user_order == 3 && S == 1)
        RC(enqueue_buckets(c, st, false));
    else if (c->sched.split_waves)
        HIPC(tx::launch_sgd_par_split(c->dm, c->sv, c->shape.vec, blocks, qf, p2, st));
    else if (c->shape.waves > 1)
        HIPC(tx::launch_sgd_par_mw(c->dm, c->sv, blocks, c->shape.waves, qf, p2, ord, day, st));
    else
        for (int s = 0; s < S; ++s)
            HIPC(tx::launch_sgd_par(c->dm, c->sv, s, c->shape.team, c->shape.vec, blocks, qf, p2, ord, 0, day, st));
    if (S > 1)
        HIPC(tx::launch_nsvd_gather(c->dm, c->iv, c->sv.gacc, c->sv.gacc_b, c->d_scale_strata, c->shape.vec, st));
    HIPC(tx::launch_rebuild(c->dm, c->n_ud, c->shape.team, c->shape.vec, st));
    return 0;
}

int enforce_thresholds(tx_ctx* c) {
    if (!c->cfg.has_ordinal) return 0;
    float th[4];
    HIPC(hipMemcpyAsync(th, c->dm.thresholds, sizeof th, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    // OrdinalHead::enforce_sorted_with_gap (src/lib.rs:519-525)
    std::sort(th, th + 4);
    for (int k = 1; k < 4; ++k) {
        const float min_ok = th[k - 1] + c->cfg.th_gap;
        if (th[k] < min_ok) th[k] = min_ok;
    }
    HIPC(hipMemcpyAsync(c->dm.thresholds, th, sizeof th, hipMemcpyHostToDevice, c->st));
    HIPC(hipStreamSynchronize(c->st));
    return 0;
}

}  // namespace

extern "C" {

const char* tx_last_error(void) { return g_err.c_str(); }

int tx_get_users(int64_t n, int32_t shuffle, uint64_t seed, int64_t epoch, int64_t* out) {
    g_err.clear();
    if (n < 0 || (n > 0 && !out)) return fail(TX_EINVAL, "tx_get_users: bad arguments");
    tx::user_order(n, shuffle != 0, seed, epoch, out);
    return 0;
}

int tx_rand_normal(uint64_t seed, float sigma, int64_t n, float* out) {
    g_err.clear();
    if (n < 0 || (n > 0 && !out) || !(sigma >= 0.0f) || !std::isfinite(sigma))
        return fail(TX_EINVAL, "tx_rand_normal: bad arguments");
    tx::ChaChaRng rng(seed);
    tx::fill_normal(rng, out, n, sigma);
    return 0;
}

int tx_create(const tx_config* cfg, const tx_ratings* tr, const tx_ratings* pr, tx_ctx** out) {
    g_err.clear();
    if (!cfg || !out) return fail(TX_EINVAL, "tx_create: NULL argument");
    *out = nullptr;
    const bool build_nbr = cfg->lr_w > 0.0f || cfg->lr_c > 0.0f;          // src/tx.rs:483-484
    const bool build_nbr_day = cfg->lr_w_day > 0.0f || cfg->lr_c_day > 0.0f;
    if (build_nbr && tr->n_ratings > 0 && !tr->raw_ratings)
        return fail(TX_EINVAL, "tx_create: the neighbourhood terms need train raw_ratings (r - b~)");
    if ((build_nbr || build_nbr_day) && cfg->max_neighbors < 0)
        return fail(TX_EINVAL, "tx_create: max_neighbors >= 0 required");
    if ((build_nbr || build_nbr_day) && static_cast<int64_t>(tr->n_items) * tr->n_items > (int64_t{1} << 32))
        return fail(TX_EUNSUPPORTED, "tx_create: the neighbourhood tables need n_items^2 <= 2^32");
    if (cfg->n_feat < 0 || cfg->n_time_bins < 1 || cfg->n_freq_bins < 1 || cfg->n_epochs < 0)
        return fail(TX_EINVAL, "tx_create: n_feat >= 0, n_time_bins >= 1, n_freq_bins >= 1 required");
    if (int rc = check_ratings(tr, "train", true)) return rc;
    if (int rc = check_ratings(pr, "heldout", false)) return rc;
    if (pr->n_users != tr->n_users || pr->n_items != tr->n_items)
        return fail(TX_EINVAL, "tx_create: train and held-out sets disagree on n_users/n_items");
    if (cfg->has_ordinal && tr->n_ratings > 0 && !tr->raw_ratings)
        return fail(TX_EINVAL, "tx_create: the ordinal head needs train raw_ratings");
    if (tr->n_ratings + pr->n_ratings >= (int64_t{1} << 32))
        return fail(TX_EUNSUPPORTED, "tx_create: more than 2^32 ratings");
    PhaseClock clk;
    Shape shape;
    if (!pick_shape(cfg->n_feat, &shape)) return fail(TX_EUNSUPPORTED, "tx_create: n_feat > 8192");

    auto ctx = std::make_unique<tx_ctx>();
    tx_ctx* c = ctx.get();
    c->cfg = *cfg;
    c->U = tr->n_users;
    c->I = tr->n_items;
    c->f = cfg->n_feat;
    c->shape = shape;
    c->fs = shape.team * shape.vec * shape.waves;
    c->F = cfg->n_freq_bins;
    c->B = cfg->n_time_bins;
    c->N = tr->n_ratings;
    c->NH = pr->n_ratings;
    const int64_t U = c->U, I = c->I, N = c->N, NH = c->NH;
    const int f = c->f, F = c->F, B = c->B;

    // initial factors in the reference's draw order (src/tx.rs:606-630):
    // ufeat, ifeat, yfeat, [yfeat_day], ifeat_freq.  They depend on the sizes
    // and the seed only, so they are drawn on a thread of their own while the
    // rating structures below are built (joined before the tables go to the
    // device; `draws` is declared after the tables, so an early return joins
    // the thread before the tables are freed).
    std::vector<float> h_ufeat, h_ifeat, h_yfeat, h_yday, h_iqf;
    struct Joiner {
        std::thread t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } draws;
    draws.t = std::thread([&h_ufeat, &h_ifeat, &h_yfeat, &h_yday, &h_iqf, U, I, f, F, cc = c->cfg]() {
        tx::ChaChaRng rng(cc.seed);
        h_ufeat.resize(static_cast<size_t>(U) * f);
        h_ifeat.resize(static_cast<size_t>(I) * f);
        h_yfeat.resize(static_cast<size_t>(I) * f);
        h_iqf.resize(static_cast<size_t>(I) * F * f);
        tx::fill_normal(rng, h_ufeat.data(), static_cast<int64_t>(h_ufeat.size()), cc.sigma_u);
        tx::fill_normal(rng, h_ifeat.data(), static_cast<int64_t>(h_ifeat.size()), cc.sigma_i);
        tx::fill_normal(rng, h_yfeat.data(), static_cast<int64_t>(h_yfeat.size()), cc.sigma_y);
        if (!cc.low_memory) {
            h_yday.resize(static_cast<size_t>(I) * f);
            tx::fill_normal(rng, h_yday.data(), static_cast<int64_t>(h_yday.size()), cc.sigma_yd);
        }
        tx::fill_normal(rng, h_iqf.data(), static_cast<int64_t>(h_iqf.size()), cc.sigma_iqf);
    });

    // calc_user_offsets (src/lib.rs:469-483): user-sorted input, spans = counts
    for (int64_t t = 1; t < N; ++t)
        if (tr->user_idxs[t] < tr->user_idxs[t - 1])
            return fail(TX_EINVAL, "dataset 'train' is not user-sorted at index %lld: user %d follows user %d",
                        (long long)t, tr->user_idxs[t], tr->user_idxs[t - 1]);
    c->t_off.assign(U + 1, 0);
    {
        std::vector<int64_t> seen(U, 0);
        for (int64_t t = 0; t < N; ++t) seen[tr->user_idxs[t]]++;
        for (int64_t u = 0; u < U; ++u) {
            if (seen[u] != tr->user_cnts[u])
                return fail(TX_EINVAL, "train user_cnts[%lld] = %d but the set holds %lld ratings of that user",
                            (long long)u, tr->user_cnts[u], (long long)seen[u]);
            c->t_off[u + 1] = c->t_off[u] + static_cast<uint32_t>(seen[u]);
        }
    }

    clk.mark("checks and user offsets");
    // tu_mean and day_range (src/tx.rs:465-475): f32 sums in file order; the
    // rows are user-sorted, so each user's sum runs over its own span, in
    // parallel over users
    std::vector<float> tu_mean(U, 0.0f);
    std::vector<int32_t> dmax(U, 0);
    parallel_for(U, [&](int64_t lo, int64_t hi) {
        for (int64_t u = lo; u < hi; ++u) {
            float sum = 0.0f;
            int32_t mx = 0;
            for (int64_t t = c->t_off[u]; t < static_cast<int64_t>(c->t_off[u + 1]); ++t) {
                sum += static_cast<float>(tr->dates[t]);
                mx = std::max<int32_t>(mx, static_cast<int32_t>(tr->dates[t]) + 1);
            }
            tu_mean[u] = tr->user_cnts[u] > 0 ? sum / static_cast<float>(tr->user_cnts[u]) : sum;
            dmax[u] = mx;
        }
    });
    int32_t day_range = 0;
    for (int64_t u = 0; u < U; ++u) day_range = std::max(day_range, dmax[u]);
    if (day_range <= 0) day_range = 1;

    clk.mark("tu_mean, day_range");
    // held-out rows grouped by user, in held-out file order
    c->h_start.assign(U + 1, 0);
    for (int64_t t = 0; t < NH; ++t) c->h_start[pr->user_idxs[t] + 1]++;
    for (int64_t u = 0; u < U; ++u) c->h_start[u + 1] += c->h_start[u];
    std::vector<uint32_t> h_item(NH);
    {
        std::vector<uint32_t> pos(c->h_start.begin(), c->h_start.end() - 1);
        for (int64_t t = 0; t < NH; ++t) h_item[pos[pr->user_idxs[t]]++] = static_cast<uint32_t>(pr->item_idxs[t]);
    }

    clk.mark("held-out grouping");
    // SparseUD (src/tx.rs:31-69): per user the sorted distinct days of train
    // and held-out ratings, with their counts
    std::vector<uint32_t> sud_start(U + 1, 0);
    std::vector<int16_t> sud_date;
    std::vector<uint32_t> sud_cnt;
    {
        std::vector<uint32_t> dstart(U + 1, 0);
        for (int64_t u = 0; u < U; ++u) dstart[u + 1] = dstart[u] + (c->t_off[u + 1] - c->t_off[u]) + (c->h_start[u + 1] - c->h_start[u]);
        std::vector<int16_t> all(dstart[U]);
        std::vector<uint32_t> pos(dstart.begin(), dstart.end() - 1);
        for (int64_t t = 0; t < N; ++t) all[pos[tr->user_idxs[t]]++] = tr->dates[t];
        for (int64_t t = 0; t < NH; ++t) all[pos[pr->user_idxs[t]]++] = pr->dates[t];
        std::vector<uint32_t> ndistinct(U, 0);
        parallel_for(U, [&](int64_t lo, int64_t hi) {
            for (int64_t u = lo; u < hi; ++u) {
                std::sort(all.begin() + dstart[u], all.begin() + dstart[u + 1]);
                uint32_t n = 0;
                for (uint32_t k = dstart[u]; k < dstart[u + 1]; ++k)
                    if (k == dstart[u] || all[k] != all[k - 1]) ++n;
                ndistinct[u] = n;
            }
        });
        for (int64_t u = 0; u < U; ++u) sud_start[u + 1] = sud_start[u] + ndistinct[u];
        c->n_ud = sud_start[U];
        sud_date.resize(c->n_ud);
        sud_cnt.resize(c->n_ud);
        parallel_for(U, [&](int64_t lo, int64_t hi) {
            for (int64_t u = lo; u < hi; ++u) {
                uint32_t w = sud_start[u];
                for (uint32_t k = dstart[u]; k < dstart[u + 1]; ++k) {
                    if (k == dstart[u] || all[k] != all[k - 1]) {
                        sud_date[w] = all[k];
                        sud_cnt[w] = 1;
                        ++w;
                    } else {
                        sud_cnt[w - 1]++;
                    }
                }
            }
        });
    }
    const int64_t D = c->n_ud;
    auto sud_index = [&](int64_t u, int16_t day) -> int64_t {
        auto b = sud_date.begin() + sud_start[u], e = sud_date.begin() + sud_start[u + 1];
        auto it = std::lower_bound(b, e, day);
        return (it != e && *it == day) ? (it - sud_date.begin()) : -1;
    };

    clk.mark("SparseUD");
    // per training rating: item, SparseUD index, time bin, dev, target
    c->t_item.resize(N);
    c->t_ud.resize(N);
    c->t_dev.resize(N);
    c->t_r.resize(N);
    c->t_bin.resize(N);
    if (cfg->has_ordinal) c->t_raw.resize(N);
    if (B > 256) return fail(TX_EUNSUPPORTED, "tx_create: n_time_bins > 256");
    parallel_for(N, [&](int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; ++t) {
            const int32_t u = tr->user_idxs[t];
            const int day = tr->dates[t];
            c->t_item[t] = static_cast<uint32_t>(tr->item_idxs[t]);
            c->t_ud[t] = static_cast<uint32_t>(sud_index(u, tr->dates[t]));
            c->t_dev[t] = dev_host(tu_mean[u], day, cfg->beta);
            c->t_r[t] = tr->residuals[t];
            const int64_t b = (static_cast<int64_t>(day) * B) / day_range;
            c->t_bin[t] = static_cast<uint8_t>(std::min<int64_t>(b, B - 1));
            if (cfg->has_ordinal) c->t_raw[t] = tr->raw_ratings[t];
        }
    });

    clk.mark("per-rating item/ud/bin/dev");
    std::vector<uint8_t>& sud_fbin = c->sud_fbin;
    sud_fbin.assign(D, 0);
    for (int64_t d = 0; d < D; ++d) sud_fbin[d] = static_cast<uint8_t>(freq_bin_host(sud_cnt[d], F));
    if (D >= (int64_t{1} << 28)) return fail(TX_EUNSUPPORTED, "tx_create: 2^28 or more (user, day) pairs");
    for (int64_t u = 0; u < U && c->days_sorted; ++u)
        for (uint32_t t = c->t_off[u] + 1; t < c->t_off[u + 1]; ++t)
            if (c->t_ud[t] < c->t_ud[t - 1]) {
                c->days_sorted = false;
                break;
            }
    // day terms train in the parallel schedule when the ratings are date-ordered
    if (!cfg->low_memory && !c->days_sorted) c->sched.mode = TX_SCHED_SERIAL;
    // neighbourhood terms: in parallel with 1..128 global neighbours
    if (build_nbr && (cfg->max_neighbors < 1 || cfg->max_neighbors > 128 || (shape.waves > 1 && cfg->has_ordinal)))
        c->sched.mode = TX_SCHED_SERIAL;

    clk.mark("freq bins, date order");
    // norms: training su (src/tx.rs:808-819) and epoch-end rebuild (src/tx.rs:410-416)
    std::vector<float> norm_train(U), norm_rebuild(U);
    for (int64_t u = 0; u < U; ++u) {
        const int64_t cnt = c->t_off[u + 1] - c->t_off[u];
        const int64_t tot = cnt + (cfg->full_su ? (c->h_start[u + 1] - c->h_start[u]) : 0);
        norm_train[u] = std::sqrt(static_cast<float>(tot));
        const int64_t rc = static_cast<int64_t>(tr->user_cnts[u]) + pr->user_cnts[u];
        norm_rebuild[u] = rc > 0 ? std::sqrt(static_cast<float>(rc)) : 0.0f;
    }

    // gbias (src/lib.rs:463-466): f64 mean of the residuals, sequential
    double rsum = 0.0;
    for (int64_t t = 0; t < N; ++t) rsum += static_cast<double>(tr->residuals[t]);
    const float gbias = N > 0 ? static_cast<float>(rsum / static_cast<double>(N)) : 0.0f;

    clk.mark("norms, gbias");
    // day-term structures (non-low-memory): held-out / train items per
    // (user, day), same-day runs with their norms
    const bool low_mem = cfg->low_memory != 0;
    std::vector<uint32_t> hd_start, hd_item, td_start, td_item, run_stop;
    std::vector<float> run_norm, sud_norm;
    if (!low_mem) {
        hd_start.assign(D + 1, 0);
        td_start.assign(D + 1, 0);
        std::vector<int64_t> h_ud(NH);
        for (int64_t t = 0; t < NH; ++t) {
            h_ud[t] = sud_index(pr->user_idxs[t], pr->dates[t]);
            hd_start[h_ud[t] + 1]++;
        }
        for (int64_t t = 0; t < N; ++t) td_start[c->t_ud[t] + 1]++;
        for (int64_t d = 0; d < D; ++d) {
            hd_start[d + 1] += hd_start[d];
            td_start[d + 1] += td_start[d];
        }
        hd_item.resize(NH);
        td_item.resize(N);
        {
            std::vector<uint32_t> ph(hd_start.begin(), hd_start.end() - 1), pt(td_start.begin(), td_start.end() - 1);
            for (int64_t t = 0; t < NH; ++t) hd_item[ph[h_ud[t]]++] = static_cast<uint32_t>(pr->item_idxs[t]);
            for (int64_t t = 0; t < N; ++t) td_item[pt[c->t_ud[t]]++] = c->t_item[t];
        }
        sud_norm.resize(D);
        for (int64_t d = 0; d < D; ++d) {
            float cnt = 0.0f;
            for (uint32_t k = hd_start[d]; k < hd_start[d + 1]; ++k) cnt += 1.0f;
            for (uint32_t k = td_start[d]; k < td_start[d + 1]; ++k) cnt += 1.0f;
            sud_norm[d] = cnt > 0.0f ? std::sqrt(cnt) : 0.0f;
        }
        // day_ranges (src/tx.rs:93-105) within each user's span
        run_stop.resize(N);
        run_norm.resize(N);
        for (int64_t u = 0; u < U; ++u) {
            const uint32_t a = c->t_off[u], b = c->t_off[u + 1];
            uint32_t rs = a;
            for (uint32_t t = a; t < b; ++t) {
                if (t + 1 == b || tr->dates[t + 1] != tr->dates[t]) {
                    float day_cnt = static_cast<float>(t - rs + 1);
                    if (cfg->full_su) {
                        const uint32_t d = c->t_ud[rs];
                        for (uint32_t k = hd_start[d]; k < hd_start[d + 1]; ++k) day_cnt += 1.0f;
                    }
                    const float nd = std::sqrt(day_cnt);
                    for (uint32_t k = rs; k <= t; ++k) {
                        run_stop[k] = t;
                        run_norm[k] = nd;
                    }
                    rs = t + 1;
                }
            }
        }
    }

    clk.mark("day-term structures");
    // neighbourhood structures (src/tx.rs:483-585): b~_i and b~_u by FP64 sums
    // over the training rows, r - b~ per training rating, N(u)\R(u) per user
    // (held-out items the user has not rated in training, sorted by item, or
    // by (item, day) with same-day terms, and deduplicated on that key), the
    // two norms, and for the same-day terms a list per SparseUD day: the
    // user's training items of that day, then its N(u)\R(u) items of that day
    c->nb = build_nbr;
    c->nbd = build_nbr_day;
    std::vector<float> t_rmb, norm_nu, norm_ru, inv_sqrt_n, sd_rmb;
    std::vector<uint32_t> imp_start, imp_item, sd_start, sd_item;
    std::vector<int16_t> imp_day;
    if (build_nbr) {
        const double mu = static_cast<double>(gbias);
        std::vector<double> isum(I, 0.0), icnt(I, 0.0), usum(U, 0.0), ucnt(U, 0.0);
        for (int64_t t = 0; t < N; ++t) {
            isum[tr->item_idxs[t]] += static_cast<double>(tr->raw_ratings[t]) - mu;
            icnt[tr->item_idxs[t]] += 1.0;
        }
        std::vector<float> bti(I), btu(U);
        for (int64_t i = 0; i < I; ++i) bti[i] = static_cast<float>(isum[i] / (icnt[i] + static_cast<double>(cfg->lambda1)));
        for (int64_t t = 0; t < N; ++t) {
            usum[tr->user_idxs[t]] += static_cast<double>(tr->raw_ratings[t]) - mu - static_cast<double>(bti[tr->item_idxs[t]]);
            ucnt[tr->user_idxs[t]] += 1.0;
        }
        for (int64_t u = 0; u < U; ++u) btu[u] = static_cast<float>(usum[u] / (ucnt[u] + static_cast<double>(cfg->lambda2)));
        t_rmb.resize(N);
        parallel_for(N, [&](int64_t lo, int64_t hi) {
            for (int64_t t = lo; t < hi; ++t) {
                const float btilde = gbias + bti[tr->item_idxs[t]] + btu[tr->user_idxs[t]];
                t_rmb[t] = static_cast<float>(tr->raw_ratings[t]) - btilde;
            }
        });
        // held-out rows by user with their days (file order), then the filter
        std::vector<int16_t> h_day(NH);
        {
            std::vector<uint32_t> pos(c->h_start.begin(), c->h_start.end() - 1);
            for (int64_t t = 0; t < NH; ++t) h_day[pos[pr->user_idxs[t]]++] = pr->dates[t];
        }
        std::vector<uint32_t> keep(NH);  // item | day bits of the kept rows, per user from h_start[u]
        c->imp_cnt.assign(U, 0);
        parallel_for(U, [&](int64_t lo, int64_t hi) {
            std::vector<uint32_t> rated;
            std::vector<std::pair<uint32_t, int16_t>> v;
            for (int64_t u = lo; u < hi; ++u) {
                rated.assign(c->t_item.begin() + c->t_off[u], c->t_item.begin() + c->t_off[u + 1]);
                std::sort(rated.begin(), rated.end());
                v.clear();
                for (uint32_t k = c->h_start[u]; k < c->h_start[u + 1]; ++k)
                    if (!std::binary_search(rated.begin(), rated.end(), h_item[k])) v.emplace_back(h_item[k], h_day[k]);
                if (build_nbr_day) {
                    std::sort(v.begin(), v.end());
                    v.erase(std::unique(v.begin(), v.end()), v.end());
                } else {
                    // key = item only: which day a duplicate keeps is irrelevant
                    // (days are read by the same-day terms only)
                    std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
                    v.erase(std::unique(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first == b.first; }),
                            v.end());
                }
                uint32_t w = c->h_start[u];
                for (const auto& e : v) {
                    keep[w] = e.first;
                    h_day[w] = e.second;
                    ++w;
                }
                c->imp_cnt[u] = static_cast<uint32_t>(v.size());
            }
        });
        imp_start.assign(U + 1, 0);
        for (int64_t u = 0; u < U; ++u) imp_start[u + 1] = imp_start[u] + c->imp_cnt[u];
        imp_item.resize(imp_start[U]);
        imp_day.resize(imp_start[U]);
        for (int64_t u = 0; u < U; ++u)
            for (uint32_t k = 0; k < c->imp_cnt[u]; ++k) {
                imp_item[imp_start[u] + k] = keep[c->h_start[u] + k];
                imp_day[imp_start[u] + k] = h_day[c->h_start[u] + k];
            }
        // norms with the host powf, as the reference (src/tx.rs:577-583)
        norm_nu.resize(U);
        norm_ru.resize(U);
        uint32_t maxn = 0;
        for (int64_t u = 0; u < U; ++u) {
            const uint32_t n_r = c->t_off[u + 1] - c->t_off[u], n_n = n_r + c->imp_cnt[u];
            norm_ru[u] = n_r > 0 ? powf(static_cast<float>(n_r), -0.5f) : 0.0f;
            norm_nu[u] = n_n > 0 ? powf(static_cast<float>(n_n), -0.5f) : 0.0f;
            maxn = std::max(maxn, n_n);
            if (n_r > 0xffffu || c->imp_cnt[u] > 0xffffu)
                return fail(TX_EUNSUPPORTED, "tx_create: a user with 65536 or more training or N(u)\\R(u) items");
        }
        // same-day counts k -> k^-0.5 (src/tx.rs:925-926)
        inv_sqrt_n.resize(static_cast<size_t>(maxn) + 2);
        for (size_t k = 0; k < inv_sqrt_n.size(); ++k) inv_sqrt_n[k] = k > 0 ? powf(static_cast<float>(k), -0.5f) : 0.0f;
        if (build_nbr_day) {
            sd_start.assign(D + 1, 0);
            for (int64_t t = 0; t < N; ++t) sd_start[c->t_ud[t] + 1]++;
            std::vector<uint32_t> imp_ud(imp_item.size());
            for (int64_t u = 0; u < U; ++u)
                for (uint32_t k = imp_start[u]; k < imp_start[u + 1]; ++k) {
                    const int64_t d = sud_index(u, imp_day[k]);
                    imp_ud[k] = static_cast<uint32_t>(d);
                    sd_start[d + 1]++;
                }
            for (int64_t d = 0; d < D; ++d) sd_start[d + 1] += sd_start[d];
            sd_item.resize(sd_start[D]);
            sd_rmb.resize(sd_start[D]);
            std::vector<uint32_t> pos(sd_start.begin(), sd_start.end() - 1);
            for (int64_t t = 0; t < N; ++t) {
                const uint32_t w = pos[c->t_ud[t]]++;
                sd_item[w] = c->t_item[t] | 0x80000000u;
                sd_rmb[w] = t_rmb[t];
            }
            for (size_t k = 0; k < imp_item.size(); ++k) {
                const uint32_t w = pos[imp_ud[k]]++;
                sd_item[w] = imp_item[k];
                sd_rmb[w] = 0.0f;
            }
        }
    }
    clk.mark("neighbourhood structures");
    draws.t.join();
    clk.mark("initial draws (waited for)");
    // ---- device ----
    HIPC(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    for (auto& e : c->ev) HIPC(hipEventCreate(&e));
    tx::DevModel& m = c->dm;
    m.U = c->U; m.I = c->I; m.f = f; m.fs = c->fs; m.F = F; m.B = B;
    m.day_range = day_range;
    m.low_mem = low_mem ? 1 : 0;
    m.full_su = cfg->full_su ? 1 : 0;
    m.has_ord = cfg->has_ordinal ? 1 : 0;
    m.sum_err_bug = cfg->sum_err_bug ? 1 : 0;
    m.skip_cut = 1;                              // cut is never trained (zero until set)
    m.skip_ibf = cfg->lr_ibf == 0.0f ? 1 : 0;    // zero-initialised, trained only by lr_ibf
    m.skip_but = cfg->lr_tu == 0.0f ? 1 : 0;
    m.skip_yb = cfg->lr_yb == 0.0f ? 1 : 0;
    m.gbias = gbias;
    m.h = make_hyper(*cfg);
    RC(zeros(c, &m.ubias, U));
    // b_i and b_i,bin(t) in one row per item (128-byte aligned rows up to 31 bins)
    m.ibs = static_cast<int32_t>((1 + B + 31) / 32 * 32);
    RC(zeros(c, &m.ibias, I * m.ibs));
    m.bit_bin = m.ibias + 1;
    RC(upload_padded(c, &m.ufeat, h_ufeat.data(), U));
    RC(upload_padded(c, &m.ufeat2, nullptr, U));
    RC(upload_padded(c, &m.ifeat, h_ifeat.data(), I));
    RC(upload_padded(c, &m.yfeat, h_yfeat.data(), I));
    RC(upload_padded(c, &m.ycache, nullptr, U));
    RC(upload_padded(c, &m.yfeat_day, low_mem ? nullptr : h_yday.data(), low_mem ? 0 : I));
    RC(c->upload(&m.tu_mean, tu_mean.data(), U));
    RC(zeros(c, &m.alpha_u, U));
    RC(zeros(c, &m.but_bin, U * B));
    RC(zeros(c, &m.ybias, I));
    RC(zeros(c, &m.ycache_bias, U));
    RC(zeros(c, &m.ubias_day, D));
    RC(upload_padded(c, &m.ufeat_day, nullptr, low_mem ? 0 : D));
    RC(upload_padded(c, &m.ycache_day, nullptr, low_mem ? 0 : D));
    RC(zeros(c, &m.ibias_freq, I * F));
    RC(upload_padded(c, &m.ifeat_freq, h_iqf.data(), I * F));
    {
        std::vector<float> ones(U, 1.0f);
        RC(c->upload(&m.cu, ones.data(), U));
    }
    RC(zeros(c, &m.cut, D));
    RC(c->upload(&m.thresholds, cfg->th_init, 4));

    clk.mark("parameter tables to device");
    uint32_t* p32;
    RC(c->upload(&p32, sud_start.data(), sud_start.size())); m.sud_start = p32;
    int16_t* p16;
    RC(c->upload(&p16, sud_date.data(), sud_date.size())); m.sud_date = p16;
    uint8_t* p8;
    RC(c->upload(&p8, sud_fbin.data(), sud_fbin.size())); m.sud_fbin = p8;
    float* pf;
    RC(c->upload(&pf, sud_norm.data(), sud_norm.size())); m.sud_norm = pf;
    RC(c->upload(&p32, c->t_off.data(), c->t_off.size())); m.t_off = p32;
    RC(c->upload(&p32, c->t_item.data(), c->t_item.size())); m.t_item = p32;
    RC(c->upload(&p32, c->t_ud.data(), c->t_ud.size())); m.t_ud = p32;
    RC(c->upload(&pf, c->t_dev.data(), c->t_dev.size())); m.t_dev = pf;
    RC(c->upload(&pf, c->t_r.data(), c->t_r.size())); m.t_r = pf;
    RC(c->upload(&p8, c->t_bin.data(), c->t_bin.size())); m.t_bin = p8;
    int8_t* pi8;
    RC(c->upload(&pi8, c->t_raw.data(), c->t_raw.size())); m.t_raw = pi8;
    {
        if (c->I > (1 << 20)) return fail(TX_EUNSUPPORTED, "tx_create: more than 2^20 items");
        std::vector<uint4> rec(N);
        parallel_for(N, [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k) rec[k] = pack_rec(c, k);
        });
        RC(c->upload(&c->d_rec, rec.data(), rec.size()));
    }
    clk.mark("ratings and records to device");
    RC(c->upload(&p32, run_stop.data(), run_stop.size())); m.t_run_stop = p32;
    RC(c->upload(&pf, run_norm.data(), run_norm.size())); m.t_run_norm = pf;
    RC(c->upload(&p32, c->h_start.data(), c->h_start.size())); m.h_start = p32;
    RC(c->upload(&p32, h_item.data(), h_item.size())); m.h_item = p32;
    RC(c->upload(&p32, hd_start.data(), hd_start.size())); m.hd_start = p32;
    RC(c->upload(&p32, hd_item.data(), hd_item.size())); m.hd_item = p32;
    RC(c->upload(&p32, td_start.data(), td_start.size())); m.td_start = p32;
    RC(c->upload(&p32, td_item.data(), td_item.size())); m.td_item = p32;
    if (!low_mem && c->days_sorted) {
        // day streams of the parallel day-term kernel (see DevModel::dg_item)
        std::vector<uint32_t> dg_off(U + 1, 0), dy_off(U + 1, 0);
        for (int64_t u = 0; u < U; ++u) {
            uint32_t ng = 0, ny = 0;
            for (uint32_t d = sud_start[u]; d < sud_start[u + 1]; ++d) {
                const uint32_t nt = td_start[d + 1] - td_start[d];
                if (nt == 0) continue;
                ng += nt + (cfg->full_su ? hd_start[d + 1] - hd_start[d] : 0) + 1;
                ny += nt + 1;
            }
            if (uint64_t{dg_off[u]} + ng > 0xffffffffu) return fail(TX_EUNSUPPORTED, "tx_create: day stream over 2^32 rows");
            dg_off[u + 1] = dg_off[u] + ng;
            dy_off[u + 1] = dy_off[u] + ny;
        }
        std::vector<uint32_t> dg(dg_off[U]), dy(dy_off[U]);
        parallel_for(U, [&](int64_t lo, int64_t hi) {
            for (int64_t u = lo; u < hi; ++u) {
                uint32_t* g = dg.data() + dg_off[u];
                uint32_t* y = dy.data() + dy_off[u];
                for (uint32_t d = sud_start[u]; d < sud_start[u + 1]; ++d) {
                    if (td_start[d + 1] == td_start[d]) continue;
                    *y++ = d | 0x80000000u;
                    for (uint32_t k = td_start[d]; k < td_start[d + 1]; ++k) *g++ = *y++ = td_item[k];
                    if (cfg->full_su)
                        for (uint32_t k = hd_start[d]; k < hd_start[d + 1]; ++k) *g++ = hd_item[k];
                    *g++ = d | 0x80000000u;
                }
            }
        });
        RC(c->upload(&p32, dg_off.data(), dg_off.size())); m.dg_off = p32;
        RC(c->upload(&p32, dg.data(), dg.size())); m.dg_item = p32;
        RC(c->upload(&p32, dy_off.data(), dy_off.size())); m.dy_off = p32;
        RC(c->upload(&p32, dy.data(), dy.size())); m.dy_item = p32;
    }
    // neighbourhood terms: lists, norms and the dense I x I tables (zero)
    m.nb_on = build_nbr ? 1 : 0;
    m.nbd_on = (build_nbr && build_nbr_day) ? 1 : 0;
    m.nb_max = cfg->max_neighbors;
    if (build_nbr) {
        RC(c->upload(&pf, t_rmb.data(), t_rmb.size())); m.t_rmb = pf;
        RC(c->upload(&p32, imp_start.data(), imp_start.size())); m.imp_start = p32;
        RC(c->upload(&p32, imp_item.data(), imp_item.size())); m.imp_item = p32;
        RC(c->upload(&p16, imp_day.data(), imp_day.size())); m.imp_day = p16;
        RC(c->upload(&pf, norm_nu.data(), norm_nu.size())); m.norm_nu = pf;
        RC(c->upload(&pf, norm_ru.data(), norm_ru.size())); m.norm_ru = pf;
        RC(c->upload(&pf, inv_sqrt_n.data(), inv_sqrt_n.size())); m.inv_sqrt_n = pf;
        m.n_isq = static_cast<int32_t>(inv_sqrt_n.size());
        RC(c->upload(&p32, sd_start.data(), sd_start.size())); m.sd_start = p32;
        RC(c->upload(&p32, sd_item.data(), sd_item.size())); m.sd_item = p32;
        RC(c->upload(&pf, sd_rmb.data(), sd_rmb.size())); m.sd_rmb = pf;
        RC(zeros(c, &m.nw, I * I));
        RC(zeros(c, &m.nc, I * I));
        // the epoch's samples: U words of counts and U x max_neighbors positions
        const int64_t slots = U * std::max<int64_t>(cfg->max_neighbors, 1);
        for (int b = 0; b < 2; ++b) {
            HIPC(hipHostMalloc(reinterpret_cast<void**>(&c->h_smp_cnt[b]), sizeof(uint32_t) * std::max<int64_t>(U, 1)));
            HIPC(hipHostMalloc(reinterpret_cast<void**>(&c->h_smp[b]), sizeof(uint16_t) * std::max<int64_t>(slots, 1)));
        }
        RC(c->alloc(&c->d_smp_cnt, static_cast<size_t>(U)));
        RC(c->alloc(&c->d_smp, static_cast<size_t>(slots)));
        m.smp_cnt = c->d_smp_cnt;
        m.smp = c->d_smp;
    }
    if (build_nbr_day) {
        RC(zeros(c, &m.nw_day, I * I));
        RC(zeros(c, &m.nc_day, I * I));
    }
    RC(c->upload(&pf, norm_train.data(), norm_train.size())); m.norm_train = pf;
    RC(c->upload(&pf, norm_rebuild.data(), norm_rebuild.size())); m.norm_rebuild = pf;
    RC(c->alloc(&c->d_order, static_cast<size_t>(U)));
    RC(c->alloc(&c->d_scratch, static_cast<size_t>(4) * c->fs + 4));
    // the device's tx_predict_noalloc scratch array (its address looked up
    // once per device, with the device current)
    HIPC(hipGetDevice(&c->device));
    HIPC(tx::noalloc_scratch(c->device, &c->d_noalloc));
    HIPC(hipDeviceSynchronize());
    clk.mark("remaining uploads");
    // default schedule: split teams (a rating wave and a streaming wave per
    // user stream) where measured faster at the same fidelity, the 64-lane
    // four-element shape (f = 129..256) of low-memory models without ordinal
    // head or neighbourhood terms (full shape, f=256: 107.8 against 110.3 ms
    // per epoch, 20-epoch holdout RMSE 1.9e-4 above the serial trainer's)
    c->sched.split_waves = (cfg->low_memory && !cfg->has_ordinal && !c->nb && c->shape.team == 64 &&
                            c->shape.vec == 4 && c->shape.waves == 1 && c->sched.mode == TX_SCHED_STRATA)
                               ? 1
                               : 0;
    *out = ctx.release();
    return 0;
}

int tx_destroy(tx_ctx* ctx) {
    g_err.clear();
    delete ctx;
    return 0;
}

int tx_set_schedule(tx_ctx* c, const tx_schedule* s) {
    g_err.clear();
    if (!c || !s) return fail(TX_EINVAL, "tx_set_schedule: NULL argument");
    if (s->mode != TX_SCHED_SERIAL && s->mode != TX_SCHED_STRATA && s->mode != TX_SCHED_TSVDX4P &&
        s->mode != TX_SCHED_PLAIN)
        return fail(TX_EINVAL, "unknown schedule mode %d", s->mode);
    if (s->mode == TX_SCHED_PLAIN) {
        if (!c->cfg.low_memory || c->cfg.has_ordinal || c->shape.waves > 1)
            return fail(TX_EUNSUPPORTED, "the plain schedule trains low_memory models without ordinal head, f <= 1024");
        if (s->blocks < 0) return fail(TX_EINVAL, "blocks must be 0 (1024) or positive");
    }
    if (s->mode == TX_SCHED_TSVDX4P) {
        if (!c->cfg.low_memory || c->cfg.has_ordinal)
            return fail(TX_EUNSUPPORTED, "the tsvdx4p schedule trains low_memory models without ordinal head");
        if (s->blocks < 0) return fail(TX_EINVAL, "blocks must be 0 (auto) or positive");
    }
    if (s->mode == TX_SCHED_STRATA) {
        if (s->n_strata < 1 || (s->n_strata > 1 && s->n_strata % 8 != 0) || s->n_strata > 64)
            return fail(TX_EINVAL, "n_strata must be 1 or a multiple of 8 (<= 64), got %d", s->n_strata);
        if (!c->cfg.low_memory && s->n_strata != 1)
            return fail(TX_EUNSUPPORTED, "day terms (low_memory = false) train in the one-stratum parallel schedule");
        if (c->shape.waves > 1 && s->n_strata != 1)
            return fail(TX_EUNSUPPORTED, "n_feat > 1024 trains in the one-stratum parallel schedule");
        if (!c->cfg.low_memory && !c->days_sorted)
            return fail(TX_EUNSUPPORTED,
                        "the parallel schedule trains day terms when every user's training ratings are sorted by "
                        "date; use TX_SCHED_SERIAL");
        if (c->cfg.has_ordinal && c->cfg.lr_t != 0.0f)
            return fail(TX_EUNSUPPORTED, "the parallel schedule keeps ordinal thresholds fixed (lr_t must be 0)");
        if (s->blocks < 0 || (s->blocks > 0 && s->blocks < s->n_strata))
            return fail(TX_EINVAL, "blocks must be 0 (auto) or at least n_strata");
    }
    if (c->nb && s->mode != TX_SCHED_SERIAL) {
        // the neighbourhood terms train in the serial schedule or the
        // one-stratum parallel one (two global neighbours per lane)
        if (s->mode != TX_SCHED_STRATA || s->n_strata != 1)
            return fail(TX_EUNSUPPORTED, "the neighbourhood terms train in TX_SCHED_SERIAL or one-stratum TX_SCHED_STRATA");
        if (c->cfg.max_neighbors < 1 || c->cfg.max_neighbors > 128)
            return fail(TX_EUNSUPPORTED, "the parallel schedule holds at most 128 global neighbours (max_neighbors 1..128)");
        if (c->shape.waves > 1 && c->cfg.has_ordinal)
            return fail(TX_EUNSUPPORTED, "neighbourhood terms with an ordinal head train in parallel up to n_feat 1024");
    }
    if (s->user_order < 0 || s->user_order > 3) return fail(TX_EINVAL, "user_order must be 0, 1, 2 or 3");
    if (s->user_order == 3 && (s->mode != TX_SCHED_STRATA || s->n_strata != 1 || c->shape.waves > 1))
        return fail(TX_EUNSUPPORTED, "length-bucketed launches (user_order 3): one-stratum TX_SCHED_STRATA, f <= 1024");
    if (s->teams_per_block < 0 || s->teams_per_block > 4)
        return fail(TX_EINVAL, "teams_per_block must be 0 (4) or 1..4, got %d", s->teams_per_block);
    if (s->heavy_min < 0) return fail(TX_EINVAL, "heavy_min must be >= 0");
    const int qm = s->update_flags & 3, ym = (s->update_flags >> 2) & 3, nm = (s->update_flags >> 5) & 3;
    if (qm > 2 || ym > 2 || nm == 1 || nm > 2 || (s->update_flags >> 7) != 0)
        return fail(TX_EINVAL, "bad update_flags %d", s->update_flags);
    if (s->split_waves != 0 && s->split_waves != 1) return fail(TX_EINVAL, "split_waves must be 0 or 1");
    if (s->split_waves &&
        (s->mode != TX_SCHED_STRATA || s->n_strata != 1 || s->user_order == 3 || ym != 2 || !c->cfg.low_memory ||
         c->cfg.has_ordinal || c->nb || c->shape.team != 64 || c->shape.waves > 1))
        return fail(TX_EUNSUPPORTED,
                    "split_waves: one-stratum TX_SCHED_STRATA, user_order 0-2, plain NSVD1 rows, a low-memory model "
                    "without ordinal head or neighbourhood terms, f <= 1024");
    const int want_S = s->mode == TX_SCHED_TSVDX4P ? 1 : s->n_strata;
    const bool relayout = s->mode != TX_SCHED_SERIAL && s->mode != TX_SCHED_PLAIN && want_S != c->built_S;
    c->sched = *s;
    if (s->mode == TX_SCHED_TSVDX4P) c->sched.n_strata = 1;
    if (relayout) {
        if (int rc = build_strata(c, want_S)) return rc;
    }
    c->sv.q_mode = qm;
    c->sv.y_mode = ym;
    c->sv.nb_mode = nm;
    if (s->user_order == 3)
        if (int rc = init_buckets(c)) return rc;
    c->sv.block_threads = (s->teams_per_block > 0 ? s->teams_per_block : 4) * 64;
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        c->graph_S = 0;
    }
    return 0;
}

int tx_fit_epoch(tx_ctx* c, int32_t epoch) {
    g_err.clear();
    if (!c) return fail(TX_EINVAL, "tx_fit_epoch: NULL ctx");
    c->bf16_stale = true;
    if (epoch < 1) return fail(TX_EINVAL, "tx_fit_epoch: epochs are 1-based");
    hipStream_t st = c->st;
    tx_epoch_stats stats{};
    stats.n_ratings = c->N;
    HIPC(hipEventRecord(c->ev[0], st));
    if (epoch == c->cfg.reset_u_epoch) {
        HIPC(tx::launch_fill_f32(c->dm.ufeat, 0.0f, static_cast<int64_t>(c->U) * c->fs, st));
        HIPC(tx::launch_fill_f32(c->dm.ufeat2, 0.0f, static_cast<int64_t>(c->U) * c->fs, st));
    }
    if (c->nb && c->smp_epoch != epoch)
        if (int rc = upload_nbr_samples(c, epoch, st)) return rc;
    bool seq_epoch = false;  // Tsvdx4pConfig::seq_epochs run the TxModel epoch (src/tsvdx4p.rs:799-800)
    for (int k = 0; k < 5; ++k) seq_epoch = seq_epoch || c->cfg.seq_epochs[k] == epoch;
    if (c->sched.mode == TX_SCHED_TSVDX4P && !seq_epoch) {
        if (int rc = tsvdx4p_epoch(c, epoch, stats, st)) return rc;
        HIPC(hipEventRecord(c->ev[3], st));
        HIPC(hipEventSynchronize(c->ev[3]));
        float t = 0;
        HIPC(hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
        stats.ms_total = t;
        stats.ms_other = t - stats.ms_sgd - stats.ms_rebuild;
    } else if (c->sched.mode == TX_SCHED_PLAIN) {
        std::vector<int64_t> order(c->U);
        tx::user_order(c->U, c->cfg.shuffle_users != 0, c->cfg.seed, epoch, order.data());
        HIPC(hipMemcpyAsync(c->d_order, order.data(), sizeof(int64_t) * c->U, hipMemcpyHostToDevice, st));
        const bool qf = c->cfg.lr_iqf != 0.0f || c->cfg.sigma_iqf != 0.0f;
        HIPC(hipEventRecord(c->ev[1], st));
        HIPC(tx::launch_sgd_plain(c->dm, c->d_order, c->shape.vec, c->sched.blocks > 0 ? c->sched.blocks : 1024, qf, st));
        HIPC(hipEventRecord(c->ev[2], st));
        HIPC(tx::launch_rebuild(c->dm, c->n_ud, c->shape.team, c->shape.vec, st));
        HIPC(hipEventRecord(c->ev[3], st));
        HIPC(hipEventSynchronize(c->ev[3]));
        float a = 0, b = 0, t = 0;
        HIPC(hipEventElapsedTime(&a, c->ev[1], c->ev[2]));
        HIPC(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
        HIPC(hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
        stats.ms_sgd = a;
        stats.ms_rebuild = b;
        stats.ms_total = t;
        stats.ms_other = t - a - b;
        stats.n_launches = 2;
    } else if (c->sched.mode == TX_SCHED_SERIAL) {
        std::vector<int64_t> order(c->U);
        tx::user_order(c->U, c->cfg.shuffle_users != 0, c->cfg.seed, epoch, order.data());
        HIPC(hipMemcpyAsync(c->d_order, order.data(), sizeof(int64_t) * c->U, hipMemcpyHostToDevice, st));
        HIPC(hipEventRecord(c->ev[1], st));
        HIPC(tx::launch_sgd_serial(c->dm, c->d_order, c->d_scratch, st));
        HIPC(hipEventRecord(c->ev[2], st));
        HIPC(tx::launch_rebuild(c->dm, c->n_ud, c->shape.team, c->shape.vec, st));
        HIPC(hipEventRecord(c->ev[3], st));
        if (int rc = draw_next_nbr_samples(c, epoch)) return rc;
        HIPC(hipEventSynchronize(c->ev[3]));
        float a = 0, b = 0, t = 0;
        HIPC(hipEventElapsedTime(&a, c->ev[1], c->ev[2]));
        HIPC(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
        HIPC(hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
        stats.ms_sgd = a;
        stats.ms_rebuild = b;
        stats.ms_total = t;
        stats.ms_other = t - a - b;
        stats.n_launches = 2;
        if (int rc = enforce_thresholds(c)) return rc;
    } else {
        if (c->built_S == 0)
            if (int rc = build_strata(c, c->sched.n_strata > 0 ? c->sched.n_strata : 1)) return rc;
        if (c->built_S == 1 && c->sched.user_order != 1)
            if (int rc = prepare_user_queue(c, epoch, st)) return rc;
        if (c->sched.use_graph) {
            if (!c->graph || c->graph_S != c->built_S) {
                if (c->graph) (void)hipGraphExecDestroy(c->graph);
                c->graph = nullptr;
                hipGraph_t g;
                HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                int rc = enqueue_strata_epoch(c, st);
                hipError_t e = hipStreamEndCapture(st, &g);
                if (rc) return rc;
                if (e != hipSuccess) return fail(TX_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
                e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) return fail(TX_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
                c->graph_S = c->built_S;
            }
            HIPC(hipEventRecord(c->ev[1], st));
            HIPC(hipGraphLaunch(c->graph, st));
            HIPC(hipEventRecord(c->ev[3], st));
            if (c->built_S == 1 && c->sched.user_order != 1) build_next_queue(c, epoch + 1);
            if (int rc = draw_next_nbr_samples(c, epoch)) return rc;
            HIPC(hipEventSynchronize(c->ev[3]));
            float a = 0, t = 0;
            HIPC(hipEventElapsedTime(&a, c->ev[1], c->ev[3]));
            HIPC(hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
            stats.ms_sgd = a;
            stats.ms_total = t;
            stats.graph = 1;
            stats.n_launches = c->built_S + 2;
        } else {
            const int S = c->built_S;
            const bool qf = c->cfg.lr_iqf != 0.0f || c->cfg.sigma_iqf != 0.0f;
            const bool p2 = c->cfg.lr_u2 != 0.0f;
            const bool ord = c->cfg.has_ordinal != 0;
            const bool day = c->cfg.low_memory == 0;
            const int blocks = c->sched.blocks > 0 ? c->sched.blocks : default_blocks(c);
            HIPC(tx::launch_fill_u32(c->sv.q_head, 0u, static_cast<int64_t>(S) * c->sv.n_cells, st));
            HIPC(hipEventRecord(c->ev[1], st));
            const bool bucketed = c->sched.user_order == 3 && S == 1;
            if (bucketed)
                RC(enqueue_buckets(c, st, true));
            else if (c->sched.split_waves)
                HIPC(tx::launch_sgd_par_split(c->dm, c->sv, c->shape.vec, blocks, qf, p2, st));
            else if (c->shape.waves > 1)
                HIPC(tx::launch_sgd_par_mw(c->dm, c->sv, blocks, c->shape.waves, qf, p2, ord, day, st));
            else
                for (int s = 0; s < S; ++s)
                    HIPC(tx::launch_sgd_par(c->dm, c->sv, s, c->shape.team, c->shape.vec, blocks, qf, p2, ord, 0, day,
                                            st));
            if (S > 1)
                HIPC(tx::launch_nsvd_gather(c->dm, c->iv, c->sv.gacc, c->sv.gacc_b, c->d_scale_strata, c->shape.vec,
                                            st));
            HIPC(hipEventRecord(c->ev[2], st));
            HIPC(tx::launch_rebuild(c->dm, c->n_ud, c->shape.team, c->shape.vec, st));
            HIPC(hipEventRecord(c->ev[3], st));
            if (c->built_S == 1 && c->sched.user_order != 1) build_next_queue(c, epoch + 1);
            if (int rc = draw_next_nbr_samples(c, epoch)) return rc;
            HIPC(hipEventSynchronize(c->ev[3]));
            float a = 0, b = 0, t = 0;
            HIPC(hipEventElapsedTime(&a, c->ev[1], c->ev[2]));
            HIPC(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
            HIPC(hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
            stats.ms_sgd = a;
            stats.ms_rebuild = b;
            stats.ms_total = t;
            stats.ms_other = t - a - b;
            stats.n_launches = S + 2;
            if (bucketed) {
                RC(read_bucket_times(c));
                stats.n_launches = TX_N_BUCKETS + 2;
            }
        }
        // OrdinalHead::enforce_sorted_with_gap at epoch end (src/tx.rs:1113):
        // the parallel schedule keeps lr_t = 0, so this only acts on th_init
        if (int rc = enforce_thresholds(c)) return rc;
    }
    c->last = stats;
    c->ycache_current = true;  // every schedule's epoch ends with the rebuild
    return 0;
}

int tx_fit_users_epoch(tx_ctx* c, int32_t epoch) {
    g_err.clear();
    if (!c) return fail(TX_EINVAL, "tx_fit_users_epoch: NULL ctx");
    c->bf16_stale = true;
    if (epoch < 1) return fail(TX_EINVAL, "tx_fit_users_epoch: epochs are 1-based");
    if (!c->cfg.low_memory || c->cfg.has_ordinal || c->nb)
        return fail(TX_EUNSUPPORTED, "tx_fit_users_epoch: low-memory models without ordinal head or neighbourhood terms");
    if (c->built_S > 1) return fail(TX_EUNSUPPORTED, "tx_fit_users_epoch: the model uses a stratified layout");
    if (c->built_S == 0)
        if (int rc = build_strata(c, 1)) return rc;
    hipStream_t st = c->st;
    tx_epoch_stats stats{};
    stats.n_ratings = c->N;
    HIPC(hipEventRecord(c->ev[0], st));
    if (int rc = prepare_user_queue(c, epoch, st)) return rc;
    const bool qf = c->cfg.lr_iqf != 0.0f || c->cfg.sigma_iqf != 0.0f;
    const bool p2 = c->cfg.lr_u2 != 0.0f;
    HIPC(tx::launch_fill_u32(c->sv.q_head, 0u, c->sv.n_cells, st));
    tx::DevModel dm = c->dm;
    tx::StrataView sv = c->sv;
    sv.su_cached = (c->cfg.full_su && c->ycache_current) ? 1 : 0;
    if (c->bf16_frozen) {
        dm.bf16_frozen = 1;
        HIPC(tx::launch_to_bf16(c->dm.ifeat, c->b_ifeat, static_cast<int64_t>(c->I) * c->fs, st));
        if (qf) HIPC(tx::launch_to_bf16(c->dm.ifeat_freq, c->b_ifeat_freq, static_cast<int64_t>(c->I) * c->F * c->fs, st));
    }
    HIPC(hipEventRecord(c->ev[1], st));
    const int blocks = c->sched.blocks > 0 && !c->sched.split_waves ? c->sched.blocks : user_pass_blocks(c);
    if (c->shape.waves > 1)
        HIPC(tx::launch_sgd_par_mw(dm, sv, blocks, c->shape.waves, qf, p2, false, false, st, 2));
    else
        HIPC(tx::launch_sgd_par(dm, sv, 0, c->shape.team, c->shape.vec, blocks, qf, p2, false, 2, false, st));
    HIPC(hipEventRecord(c->ev[2], st));
    HIPC(tx::launch_rebuild(c->dm, c->n_ud, c->shape.team, c->shape.vec, st));
    HIPC(hipEventRecord(c->ev[3], st));
    HIPC(hipEventSynchronize(c->ev[3]));
    float a = 0, b = 0, t = 0;
    HIPC(hipEventElapsedTime(&a, c->ev[1], c->ev[2]));
    HIPC(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
    HIPC(hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
    stats.ms_sgd = a;
    stats.ms_rebuild = b;
    stats.ms_total = t;
    stats.ms_other = t - a - b;
    stats.n_launches = 2;
    c->last = stats;
    c->ycache_current = true;
    return 0;
}

int tx_autotune(tx_ctx* c, int32_t trial_epochs, tx_schedule* best_out) {
    g_err.clear();
    if (!c) return fail(TX_EINVAL, "tx_autotune: NULL ctx");
    c->ycache_current = false;  // the trial epochs end with a restore of every table
    if (c->sched.mode != TX_SCHED_STRATA || c->sched.n_strata != 1)
        return fail(TX_EUNSUPPORTED, "tx_autotune tunes the one-stratum parallel schedule");
    if (trial_epochs < 1) trial_epochs = 1;
    hipStream_t st = c->st;
    // snapshot of every parameter table (device to device)
    struct Snap { float* src; float* copy; size_t n; };
    std::vector<Snap> snaps;
    auto release = [&]() {
        for (auto& s : snaps) (void)hipFree(s.copy);
        snaps.clear();
    };
    for (int id = 0; id < TX_P_COUNT; ++id) {
        const ParamDesc pd = param_desc(c, id);
        if (!pd.p || pd.rows <= 0) continue;
        // from the first element to the last (the item-bias rows hold two tables)
        const size_t n = static_cast<size_t>(pd.rows - 1) * pd.pitch + pd.cols;
        float* cp = nullptr;
        if (hipMalloc(&cp, n * sizeof(float)) != hipSuccess) {
            release();
            return fail(TX_EDEVICE, "tx_autotune: no device memory for the parameter snapshot");
        }
        snaps.push_back({pd.p, cp, n});
        if (hipMemcpyAsync(cp, pd.p, n * sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess) {
            release();
            return fail(TX_EDEVICE, "tx_autotune: snapshot copy failed");
        }
    }
    auto restore = [&]() -> int {
        for (auto& s : snaps)
            HIPC(hipMemcpyAsync(s.src, s.copy, s.n * sizeof(float), hipMemcpyDeviceToDevice, st));
        HIPC(hipStreamSynchronize(st));
        return 0;
    };
    const tx_schedule orig = c->sched;
    const int tpb = teams_per_block(c), cap = target_teams(c);
    if (orig.user_order == 3) {
        // length-bucketed launches: per bucket (most ratings first) its team
        // shape and teams in flight, one at a time with the others held, each
        // trial an epoch from the snapshot; teams in all stay within the
        // fidelity bound
        if (int rc = init_buckets(c)) { release(); return rc; }
        tx_schedule s1 = orig;
        s1.use_graph = 0;
        if (int rc = tx_set_schedule(c, &s1)) { release(); return rc; }
        c->trials.clear();
        auto trial = [&](double* ms) -> int {
            if (int rc = restore()) return rc;
            if (int rc = tx_fit_epoch(c, 1)) return rc;
            *ms = c->last.ms_total;
            return 0;
        };
        auto record = [&](int b, double ms) {
            tx_autotune_trial t{};
            t.bucket = b;
            t.n_strata = 1;
            t.blocks = c->sched.blocks;
            t.user_order = 3;
            t.heavy_min = c->sched.heavy_min;
            if (b >= 0) {
                t.lanes_per_user = c->bk[b].lanes;
                t.waves_per_user = c->bk[b].waves;
                t.teams = c->bk[b].teams;
            }
            t.ms = ms;
            c->trials.push_back(t);
        };
        int order[TX_N_BUCKETS];
        std::iota(order, order + TX_N_BUCKETS, 0);
        std::sort(order, order + TX_N_BUCKETS, [&](int a, int b) { return c->bk[a].ratings > c->bk[b].ratings; });
        double best_ms = 0.0;
        int rc = trial(&best_ms);
        if (!rc) record(-1, best_ms);
        for (int b : order) {
            if (rc) break;
            auto& k = c->bk[b];
            if (k.users == 0) continue;
            const tx_ctx::Bucket keep = k;
            tx_ctx::Bucket best = k;
            int others = 0;
            for (int o = 0; o < TX_N_BUCKETS; ++o)
                if (o != b) others += c->bk[o].teams;
            std::vector<std::pair<int, int>> shapes = {{keep.lanes, keep.waves}};
            if (bucket_shapes_ok(c)) {
                if (c->fs <= 256 && keep.lanes != 16) shapes.push_back({16, 1});
                if (keep.lanes != 64 || keep.waves != 1) shapes.push_back({64, 1});
                if (c->fs >= 128 && c->fs <= 1024 && keep.waves != 2) shapes.push_back({64, 2});
            }
            for (const auto& sh : shapes)
                for (double mult : {0.5, 1.0, 1.5}) {
                    if (rc) break;
                    const int t = std::max(1, static_cast<int>(std::min<int64_t>(keep.teams * mult + 0.5, k.users)));
                    if (sh.first == keep.lanes && sh.second == keep.waves && t == keep.teams) continue;
                    if (others + t > cap) continue;
                    k.lanes = sh.first;
                    k.waves = sh.second;
                    k.teams = t;
                    double ms = 0.0;
                    if ((rc = trial(&ms))) break;
                    record(b, ms);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = k;
                    }
                }
            k = best;
        }
        // the chosen shape of each bucket
        for (auto& t : c->trials)
            t.chosen = t.bucket >= 0 && t.lanes_per_user == c->bk[t.bucket].lanes &&
                       t.waves_per_user == c->bk[t.bucket].waves && t.teams == c->bk[t.bucket].teams;
        const int rc2 = restore();
        release();
        if (rc) return rc;
        if (rc2) return rc2;
        if (int r = tx_set_schedule(c, &orig)) return r;  // back to the caller's graph setting
        c->bf16_stale = true;
        if (best_out) *best_out = orig;
        return 0;
    }
    std::vector<tx_schedule> cands;
    // team form: one wave per team, and split teams where the model takes them
    const bool split_ok = c->cfg.low_memory && !c->cfg.has_ordinal && !c->nb && c->shape.team == 64 &&
                          c->shape.waves == 1 && ((orig.update_flags >> 2) & 3) == 2;
    for (int split : {0, 1}) {
        if (split && !split_ok) continue;
        const int tpb_s = split ? 1 : (c->shape.waves > 1 ? 1 : (orig.teams_per_block > 0 ? orig.teams_per_block
                                                                                         : 256 / c->shape.team));
        for (int quarters : {2, 3, 4}) {
            int blocks = (cap * quarters / 4 + tpb_s - 1) / tpb_s;
            blocks = std::max(8, (blocks + 7) / 8 * 8);
            const int orders[4][2] = {{0, 0}, {2, 256}, {2, 1024}, {2, 4096}};
            for (const auto& o : orders) {
                tx_schedule s = orig;
                s.split_waves = split;
                s.blocks = blocks;
                s.user_order = o[0];
                s.heavy_min = o[1];
                s.use_graph = 0;
                cands.push_back(s);
            }
        }
    }
    (void)tpb;
    double best_ms = 1e30;
    tx_schedule best = orig;
    int rc = 0;
    c->trials.clear();
    auto run = [&](const tx_schedule& s) -> int {
        if (int r = tx_set_schedule(c, &s)) return r;
        std::vector<double> ms;
        for (int e = 1; e <= trial_epochs; ++e) {
            if (int r = restore()) return r;
            if (int r = tx_fit_epoch(c, e)) return r;
            ms.push_back(c->last.ms_total);
        }
        std::sort(ms.begin(), ms.end());
        const double med = ms[ms.size() / 2];
        tx_autotune_trial t{};
        t.bucket = -1;
        t.split_waves = s.split_waves;
        t.n_strata = s.n_strata;
        t.blocks = c->sched.blocks > 0 ? c->sched.blocks : default_blocks(c);
        t.user_order = s.user_order;
        t.heavy_min = s.heavy_min;
        t.teams = t.blocks * teams_per_block(c);
        t.ms = med;
        c->trials.push_back(t);
        if (med < best_ms) {
            best_ms = med;
            best = s;
        }
        return 0;
    };
    for (const auto& s : cands)
        if ((rc = run(s))) break;
    // the stratum grid: the fastest one-stratum candidate over 8 and 64
    // user x item blocks (count-balanced, one launch per stratum, NSVD1 steps
    // gathered per item after the last stratum) where the model supports it
    if (!rc && best.n_strata == 1 && c->cfg.low_memory && !c->nb && !c->cfg.has_ordinal) {
        const tx_schedule one = best;
        for (int S : {8, 64}) {
            tx_schedule s = one;
            s.n_strata = S;
            s.blocks = 0;
            s.split_waves = 0;  // the stratified schedule runs one-wave teams
            if ((rc = run(s)) == TX_EUNSUPPORTED) {  // a grid this model cannot take
                rc = 0;
                g_err.clear();
                continue;
            }
            if (rc) break;
        }
    }
    for (auto& t : c->trials)
        t.chosen = t.n_strata == best.n_strata && t.user_order == best.user_order && t.heavy_min == best.heavy_min &&
                   t.split_waves == best.split_waves && (best.n_strata > 1 || t.blocks == best.blocks);
    const int rc2 = restore();
    release();
    if (rc) return rc;
    if (rc2) return rc2;
    best.use_graph = orig.use_graph;
    if (int r = tx_set_schedule(c, &best)) return r;
    c->bf16_stale = true;
    if (best_out) *best_out = best;
    return 0;
}

int tx_get_schedule(tx_ctx* c, tx_schedule* out) {
    g_err.clear();
    if (!c || !out) return fail(TX_EINVAL, "tx_get_schedule: NULL argument");
    *out = c->sched;
    return 0;
}

int tx_get_autotune_trials(tx_ctx* c, tx_autotune_trial* out, int32_t max, int32_t* n) {
    g_err.clear();
    if (!c || !n || (max > 0 && !out)) return fail(TX_EINVAL, "tx_get_autotune_trials: NULL argument");
    *n = static_cast<int32_t>(c->trials.size());
    for (int32_t k = 0; k < max && k < *n; ++k) out[k] = c->trials[static_cast<size_t>(k)];
    return 0;
}

int tx_set_table_precision(tx_ctx* c, int32_t precision) {
    g_err.clear();
    if (!c) return fail(TX_EINVAL, "tx_set_table_precision: NULL ctx");
    if (precision != TX_PREC_F32 && precision != TX_PREC_BF16)
        return fail(TX_EINVAL, "tx_set_table_precision: unknown precision %d", precision);
    if (precision == TX_PREC_BF16)
        if (int rc = alloc_bf16(c)) return rc;
    c->bf16 = precision == TX_PREC_BF16;
    c->dm.bf16_pred = c->bf16 ? 1 : 0;
    c->bf16_stale = true;
    return 0;
}

int tx_set_frozen_precision(tx_ctx* c, int32_t precision) {
    g_err.clear();
    if (!c) return fail(TX_EINVAL, "tx_set_frozen_precision: NULL ctx");
    if (precision != TX_PREC_F32 && precision != TX_PREC_BF16)
        return fail(TX_EINVAL, "tx_set_frozen_precision: unknown precision %d", precision);
    if (precision == TX_PREC_BF16) {
        if (c->shape.waves > 1)
            return fail(TX_EUNSUPPORTED, "tx_set_frozen_precision: BF16 frozen sides need f <= 1024");
        if (int rc = alloc_bf16(c)) return rc;
    }
    c->bf16_frozen = precision == TX_PREC_BF16;
    return 0;
}

int tx_rebuild_ycache(tx_ctx* c) {
    g_err.clear();
    if (!c) return fail(TX_EINVAL, "tx_rebuild_ycache: NULL ctx");
    c->bf16_stale = true;
    HIPC(tx::launch_rebuild(c->dm, c->n_ud, c->shape.team, c->shape.vec, c->st));
    HIPC(hipStreamSynchronize(c->st));
    c->ycache_current = true;
    return 0;
}

int tx_get_buckets(tx_ctx* c, tx_bucket_info* out) {
    g_err.clear();
    if (!c || !out) return fail(TX_EINVAL, "tx_get_buckets: NULL argument");
    if (!c->bk_ready) return fail(TX_EINVAL, "tx_get_buckets: no schedule with user_order 3 has been set");
    for (int b = 0; b < TX_N_BUCKETS; ++b) {
        const auto& k = c->bk[b];
        out[b] = tx_bucket_info{k.lo, k.hi, k.users, k.ratings, k.lanes, k.waves, k.teams, c->bk_one_launch ? 1 : 0, k.ms};
    }
    return 0;
}

int tx_set_bucket_shape(tx_ctx* c, int32_t bucket, int32_t lanes, int32_t waves, int32_t teams) {
    g_err.clear();
    if (!c) return fail(TX_EINVAL, "tx_set_bucket_shape: NULL ctx");
    if (!c->bk_ready) return fail(TX_EINVAL, "tx_set_bucket_shape: set a schedule with user_order 3 first");
    if (bucket < 0 || bucket >= TX_N_BUCKETS) return fail(TX_EINVAL, "bucket must be 0..%d", TX_N_BUCKETS - 1);
    if ((lanes != 16 && lanes != 64) || (waves != 1 && waves != 2) || (lanes == 16 && waves == 2) || teams < 1)
        return fail(TX_EINVAL, "lanes_per_user 16 or 64, waves_per_user 1 or 2 (not with 16 lanes), teams >= 1");
    if ((lanes == 16 || waves == 2) && !bucket_shapes_ok(c))
        return fail(TX_EUNSUPPORTED, "bucket shapes other than one wave: low-memory models without ordinal head or "
                                     "neighbourhood terms");
    if (lanes == 16 && c->fs > 256) return fail(TX_EUNSUPPORTED, "16-lane teams need a padded row of at most 256");
    if (waves == 2 && (c->fs < 128 || c->fs > 1024))
        return fail(TX_EUNSUPPORTED, "two-wave teams need a padded row of 128 to 1024");
    auto& k = c->bk[bucket];
    k.lanes = lanes;
    k.waves = waves;
    k.teams = teams;
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        c->graph_S = 0;
    }
    return 0;
}

int tx_last_epoch_stats(tx_ctx* c, tx_epoch_stats* out) {
    g_err.clear();
    if (!c || !out) return fail(TX_EINVAL, "tx_last_epoch_stats: NULL argument");
    *out = c->last;
    return 0;
}

int tx_predict_device(tx_ctx* c, int64_t n, const int32_t* d_u, const int32_t* d_i, const int16_t* d_d, float* d_out,
                      void* stream) {
    g_err.clear();
    if (!c || n < 0 || (n > 0 && (!d_u || !d_i || !d_d || !d_out))) return fail(TX_EINVAL, "tx_predict_device: bad arguments");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->st;
    if (int rc = ensure_bf16(c)) return rc;
    HIPC(tx::launch_predict(c->dm, 0, n, d_u, d_i, d_d, d_out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            c->shape.team, c->shape.vec, st));
    return 0;
}

namespace {

int check_queries(tx_ctx* c, int64_t n, const int32_t* u, const int32_t* i) {
    for (int64_t k = 0; k < n; ++k) {
        if (u[k] < 0 || u[k] >= c->U) return fail(TX_EINVAL, "user index %d out of range at %lld", u[k], (long long)k);
        if (i[k] < 0 || i[k] >= c->I) return fail(TX_EINVAL, "item index %d out of range at %lld", i[k], (long long)k);
    }
    return 0;
}

// Device memory for host-buffer queries: one chunk of `rows` queries on a
// stream, plus two FP64 sums for calc_rmse.
struct Staging {
    hipStream_t st;
    char* base;
    int64_t rows;
    double* acc;
};

size_t round256(size_t bytes) { return (bytes + 255) / 256 * 256; }

// Bytes of staging for `rows` queries: user, item (i32), day (i16), four
// float outputs (the first doubles as calc_rmse's target), is_test (i8).
size_t staging_bytes(int64_t rows) {
    const size_t r = static_cast<size_t>(rows);
    return round256(r * 4) * 2 + round256(r * 2) + round256(r * 4) * 4 + round256(r);
}

// Host-buffer predict / subscores / rmse in chunks through `sg`.
int run_host(tx_ctx* c, const Staging& sg, int mode, int64_t n, const int32_t* u, const int32_t* i, const int16_t* d,
             float* o0, float* o1, float* o2, float* o3, const float* target, const int8_t* is_test, double* acc) {
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(n, 1), sg.rows);
    char* base = sg.base;
    auto carve = [&](size_t bytes) {
        char* p = base;
        base += round256(bytes);
        return p;
    };
    int32_t* du = reinterpret_cast<int32_t*>(carve(chunk * 4));
    int32_t* di = reinterpret_cast<int32_t*>(carve(chunk * 4));
    int16_t* dd = reinterpret_cast<int16_t*>(carve(chunk * 2));
    float* d0 = reinterpret_cast<float*>(carve(chunk * 4));
    float* d1 = reinterpret_cast<float*>(carve(chunk * 4));
    float* d2 = reinterpret_cast<float*>(carve(chunk * 4));
    float* d3 = reinterpret_cast<float*>(carve(chunk * 4));
    int8_t* dt = reinterpret_cast<int8_t*>(carve(chunk));
    hipStream_t st = sg.st;
    if (mode == 2) HIPC(tx::launch_fill_f64(sg.acc, 0.0, 2, st));
    for (int64_t lo = 0; lo < n; lo += chunk) {
        const int64_t m = std::min(chunk, n - lo);
        HIPC(hipMemcpyAsync(du, u + lo, m * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(di, i + lo, m * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(dd, d + lo, m * 2, hipMemcpyHostToDevice, st));
        if (mode == 2) {
            HIPC(hipMemcpyAsync(d0, target + lo, m * 4, hipMemcpyHostToDevice, st));
            if (is_test) HIPC(hipMemcpyAsync(dt, is_test + lo, m, hipMemcpyHostToDevice, st));
        }
        HIPC(tx::launch_predict(c->dm, mode, m, du, di, dd, d0, d1, d2, d3, mode == 2 ? d0 : nullptr,
                                (mode == 2 && is_test) ? dt : nullptr, sg.acc, c->shape.team, c->shape.vec, st));
        if (mode == 0) HIPC(hipMemcpyAsync(o0 + lo, d0, m * 4, hipMemcpyDeviceToHost, st));
        if (mode == 1) {
            HIPC(hipMemcpyAsync(o0 + lo, d0, m * 4, hipMemcpyDeviceToHost, st));
            HIPC(hipMemcpyAsync(o1 + lo, d1, m * 4, hipMemcpyDeviceToHost, st));
            HIPC(hipMemcpyAsync(o2 + lo, d2, m * 4, hipMemcpyDeviceToHost, st));
            if (o3) HIPC(hipMemcpyAsync(o3 + lo, d3, m * 4, hipMemcpyDeviceToHost, st));
        }
        HIPC(hipStreamSynchronize(st));
    }
    if (mode == 2) {
        HIPC(hipMemcpyAsync(acc, sg.acc, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
    }
    return 0;
}

// Take a stager with room for a chunk of min(n, 4M) queries (a new one when
// every stager is in use), ordered after the work queued on the model's
// stream so far.
int checkout(tx_ctx* c, int64_t n, tx_ctx::Stager** out) {
    const int64_t rows = std::min<int64_t>(std::max<int64_t>(n, 1), int64_t{1} << 22);
    const size_t need = staging_bytes(rows);
    tx_ctx::Stager* g = nullptr;
    {
        std::lock_guard<std::mutex> lock(c->stage_mu);
        if (!c->free_stagers.empty()) {
            g = c->free_stagers.back();
            c->free_stagers.pop_back();
        } else {
            c->stagers.push_back(std::make_unique<tx_ctx::Stager>());
            g = c->stagers.back().get();
        }
    }
    auto give_back = [&](int rc) {
        std::lock_guard<std::mutex> lock(c->stage_mu);
        c->free_stagers.push_back(g);
        return rc;
    };
    if (!g->st) {
        if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess)
            return give_back(fail(TX_EDEVICE, "hipStreamCreate failed"));
        if (hipEventCreateWithFlags(&g->ready, hipEventDisableTiming) != hipSuccess)
            return give_back(fail(TX_EDEVICE, "hipEventCreate failed"));
        if (hipMalloc(&g->acc, 2 * sizeof(double)) != hipSuccess)
            return give_back(fail(TX_ENOMEM, "hipMalloc of the staging sums failed"));
    }
    if (g->bytes < need) {
        if (g->buf) (void)hipFree(g->buf);
        g->buf = nullptr;
        g->bytes = 0;
        hipError_t e = hipMalloc(&g->buf, need);
        if (e != hipSuccess) return give_back(fail(TX_ENOMEM, "hipMalloc(%zu bytes): %s", need, hipGetErrorString(e)));
        g->bytes = need;
    }
    if (hipEventRecord(g->ready, c->st) != hipSuccess || hipStreamWaitEvent(g->st, g->ready, 0) != hipSuccess)
        return give_back(fail(TX_EDEVICE, "ordering the staging stream after the model's stream failed"));
    *out = g;
    return 0;
}

// run_host through a stager of the pool.
int run_staged(tx_ctx* c, int mode, int64_t n, const int32_t* u, const int32_t* i, const int16_t* d, float* o0,
               float* o1, float* o2, float* o3, const float* target, const int8_t* is_test, double* acc) {
    if (int rc = check_queries(c, n, u, i)) return rc;
    if (int rc = ensure_bf16(c)) return rc;
    tx_ctx::Stager* g = nullptr;
    if (int rc = checkout(c, n, &g)) return rc;
    const int64_t rows = std::min<int64_t>(std::max<int64_t>(n, 1), int64_t{1} << 22);
    const Staging sg{g->st, static_cast<char*>(g->buf), rows, g->acc};
    const int rc = run_host(c, sg, mode, n, u, i, d, o0, o1, o2, o3, target, is_test, acc);
    if (rc) (void)hipStreamSynchronize(g->st);
    std::lock_guard<std::mutex> lock(c->stage_mu);
    c->free_stagers.push_back(g);
    return rc;
}

}  // namespace

int tx_predict(tx_ctx* c, int64_t n, const int32_t* u, const int32_t* i, const int16_t* d, float* out) {
    g_err.clear();
    if (!c || n < 0 || (n > 0 && (!u || !i || !d || !out))) return fail(TX_EINVAL, "tx_predict: bad arguments");
    return run_staged(c, 0, n, u, i, d, out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

int tx_predict_subscores(tx_ctx* c, int64_t n, const int32_t* u, const int32_t* i, const int16_t* d, float* bias,
                         float* mf, float* nsvd1, float* nbr) {
    g_err.clear();
    if (!c || n < 0 || (n > 0 && (!u || !i || !d || !bias || !mf || !nsvd1 || (c->nb && !nbr))))
        return fail(TX_EINVAL, "tx_predict_subscores: bad arguments (nbr is required with the neighbourhood terms)");
    return run_staged(c, 1, n, u, i, d, bias, mf, nsvd1, c->nb ? nbr : nullptr, nullptr, nullptr, nullptr);
}

int tx_n_subscores(tx_ctx* c, int32_t* out) {
    g_err.clear();
    if (!c || !out) return fail(TX_EINVAL, "tx_n_subscores: NULL argument");
    *out = c->nb ? 4 : 3;
    return 0;
}

int tx_predict_noalloc(tx_ctx* c, int64_t n, const int32_t* u, const int32_t* i, const int16_t* d, float* out) {
    g_err.clear();
    if (!c || n < 0 || (n > 0 && (!u || !i || !d || !out))) return fail(TX_EINVAL, "tx_predict_noalloc: bad arguments");
    if (n > TX_NOALLOC_MAX_ROWS)
        return fail(TX_EINVAL, "tx_predict_noalloc: %lld queries exceed the maximum of %d per call", (long long)n,
                    TX_NOALLOC_MAX_ROWS);
    if (!c->d_noalloc) return fail(TX_EDEVICE, "tx_predict_noalloc: no scratch array on device %d", c->device);
    if (int rc = check_queries(c, n, u, i)) return rc;
    if (int rc = ensure_bf16(c)) return rc;
    // the staging of one chunk of TX_NOALLOC_MAX_ROWS queries, then the sums
    char* base = static_cast<char*>(c->d_noalloc);
    static_assert(tx::kNoAllocRows == TX_NOALLOC_MAX_ROWS, "scratch sized for the header's limit");
    const Staging sg{c->st, base, TX_NOALLOC_MAX_ROWS,
                     reinterpret_cast<double*>(base + staging_bytes(TX_NOALLOC_MAX_ROWS))};
    return run_host(c, sg, 0, n, u, i, d, out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

int tx_calc_rmse(tx_ctx* c, const tx_ratings* ds, double* out) {
    g_err.clear();
    if (!c || !ds || !out) return fail(TX_EINVAL, "tx_calc_rmse: NULL argument");
    if (ds->n_ratings > 0 && (!ds->user_idxs || !ds->item_idxs || !ds->dates || !ds->residuals))
        return fail(TX_EINVAL, "tx_calc_rmse: rating set lacks indices or residuals");
    double acc[2] = {0.0, 0.0};
    if (int rc = run_staged(c, 2, ds->n_ratings, ds->user_idxs, ds->item_idxs, ds->dates, nullptr, nullptr, nullptr,
                          nullptr, ds->residuals, ds->is_test, acc))
        return rc;
    *out = std::sqrt(acc[0] / acc[1]);
    return 0;
}

int tx_param_len(tx_ctx* c, int32_t id, int64_t* n) {
    g_err.clear();
    if (!c || !n) return fail(TX_EINVAL, "tx_param_len: NULL argument");
    const ParamDesc p = param_desc(c, id);
    if (p.rows < 0) return fail(TX_EINVAL, "unknown parameter id %d", id);
    *n = p.rows * p.cols;
    return 0;
}

int tx_get_param(tx_ctx* c, int32_t id, float* out, int64_t n) {
    g_err.clear();
    if (!c || (n > 0 && !out)) return fail(TX_EINVAL, "tx_get_param: NULL argument");
    const ParamDesc p = param_desc(c, id);
    if (p.rows < 0) return fail(TX_EINVAL, "unknown parameter id %d", id);
    if (n != p.rows * p.cols) return fail(TX_EINVAL, "parameter %d holds %lld elements, caller passed %lld", id,
                                          (long long)(p.rows * p.cols), (long long)n);
    if (id == TX_P_GBIAS) {
        out[0] = c->dm.gbias;
        return 0;
    }
    if (n == 0) return 0;
    HIPC(hipStreamSynchronize(c->st));
    if (int rc = rows_to_host(out, p.p, p.pitch, p.cols, p.rows)) return rc;
    return 0;
}

int tx_set_param(tx_ctx* c, int32_t id, const float* in, int64_t n) {
    g_err.clear();
    if (!c || (n > 0 && !in)) return fail(TX_EINVAL, "tx_set_param: NULL argument");
    c->bf16_stale = true;
    const ParamDesc p = param_desc(c, id);
    if (p.rows < 0) return fail(TX_EINVAL, "unknown parameter id %d", id);
    if (n != p.rows * p.cols) return fail(TX_EINVAL, "parameter %d holds %lld elements, caller passed %lld", id,
                                          (long long)(p.rows * p.cols), (long long)n);
    if (id == TX_P_GBIAS) {
        c->dm.gbias = in[0];
        return 0;
    }
    if (n == 0) return 0;
    HIPC(hipStreamSynchronize(c->st));
    if (int rc = rows_to_device(p.p, p.pitch, in, p.cols, p.rows)) return rc;
    // tables the parallel kernel may skip while they are all zero
    auto all_zero = [&]() {
        for (int64_t k = 0; k < n; ++k)
            if (in[k] != 0.0f) return false;
        return true;
    };
    if (id == TX_P_YFEAT || id == TX_P_YBIAS || id == TX_P_YCACHE || id == TX_P_YCACHE_BIAS) c->ycache_current = false;
    if (id == TX_P_CUT) c->dm.skip_cut = all_zero() ? 1 : 0;
    if (id == TX_P_IBIAS_FREQ) c->dm.skip_ibf = (c->cfg.lr_ibf == 0.0f && all_zero()) ? 1 : 0;
    if (id == TX_P_BUT_BIN) c->dm.skip_but = (c->cfg.lr_tu == 0.0f && all_zero()) ? 1 : 0;
    if (id == TX_P_YBIAS) c->dm.skip_yb = (c->cfg.lr_yb == 0.0f && all_zero()) ? 1 : 0;
    // a captured epoch embeds the model by value: re-capture after a change
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        c->graph_S = 0;
    }
    return 0;
}

int tx_set_param_device(tx_ctx* c, int32_t id, const float* in, int64_t rows, int64_t ld, void* stream) {
    g_err.clear();
    if (!c || (rows > 0 && !in)) return fail(TX_EINVAL, "tx_set_param_device: NULL argument");
    // tables whose all-zero state lets the trainer skip work are set from host memory only
    if (id == TX_P_GBIAS || id == TX_P_CUT || id == TX_P_IBIAS_FREQ || id == TX_P_BUT_BIN || id == TX_P_YBIAS)
        return fail(TX_EINVAL, "tx_set_param_device: parameter %d is set with tx_set_param", id);
    const ParamDesc p = param_desc(c, id);
    if (p.rows < 0) return fail(TX_EINVAL, "unknown parameter id %d", id);
    if (rows != p.rows) return fail(TX_EINVAL, "parameter %d has %lld rows, caller passed %lld", id,
                                    (long long)p.rows, (long long)rows);
    if (ld < p.cols) return fail(TX_EINVAL, "tx_set_param_device: row pitch %lld < %lld columns", (long long)ld,
                                 (long long)p.cols);
    if (rows == 0) return 0;
    c->bf16_stale = true;
    if (id == TX_P_YFEAT || id == TX_P_YCACHE || id == TX_P_YCACHE_BIAS) c->ycache_current = false;
    // the caller's producer work first, then every ctx stream's work, then the copy
    HIPC(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    HIPC(hipStreamSynchronize(c->st));
    const size_t dpitch = p.pitch * sizeof(float);
    HIPC(hipMemcpy2D(p.p, dpitch, in, ld * sizeof(float), p.cols * sizeof(float), p.rows, hipMemcpyDeviceToDevice));
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        c->graph_S = 0;
    }
    return 0;
}

}  // extern "C"
