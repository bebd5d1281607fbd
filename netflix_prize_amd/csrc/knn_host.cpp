// Host side of the GPU KnnfModel, KnnsModel and KnnModel (`__knn-d` form):
// the C ABI declared in include/tsvd_knn.h (Knn3Model: knn3_host.cpp).
//
// knnf_create restates KnnfModel::new_with_factors (src/knnf.rs:60-90): the
// factor table goes to the device zero-padded to a multiple of 64 rows and 32
// columns, row norms and the item-item similarity max(0, cos) are computed
// there (knn_kernels.hip).  knns_create restates KnnsModel::new with
// SuppSource::Compute (src/knns.rs:78-100): the co-occurrence support of the
// training and held-out items per user, turned into the shrunk-lift
// similarity on the device.  knnd_create restates KnnModel::new for the
// Support similarity with use_stat1 over two given stat tables.  All keep
// each user's training ratings (item,
// residual, day) as CSR arrays; predictions (KnnfModel::predict and
// KnnsModel::predict are the same rule over their similarity) run batched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "knn_launch.hpp"
#include "tsvd_knn.h"

namespace tx {
void set_error(const std::string& msg);
void clear_error();
}  // namespace tx

struct knnf_ctx {
    knnf_config cfg{};
    int32_t n_items = 0, n_users = 0;
    int ip = 0, fp = 0;
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    knn::KnnView v{};
    std::vector<uint32_t> h_off;  // host copy of the training CSR offsets
    ~knnf_ctx() {
        for (void* p : allocs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(T) * (n > 0 ? n : 1));
        if (e == hipSuccess) {
            allocs.push_back(q);
            *p = static_cast<T*>(q);
        }
        return e;
    }
};

namespace {

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    tx::set_error(buf);
    return code;
}

#define HIPC(expr)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(TX_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Training ratings as per-user CSR on the device (user-sorted input, as the
// reference asserts) and the similarity matrix allocation (ip x ip).
int setup_train(knnf_ctx* c, const tx_ratings* train, int32_t n_items, const char* what) {
    const int64_t N = train->n_ratings;
    const int32_t U = train->n_users;
    c->n_items = n_items;
    c->n_users = U;
    c->ip = (n_items + 63) / 64 * 64;
    if (static_cast<int64_t>(c->ip) * c->ip > (int64_t{1} << 34))
        return fail(TX_EUNSUPPORTED, "%s: %d items exceed the similarity matrix budget", what, n_items);
    c->h_off.assign(static_cast<size_t>(U) + 1, 0);
    for (int64_t t = 0; t < N; ++t) {
        const int32_t u = train->user_idxs[t];
        if (u < 0 || u >= U) return fail(TX_EINVAL, "%s: user index %d out of range", what, u);
        if (t > 0 && u < train->user_idxs[t - 1]) return fail(TX_EINVAL, "%s: training rows not user-sorted", what);
        const int32_t i = train->item_idxs[t];
        if (i < 0 || i >= n_items) return fail(TX_EINVAL, "%s: item index %d out of range", what, i);
        c->h_off[u + 1]++;
    }
    for (int32_t u = 0; u < U; ++u) c->h_off[u + 1] += c->h_off[u];
    std::vector<float> day(static_cast<size_t>(N));
    for (int64_t t = 0; t < N; ++t) day[t] = static_cast<float>(train->dates[t]);
    HIPC(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    float *dS = nullptr, *dres = nullptr, *dday = nullptr;
    uint32_t* doff = nullptr;
    int32_t* ditem = nullptr;
    HIPC(c->alloc(&dS, static_cast<size_t>(c->ip) * c->ip));
    HIPC(c->alloc(&doff, c->h_off.size()));
    HIPC(c->alloc(&ditem, static_cast<size_t>(N)));
    HIPC(c->alloc(&dres, static_cast<size_t>(N)));
    HIPC(c->alloc(&dday, static_cast<size_t>(N)));
    HIPC(hipMemcpyAsync(doff, c->h_off.data(), sizeof(uint32_t) * c->h_off.size(), hipMemcpyHostToDevice, c->st));
    if (N > 0) {
        HIPC(hipMemcpyAsync(ditem, train->item_idxs, sizeof(int32_t) * N, hipMemcpyHostToDevice, c->st));
        HIPC(hipMemcpyAsync(dres, train->residuals, sizeof(float) * N, hipMemcpyHostToDevice, c->st));
        HIPC(hipMemcpyAsync(dday, day.data(), sizeof(float) * N, hipMemcpyHostToDevice, c->st));
    }
    HIPC(hipStreamSynchronize(c->st));  // day is a host temporary
    c->v = knn::KnnView{dS, c->ip, doff, ditem, dres, dday, 1, 1.0f, 0.0f};
    return 0;
}

}  // namespace

extern "C" {

int knnf_create(const knnf_config* cfg, const float* factors, int32_t n_items, int32_t n_feat,
                const tx_ratings* train, knnf_ctx** out) {
    tx::clear_error();
    if (!cfg || !factors || !train || !out) return fail(TX_EINVAL, "knnf_create: NULL argument");
    if (cfg->k < 1) return fail(TX_EINVAL, "knnf_create: k must be >= 1");
    if (n_items < 1 || n_feat < 1) return fail(TX_EINVAL, "knnf_create: empty factor table");
    if (train->n_items > n_items)
        return fail(TX_EINVAL, "knnf_create: factor table has %d rows for %d items", n_items, train->n_items);
    if (!train->residuals) return fail(TX_EINVAL, "knnf_create: training residuals missing");
    auto c = std::make_unique<knnf_ctx>();
    c->cfg = *cfg;
    c->fp = (n_feat + 31) / 32 * 32;
    if (int rc = setup_train(c.get(), train, n_items, "knnf_create")) return rc;
    // padded factors
    std::vector<float> F(static_cast<size_t>(c->ip) * c->fp, 0.0f);
    for (int32_t i = 0; i < n_items; ++i)
        for (int32_t k = 0; k < n_feat; ++k) F[static_cast<size_t>(i) * c->fp + k] = factors[static_cast<size_t>(i) * n_feat + k];
    float *dF = nullptr, *dn = nullptr;
    HIPC(c->alloc(&dF, F.size()));
    HIPC(c->alloc(&dn, static_cast<size_t>(c->ip)));
    HIPC(hipMemcpyAsync(dF, F.data(), sizeof(float) * F.size(), hipMemcpyHostToDevice, c->st));
    HIPC(knn::launch_norms(dF, c->fp, c->ip, dn, c->st));
    HIPC(knn::launch_cosine(dF, c->fp, dn, const_cast<float*>(c->v.S), c->ip, c->st));
    HIPC(hipStreamSynchronize(c->st));
    c->v.k = cfg->k;
    c->v.scaling = cfg->scaling;
    c->v.tau = cfg->tau;
    *out = c.release();
    return 0;
}

int knns_create(const knns_config* cfg, const tx_ratings* train, const tx_ratings* heldout, knnf_ctx** out) {
    tx::clear_error();
    if (!cfg || !train || !heldout || !out) return fail(TX_EINVAL, "knns_create: NULL argument");
    if (cfg->k < 1) return fail(TX_EINVAL, "knns_create: k must be >= 1");
    if (heldout->n_users != train->n_users) return fail(TX_EINVAL, "knns_create: user counts differ");
    const int32_t I = train->n_items > heldout->n_items ? train->n_items : heldout->n_items;
    if (I < 1) return fail(TX_EINVAL, "knns_create: no items");
    if (!train->residuals) return fail(TX_EINVAL, "knns_create: training residuals missing");
    auto c = std::make_unique<knnf_ctx>();
    if (int rc = setup_train(c.get(), train, I, "knns_create")) return rc;
    // per-user item lists: training items, then held-out items (build_bin_nbstats)
    const int32_t U = train->n_users;
    const int64_t NH = heldout->n_ratings;
    std::vector<uint32_t> hcnt(static_cast<size_t>(U) + 1, 0);
    for (int64_t t = 0; t < NH; ++t) {
        const int32_t u = heldout->user_idxs[t], i = heldout->item_idxs[t];
        if (u < 0 || u >= U || i < 0 || i >= I) return fail(TX_EINVAL, "knns_create: held-out row %lld out of range",
                                                            static_cast<long long>(t));
        hcnt[u + 1]++;
    }
    std::vector<uint32_t> uoff(static_cast<size_t>(U) + 1, 0);
    // a user without training ratings is skipped, held-out items too (src/nbstats.rs:79-80)
    for (int32_t u = 0; u < U; ++u)
        uoff[u + 1] = uoff[u] + (c->h_off[u + 1] > c->h_off[u] ? (c->h_off[u + 1] - c->h_off[u]) + hcnt[u + 1] : 0);
    std::vector<int32_t> items(uoff[U]);
    std::vector<uint32_t> pos(uoff.begin(), uoff.end() - 1);
    for (int32_t u = 0; u < U; ++u)
        for (uint32_t t = c->h_off[u]; t < c->h_off[u + 1]; ++t) items[pos[u]++] = train->item_idxs[t];
    for (int64_t t = 0; t < NH; ++t) {
        const int32_t u = heldout->user_idxs[t];
        if (c->h_off[u + 1] > c->h_off[u]) items[pos[u]++] = heldout->item_idxs[t];
    }
    // sorted by item: the device counts each item pair once, in the upper triangle
    {
        unsigned nt = std::thread::hardware_concurrency();
        nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
        std::vector<std::thread> pool;
        for (unsigned w = 0; w < nt; ++w)
            pool.emplace_back([&, w] {
                for (int64_t u = w; u < U; u += nt) std::sort(items.begin() + uoff[u], items.begin() + uoff[u + 1]);
            });
        for (auto& th : pool) th.join();
    }
    uint32_t* duoff = nullptr;
    int32_t* ditems = nullptr;
    float* ddiag = nullptr;
    HIPC(c->alloc(&duoff, uoff.size()));
    HIPC(c->alloc(&ditems, items.size()));
    HIPC(c->alloc(&ddiag, static_cast<size_t>(c->ip)));
    HIPC(hipMemcpyAsync(duoff, uoff.data(), sizeof(uint32_t) * uoff.size(), hipMemcpyHostToDevice, c->st));
    if (!items.empty())
        HIPC(hipMemcpyAsync(ditems, items.data(), sizeof(int32_t) * items.size(), hipMemcpyHostToDevice, c->st));
    float* S = const_cast<float*>(c->v.S);
    uint32_t* scratch = nullptr;
    HIPC(hipMalloc(&scratch, sizeof(uint32_t) * knn::support_scratch_words(static_cast<int64_t>(items.size()), c->ip)));
    std::unique_ptr<void, void (*)(void*)> g(scratch, [](void* p) { (void)hipFree(p); });
    HIPC(knn::launch_knns_sim(duoff, ditems, U, static_cast<int64_t>(items.size()), S, ddiag, c->ip, cfg->shrinkage,
                              scratch, c->st));
    HIPC(hipStreamSynchronize(c->st));
    c->v.k = cfg->k;
    c->v.scaling = cfg->scaling;
    c->v.tau = cfg->tau;
    *out = c.release();
    return 0;
}

int knnd_create(const knnd_config* cfg, const float* stat0, const float* stat1, int32_t n_items,
                const tx_ratings* train, knnf_ctx** out) {
    tx::clear_error();
    if (!cfg || !stat0 || !stat1 || !train || !out) return fail(TX_EINVAL, "knnd_create: NULL argument");
    if (cfg->k < 1) return fail(TX_EINVAL, "knnd_create: k must be >= 1");
    if (n_items < 1 || train->n_items > n_items)
        return fail(TX_EINVAL, "knnd_create: stats have %d items for %d", n_items, train->n_items);
    if (!train->residuals) return fail(TX_EINVAL, "knnd_create: training residuals missing");
    auto c = std::make_unique<knnf_ctx>();
    if (int rc = setup_train(c.get(), train, n_items, "knnd_create")) return rc;
    const size_t ip = static_cast<size_t>(c->ip);
    float *dD = nullptr, *ddiag = nullptr;
    HIPC(c->alloc(&dD, ip * ip));
    HIPC(c->alloc(&ddiag, ip));
    float* S = const_cast<float*>(c->v.S);
    HIPC(hipMemsetAsync(S, 0, sizeof(float) * ip * ip, c->st));
    HIPC(hipMemsetAsync(dD, 0, sizeof(float) * ip * ip, c->st));
    // the n_items x n_items tables into the padded ip x ip layout
    HIPC(hipMemcpy2DAsync(S, sizeof(float) * ip, stat0, sizeof(float) * n_items, sizeof(float) * n_items, n_items,
                          hipMemcpyHostToDevice, c->st));
    HIPC(hipMemcpy2DAsync(dD, sizeof(float) * ip, stat1, sizeof(float) * n_items, sizeof(float) * n_items, n_items,
                          hipMemcpyHostToDevice, c->st));
    HIPC(knn::launch_stat_sim(S, dD, ddiag, c->ip, cfg->shrinkage, c->st));
    HIPC(hipStreamSynchronize(c->st));
    c->v.k = cfg->k;
    c->v.scaling = cfg->scaling;
    c->v.tau = cfg->tau;
    c->v.D = dD;
    *out = c.release();
    return 0;
}

int knnd_predict(knnf_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days,
                 float* out) {
    return knnf_predict(ctx, n, users, items, days, out);
}

int knnd_destroy(knnf_ctx* ctx) { return knnf_destroy(ctx); }

int knns_predict(knnf_ctx* ctx, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days,
                 float* out) {
    return knnf_predict(ctx, n, users, items, days, out);
}

int knns_destroy(knnf_ctx* ctx) { return knnf_destroy(ctx); }

int knnf_predict(knnf_ctx* c, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days, float* out) {
    tx::clear_error();
    if (!c || (n > 0 && (!users || !items || !days || !out))) return fail(TX_EINVAL, "knnf_predict: NULL argument");
    for (int64_t q = 0; q < n; ++q) {
        if (users[q] < 0 || users[q] >= c->n_users) return fail(TX_EINVAL, "knnf_predict: user %d out of range", users[q]);
        if (items[q] < 0 || items[q] >= c->n_items) return fail(TX_EINVAL, "knnf_predict: item %d out of range", items[q]);
    }
    const int64_t chunk = int64_t{1} << 22;
    int32_t *du = nullptr, *di = nullptr;
    int16_t* dd = nullptr;
    float* dout = nullptr;
    const size_t cap = static_cast<size_t>(n < chunk ? (n > 0 ? n : 1) : chunk);
    HIPC(hipMalloc(&du, sizeof(int32_t) * cap));
    std::unique_ptr<void, void (*)(void*)> g1(du, [](void* p) { (void)hipFree(p); });
    HIPC(hipMalloc(&di, sizeof(int32_t) * cap));
    std::unique_ptr<void, void (*)(void*)> g2(di, [](void* p) { (void)hipFree(p); });
    HIPC(hipMalloc(&dd, sizeof(int16_t) * cap));
    std::unique_ptr<void, void (*)(void*)> g3(dd, [](void* p) { (void)hipFree(p); });
    HIPC(hipMalloc(&dout, sizeof(float) * cap));
    std::unique_ptr<void, void (*)(void*)> g4(dout, [](void* p) { (void)hipFree(p); });
    for (int64_t q0 = 0; q0 < n; q0 += chunk) {
        const int64_t m = n - q0 < chunk ? n - q0 : chunk;
        HIPC(hipMemcpyAsync(du, users + q0, sizeof(int32_t) * m, hipMemcpyHostToDevice, c->st));
        HIPC(hipMemcpyAsync(di, items + q0, sizeof(int32_t) * m, hipMemcpyHostToDevice, c->st));
        HIPC(hipMemcpyAsync(dd, days + q0, sizeof(int16_t) * m, hipMemcpyHostToDevice, c->st));
        HIPC(knn::launch_knnf_predict(c->v, m, du, di, dd, dout, c->st));
        HIPC(hipMemcpyAsync(out + q0, dout, sizeof(float) * m, hipMemcpyDeviceToHost, c->st));
        HIPC(hipStreamSynchronize(c->st));
    }
    return 0;
}

int knnf_similarity(knnf_ctx* c, int64_t n, const int32_t* i, const int32_t* j, float* out) {
    tx::clear_error();
    if (!c || (n > 0 && (!i || !j || !out))) return fail(TX_EINVAL, "knnf_similarity: NULL argument");
    for (int64_t q = 0; q < n; ++q) {
        if (i[q] < 0 || i[q] >= c->n_items || j[q] < 0 || j[q] >= c->n_items)
            return fail(TX_EINVAL, "knnf_similarity: item out of range");
        HIPC(hipMemcpyAsync(out + q, c->v.S + static_cast<size_t>(i[q]) * c->ip + j[q], sizeof(float),
                            hipMemcpyDeviceToHost, c->st));
    }
    HIPC(hipStreamSynchronize(c->st));
    return 0;
}

int knnf_destroy(knnf_ctx* c) {
    tx::clear_error();
    delete c;
    return 0;
}

}  // extern "C"
