// Host side of the GPU Knn3Model: the knn3_* C ABI declared in
// include/tsvd_knn.h.
//
// knn3_create restates Knn3Model::new (src/knn3.rs:350-362) around the device
// build (knn3_kernels.hip): the regularised baseline and r_prime in the
// reference's order of operations (compute_baseline, src/knn3.rs:123-165 and
// :295-300, a serial O(N) pass), each user's non-test ratings sorted by item
// (UserErrors, :37-77) and each item's users (the inverted index, :303-310);
// then on the device the co-rating counts, the candidates, the normal
// equations and the NNLS weights.  knn3_predict runs the neighbour sums
// batched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
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

struct knn3_ctx {
    knn3_config cfg{};
    int32_t n_items = 0, n_users = 0;
    float mu = 0.0f;
    std::vector<float> bm, bu;
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    knn::Knn3View v{};
    ~knn3_ctx() {
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
    // frees a build-only buffer early
    void release(const void* p) {
        for (auto it = allocs.begin(); it != allocs.end(); ++it)
            if (*it == p) {
                (void)hipFree(*it);
                allocs.erase(it);
                return;
            }
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

struct Rated {
    int32_t item;
    float err, rp;
};

unsigned worker_count() {
    const unsigned nt = std::thread::hardware_concurrency();
    return nt == 0 ? 1 : (nt > 16 ? 16 : nt);
}

template <typename F>
void run_workers(unsigned nt, F&& f) {
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; ++w) pool.emplace_back([&, w] { f(w); });
    for (auto& th : pool) th.join();
}

// f(u) for every user, users interleaved over the workers
template <typename F>
void parallel_users(int32_t U, F&& f) {
    const unsigned nt = worker_count();
    run_workers(nt, [&](unsigned w) {
        for (int64_t u = w; u < U; u += nt) f(static_cast<int32_t>(u));
    });
}

}  // namespace

extern "C" {

int knn3_create(const knn3_config* cfg, const tx_ratings* train, knn3_ctx** out) {
    tx::clear_error();
    if (!cfg || !train || !out) return fail(TX_EINVAL, "knn3_create: NULL argument");
    if (cfg->k_min < 0 || cfg->k_max < 0) return fail(TX_EINVAL, "knn3_create: k_min and k_max must be >= 0");
    if (cfg->k_max > 64) return fail(TX_EUNSUPPORTED, "knn3_create: k_max %d above 64", cfg->k_max);
    if (!train->residuals || !train->raw_ratings)
        return fail(TX_EINVAL, "knn3_create: training residuals and raw ratings are required");
    const int64_t N = train->n_ratings;
    const int32_t U = train->n_users, I = train->n_items;
    if (I < 1 || U < 1) return fail(TX_EINVAL, "knn3_create: no users or items");
    const int ip = (I + 63) / 64 * 64;
    if (ip > 32768) return fail(TX_EUNSUPPORTED, "knn3_create: %d items above 32,768", I);
    auto c = std::make_unique<knn3_ctx>();
    c->cfg = *cfg;
    c->n_items = I;
    c->n_users = U;
    const int8_t* test = train->is_test;
    for (int64_t t = 0; t < N; ++t) {
        const int32_t u = train->user_idxs[t], i = train->item_idxs[t];
        if (u < 0 || u >= U) return fail(TX_EINVAL, "knn3_create: user index %d out of range", u);
        if (i < 0 || i >= I) return fail(TX_EINVAL, "knn3_create: item index %d out of range", i);
        if (t > 0 && u < train->user_idxs[t - 1]) return fail(TX_EINVAL, "knn3_create: training rows not user-sorted");
    }
    // compute_baseline (src/knn3.rs:123-165): FP64 sums in rating order
    const float* res = train->residuals;
    const int8_t* raw = train->raw_ratings;
    double mu_sum = 0.0;
    int64_t mu_cnt = 0;
    for (int64_t t = 0; t < N; ++t) {
        if (test && test[t] != 0) continue;
        mu_sum += static_cast<double>(raw[t]);
        ++mu_cnt;
    }
    const float mu = static_cast<float>(mu_sum / static_cast<double>(mu_cnt));
    std::vector<double> bm_sum(I, 0.0);
    std::vector<int64_t> bm_cnt(I, 0);
    for (int64_t t = 0; t < N; ++t) {
        if (test && test[t] != 0) continue;
        const int32_t m = train->item_idxs[t];
        bm_sum[m] += static_cast<double>(static_cast<float>(raw[t]) - mu);
        ++bm_cnt[m];
    }
    c->bm.resize(I);
    for (int32_t m = 0; m < I; ++m)
        c->bm[m] = static_cast<float>(bm_sum[m] / (static_cast<double>(bm_cnt[m]) + static_cast<double>(cfg->bl_reg_m)));
    c->bu.assign(U, 0.0f);
    c->mu = mu;
    // bu per user (each sum sequential in rating order; users in parallel)
    std::vector<int64_t> row_off(static_cast<size_t>(U) + 1, 0);
    for (int64_t t = 0; t < N; ++t) row_off[train->user_idxs[t] + 1]++;
    for (int32_t u = 0; u < U; ++u) row_off[u + 1] += row_off[u];
    std::vector<uint32_t> u_off(static_cast<size_t>(U) + 1, 0);
    parallel_users(U, [&](int32_t u) {
        double sum = 0.0;
        int64_t cnt = 0;
        for (int64_t t = row_off[u]; t < row_off[u + 1]; ++t) {
            if (test && test[t] != 0) continue;
            sum += static_cast<double>((static_cast<float>(raw[t]) - mu) - c->bm[train->item_idxs[t]]);
            ++cnt;
        }
        if (cnt > 0) c->bu[u] = static_cast<float>(sum / (static_cast<double>(cnt) + static_cast<double>(cfg->bl_reg_u)));
        u_off[u + 1] = static_cast<uint32_t>(cnt);
    });
    for (int32_t u = 0; u < U; ++u) u_off[u + 1] += u_off[u];
    // per-user non-test ratings with err and r_prime (src/knn3.rs:295-300), sorted by item
    const uint32_t NR = u_off[U];
    std::vector<int32_t> u_item(NR);
    std::vector<float> u_err(NR), u_rp(NR);
    std::atomic<int32_t> dup_user{-1};
    parallel_users(U, [&](int32_t u) {
        thread_local std::vector<Rated> buf;
        buf.clear();
        for (int64_t t = row_off[u]; t < row_off[u + 1]; ++t) {
            if (test && test[t] != 0) continue;
            const int32_t m = train->item_idxs[t];
            buf.push_back(Rated{m, res[t], ((static_cast<float>(raw[t]) - mu) - c->bm[m]) - c->bu[u]});
        }
        std::sort(buf.begin(), buf.end(), [](const Rated& x, const Rated& y) { return x.item < y.item; });
        uint32_t p = u_off[u];
        for (size_t k = 0; k < buf.size(); ++k, ++p) {
            if (k > 0 && buf[k].item == buf[k - 1].item) dup_user.store(u);  // any one of them is reported
            u_item[p] = buf[k].item;
            u_err[p] = buf[k].err;
            u_rp[p] = buf[k].rp;
        }
    });
    if (dup_user.load() >= 0)
        return fail(TX_EINVAL, "knn3_create: user %d rates an item twice (one rating per user and item expected)",
                    dup_user.load());
    // each item's users in ascending order (the inverted index, src/knn3.rs:303-310): a
    // stable counting sort over contiguous user ranges, one per thread
    std::vector<uint32_t> i_off(static_cast<size_t>(I) + 1, 0);
    std::vector<int32_t> i_user(NR);
    std::vector<float> i_err(NR), i_rp(NR);
    {
        const unsigned T = worker_count();
        std::vector<std::vector<uint32_t>> tpos(T, std::vector<uint32_t>(static_cast<size_t>(I), 0));
        auto range = [&](unsigned k) { return static_cast<int32_t>(static_cast<int64_t>(U) * k / T); };
        run_workers(T, [&](unsigned k) {
            for (uint32_t p = u_off[range(k)]; p < u_off[range(k + 1)]; ++p) tpos[k][u_item[p]]++;
        });
        uint32_t run = 0;
        for (int32_t m = 0; m < I; ++m) {
            i_off[m] = run;
            for (unsigned k = 0; k < T; ++k) {
                const uint32_t n = tpos[k][m];
                tpos[k][m] = run;
                run += n;
            }
        }
        i_off[I] = run;
        run_workers(T, [&](unsigned k) {
            std::vector<uint32_t>& pos = tpos[k];
            for (int32_t u = range(k); u < range(k + 1); ++u)
                for (uint32_t p = u_off[u]; p < u_off[u + 1]; ++p) {
                    const uint32_t q = pos[u_item[p]]++;
                    i_user[q] = u;
                    i_err[q] = u_err[p];
                    i_rp[q] = u_rp[p];
                }
        });
    }

    HIPC(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    knn::Knn3View& v = c->v;
    v.n_items = I;
    v.ip = ip;
    v.n_users = U;
    float *cnt = nullptr, *diag = nullptr, *d_uerr = nullptr, *d_urp = nullptr, *d_ierr = nullptr, *d_irp = nullptr;
    uint32_t *d_uoff = nullptr, *d_ioff = nullptr;
    int32_t *d_uitem = nullptr, *d_iuser = nullptr;
    HIPC(c->alloc(&cnt, static_cast<size_t>(ip) * ip));
    HIPC(c->alloc(&diag, static_cast<size_t>(ip)));
    HIPC(c->alloc(&d_uoff, u_off.size()));
    HIPC(c->alloc(&d_uitem, NR));
    HIPC(c->alloc(&d_uerr, NR));
    HIPC(c->alloc(&d_urp, NR));
    HIPC(c->alloc(&d_ioff, i_off.size()));
    HIPC(c->alloc(&d_iuser, NR));
    HIPC(c->alloc(&d_ierr, NR));
    HIPC(c->alloc(&d_irp, NR));
    HIPC(c->alloc(&v.nb_cnt, static_cast<size_t>(I)));
    HIPC(c->alloc(&v.nb_item, static_cast<size_t>(I) * 64));
    HIPC(c->alloc(&v.sys, static_cast<size_t>(I) * 64 * 64));
    HIPC(c->alloc(&v.rhs, static_cast<size_t>(I) * 64));
    HIPC(c->alloc(&v.w_cnt, static_cast<size_t>(I)));
    HIPC(c->alloc(&v.w_item, static_cast<size_t>(I) * 64));
    HIPC(c->alloc(&v.w_val, static_cast<size_t>(I) * 64));
    auto up = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->st) : hipSuccess;
    };
    HIPC(up(d_uoff, u_off.data(), sizeof(uint32_t) * u_off.size()));
    HIPC(up(d_uitem, u_item.data(), sizeof(int32_t) * NR));
    HIPC(up(d_uerr, u_err.data(), sizeof(float) * NR));
    HIPC(up(d_urp, u_rp.data(), sizeof(float) * NR));
    HIPC(up(d_ioff, i_off.data(), sizeof(uint32_t) * i_off.size()));
    HIPC(up(d_iuser, i_user.data(), sizeof(int32_t) * NR));
    HIPC(up(d_ierr, i_err.data(), sizeof(float) * NR));
    HIPC(up(d_irp, i_rp.data(), sizeof(float) * NR));
    std::vector<int32_t> heavy(I);
    for (int32_t m = 0; m < I; ++m) heavy[m] = m;
    std::stable_sort(heavy.begin(), heavy.end(),
                     [&](int32_t a, int32_t b) { return i_off[a + 1] - i_off[a] > i_off[b + 1] - i_off[b]; });
    int32_t* d_heavy = nullptr;
    HIPC(c->alloc(&d_heavy, static_cast<size_t>(I)));
    HIPC(up(d_heavy, heavy.data(), sizeof(int32_t) * I));
    HIPC(hipMemsetAsync(v.w_cnt, 0, sizeof(int32_t) * I, c->st));
    // co-rating counts of the non-test ratings (the diagonal: each item's users)
    uint32_t* scratch = nullptr;
    HIPC(c->alloc(&scratch, knn::support_scratch_words(static_cast<int64_t>(NR), ip)));
    HIPC(knn::launch_support(d_uoff, d_uitem, U, static_cast<int64_t>(NR), cnt, diag, ip, scratch, c->st));
    v.cnt = cnt;
    v.diag = diag;
    v.u_off = d_uoff;
    v.u_item = d_uitem;
    v.u_err = d_uerr;
    v.u_rp = d_urp;
    v.i_off = d_ioff;
    v.i_user = d_iuser;
    v.i_err = d_ierr;
    v.i_rp = d_irp;
    v.heavy_first = d_heavy;
    v.threshold = cfg->threshold;
    v.shrinkage = cfg->shrinkage;
    v.k_min = cfg->k_min;
    v.k_max = cfg->k_max;
    v.x = static_cast<double>(cfg->x);
    v.reg = static_cast<double>(cfg->reg);
    HIPC(knn::launch_knn3_build(v, c->st));
    HIPC(hipStreamSynchronize(c->st));  // host arrays above are temporaries
    // build-only device buffers
    for (const void* p : {static_cast<const void*>(cnt), static_cast<const void*>(scratch), static_cast<const void*>(v.sys),
                          static_cast<const void*>(v.rhs), static_cast<const void*>(d_ioff),
                          static_cast<const void*>(d_iuser), static_cast<const void*>(d_ierr),
                          static_cast<const void*>(d_irp), static_cast<const void*>(d_heavy)})
        c->release(p);
    v.cnt = nullptr;
    v.sys = nullptr;
    v.rhs = nullptr;
    v.i_off = nullptr;
    v.i_user = nullptr;
    v.i_err = nullptr;
    v.i_rp = nullptr;
    v.heavy_first = nullptr;
    *out = c.release();
    return 0;
}

int knn3_predict(knn3_ctx* c, int64_t n, const int32_t* users, const int32_t* items, const int16_t* days, float* out) {
    (void)days;  // Knn3Model::predict ignores the day (src/knn3.rs:379)
    tx::clear_error();
    if (!c || (n > 0 && (!users || !items || !out))) return fail(TX_EINVAL, "knn3_predict: NULL argument");
    for (int64_t q = 0; q < n; ++q) {
        if (users[q] < 0 || users[q] >= c->n_users) return fail(TX_EINVAL, "knn3_predict: user %d out of range", users[q]);
        if (items[q] < 0 || items[q] >= c->n_items) return fail(TX_EINVAL, "knn3_predict: item %d out of range", items[q]);
    }
    if (n == 0) return 0;
    const int64_t chunk = int64_t{1} << 22;
    const size_t cap = static_cast<size_t>(n < chunk ? n : chunk);
    int32_t *du = nullptr, *di = nullptr;
    float* dout = nullptr;
    HIPC(hipMalloc(&du, sizeof(int32_t) * cap));
    std::unique_ptr<void, void (*)(void*)> g1(du, [](void* p) { (void)hipFree(p); });
    HIPC(hipMalloc(&di, sizeof(int32_t) * cap));
    std::unique_ptr<void, void (*)(void*)> g2(di, [](void* p) { (void)hipFree(p); });
    HIPC(hipMalloc(&dout, sizeof(float) * cap));
    std::unique_ptr<void, void (*)(void*)> g3(dout, [](void* p) { (void)hipFree(p); });
    for (int64_t q0 = 0; q0 < n; q0 += chunk) {
        const int64_t m = n - q0 < chunk ? n - q0 : chunk;
        HIPC(hipMemcpyAsync(du, users + q0, sizeof(int32_t) * m, hipMemcpyHostToDevice, c->st));
        HIPC(hipMemcpyAsync(di, items + q0, sizeof(int32_t) * m, hipMemcpyHostToDevice, c->st));
        HIPC(knn::launch_knn3_predict(c->v, m, du, di, dout, c->st));
        HIPC(hipMemcpyAsync(out + q0, dout, sizeof(float) * m, hipMemcpyDeviceToHost, c->st));
        HIPC(hipStreamSynchronize(c->st));
    }
    return 0;
}

int knn3_weights(knn3_ctx* c, int32_t* counts, int32_t* items, float* weights) {
    tx::clear_error();
    if (!c || !counts || !items || !weights) return fail(TX_EINVAL, "knn3_weights: NULL argument");
    const size_t I = static_cast<size_t>(c->n_items);
    HIPC(hipMemcpyAsync(counts, c->v.w_cnt, sizeof(int32_t) * I, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(items, c->v.w_item, sizeof(int32_t) * I * 64, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(weights, c->v.w_val, sizeof(float) * I * 64, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    return 0;
}

int knn3_candidates(knn3_ctx* c, int32_t* counts, int32_t* items) {
    tx::clear_error();
    if (!c || !counts || !items) return fail(TX_EINVAL, "knn3_candidates: NULL argument");
    const size_t I = static_cast<size_t>(c->n_items);
    HIPC(hipMemcpyAsync(counts, c->v.nb_cnt, sizeof(int32_t) * I, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(items, c->v.nb_item, sizeof(int32_t) * I * 64, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    return 0;
}

int knn3_baseline(knn3_ctx* c, float* mu, float* bm, float* bu) {
    tx::clear_error();
    if (!c || !mu || !bm || !bu) return fail(TX_EINVAL, "knn3_baseline: NULL argument");
    *mu = c->mu;
    std::copy(c->bm.begin(), c->bm.end(), bm);
    std::copy(c->bu.begin(), c->bu.end(), bu);
    return 0;
}

int knn3_destroy(knn3_ctx* c) {
    tx::clear_error();
    delete c;
    return 0;
}

}  // extern "C"
