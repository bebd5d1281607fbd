// Reference-compatible random streams for TxModel::new and get_users.
//
// The reference draws its initial factors from rand 0.9.2's StdRng
// (ChaCha12, rand_chacha 0.9.0) seeded with SeedableRng::seed_from_u64
// (rand_core 0.9.3: PCG32 fills the 32-byte key), samples Normal<f32>
// (rand_distr 0.5.1: f64 ZIGNOR ziggurat, cast to f32) and shuffles users
// with SliceRandom::shuffle (inside-out Fisher-Yates with several indices
// drawn per u32).  Reproducing these streams lets a model created here start
// from exactly the reference's parameters (src/tx.rs:461, 600-648) and visit
// users in exactly its order (src/lib.rs:486-494).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace tx {

class ChaChaRng {
public:
    explicit ChaChaRng(uint64_t seed, int rounds = 12) : rounds_(rounds) {
        // seed_from_u64: PCG32 (MUL/INC of rand_core) output words form the key.
        uint64_t s = seed;
        for (int w = 0; w < 8; ++w) {
            s = s * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = static_cast<uint32_t>(((s >> 18) ^ s) >> 27);
            const uint32_t rot = static_cast<uint32_t>(s >> 59);
            key_[w] = (xs >> rot) | (xs << ((32u - rot) & 31u));
        }
    }

    ChaChaRng(const ChaChaRng&) = delete;
    ChaChaRng& operator=(const ChaChaRng&) = delete;

    uint32_t next_u32() {
        if (pos_ == len_) refill();
        return cur_[pos_++];
    }

    // Bulk mode: the keystream is produced kBulkBlocks blocks at a time by
    // several threads (a block depends only on its counter) and served in
    // stream order, so the words are those of the four-block buffer, only
    // computed ahead.  Words already produced are served first either way.
    void set_bulk(bool on) { bulk_ = on; }

    // BlockRng::next_u64: two consecutive words, low first.
    uint64_t next_u64() {
        if (pos_ + 2 <= len_) {
            const uint64_t v = cur_[pos_] | (static_cast<uint64_t>(cur_[pos_ + 1]) << 32);
            pos_ += 2;
            return v;
        }
        const uint64_t lo = next_u32();
        const uint64_t hi = next_u32();
        return lo | (hi << 32);
    }

private:
    static constexpr int kWords = 64;  // rand_chacha buffers 4 blocks

    static uint32_t rotl(uint32_t v, int n) { return (v << n) | (v >> (32 - n)); }

    static constexpr int64_t kBulkBlocks = int64_t{1} << 18;  // 16 MB of keystream

    void refill() {
        if (bulk_) {
            big_.resize(static_cast<size_t>(kBulkBlocks) * 16);
            const uint64_t c0 = ctr_;
            unsigned nt = std::thread::hardware_concurrency();
            nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
            std::vector<std::thread> th;
            for (unsigned t = 0; t < nt; ++t)
                th.emplace_back([this, t, nt, c0]() {  // a contiguous range of blocks per thread
                    const int64_t lo = kBulkBlocks * t / nt, hi = kBulkBlocks * (t + 1) / nt;
                    for (int64_t b = lo; b < hi; ++b) block(c0 + static_cast<uint64_t>(b), big_.data() + 16 * b);
                });
            for (auto& x : th) x.join();
            ctr_ += static_cast<uint64_t>(kBulkBlocks);
            cur_ = big_.data();
            len_ = static_cast<int64_t>(big_.size());
        } else {
            for (int b = 0; b < 4; ++b) block(ctr_ + static_cast<uint64_t>(b), buf_ + 16 * b);
            ctr_ += 4;
            cur_ = buf_;
            len_ = kWords;
        }
        pos_ = 0;
    }

    void block(uint64_t ctr, uint32_t* out) const {
        uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                           key_[0], key_[1], key_[2], key_[3], key_[4], key_[5], key_[6], key_[7],
                           static_cast<uint32_t>(ctr), static_cast<uint32_t>(ctr >> 32), 0u, 0u};
        uint32_t x[16];
        std::memcpy(x, st, sizeof x);
        auto qr = [&](int a, int b, int c, int d) {
            x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
            x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
            x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
            x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
        };
        for (int r = 0; r < rounds_; r += 2) {
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) out[i] = x[i] + st[i];
    }

    uint32_t key_[8];
    uint32_t buf_[kWords];
    std::vector<uint32_t> big_;
    uint64_t ctr_ = 0;
    const uint32_t* cur_ = buf_;  // buf_ or big_ (the object is not copyable)
    int64_t pos_ = 0, len_ = 0;
    bool bulk_ = false;
    int rounds_;
};

// ZIGNOR tables for N(0,1), 256 layers (R = 3.6541528853610088,
// V = 0.00492867323399), rounded to 18 decimals as rand_distr's literals are.
class NormalTables {
public:
    static const NormalTables& get() {
        static const NormalTables t;
        return t;
    }
    double x[257];
    double f[257];

private:
    static double lit(double v) {
        char s[64];
        std::snprintf(s, sizeof s, "%.18f", v);
        return std::strtod(s, nullptr);
    }
    NormalTables() {
        const double R = 3.6541528853610088, V = 0.00492867323399;
        auto pdf = [](double v) { return std::exp(-v * v / 2.0); };
        double raw[257];
        raw[0] = V / pdf(R);
        raw[1] = R;
        for (int i = 2; i < 256; ++i) raw[i] = std::sqrt(-2.0 * std::log(V / raw[i - 1] + pdf(raw[i - 1])));
        raw[256] = 0.0;
        for (int i = 0; i < 257; ++i) {
            f[i] = lit(pdf(raw[i]));
            x[i] = lit(raw[i]);
        }
    }
};

inline double bits_to_f64(uint64_t b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// StandardNormal for f64 (ziggurat with exact-tail fallback).
inline double standard_normal(ChaChaRng& rng) {
    const NormalTables& T = NormalTables::get();
    constexpr double kR = 3.654152885361008796;
    for (;;) {
        const uint64_t bits = rng.next_u64();
        const int i = static_cast<int>(bits & 0xff);
        const double u = bits_to_f64((1024ull << 52) | (bits >> 12)) - 3.0;   // [-1, 1)
        const double v = u * T.x[i];
        if (std::fabs(v) < T.x[i + 1]) return v;
        if (i == 0) {
            double tx = 1.0, ty = 0.0;
            while (-2.0 * ty < tx * tx) {
                const double a = bits_to_f64((1023ull << 52) | (rng.next_u64() >> 12)) - (1.0 - 0x1p-53);
                const double b = bits_to_f64((1023ull << 52) | (rng.next_u64() >> 12)) - (1.0 - 0x1p-53);
                tx = std::log(a) / kR;
                ty = std::log(b);
            }
            return u < 0.0 ? tx - kR : kR - tx;
        }
        const double unif = static_cast<double>(rng.next_u64() >> 11) * 0x1p-53;
        if (T.f[i + 1] + (T.f[i] - T.f[i + 1]) * unif < std::exp(-v * v / 2.0)) return v;
    }
}

// rand_array2: Normal<f32>(0, sigma) = 0 + sigma * (f32)z, one draw per element.
inline void fill_normal(ChaChaRng& rng, float* out, int64_t n, float sigma) {
    rng.set_bulk(n >= (int64_t{1} << 20));
    for (int64_t k = 0; k < n; ++k) {
        const float z = static_cast<float>(standard_normal(rng));
        out[k] = 0.0f + sigma * z;
    }
    rng.set_bulk(false);
}

// UniformInt<u32> single sample in [0, range) (rand 0.9's widening-multiply
// method with one extra draw when the low half is biased).
inline uint32_t uniform_below(ChaChaRng& rng, uint32_t range) {
    const uint64_t m = static_cast<uint64_t>(rng.next_u32()) * range;
    uint32_t hi = static_cast<uint32_t>(m >> 32);
    const uint32_t lo = static_cast<uint32_t>(m);
    if (lo > 0u - range) {
        const uint32_t hi2 = static_cast<uint32_t>((static_cast<uint64_t>(rng.next_u32()) * range) >> 32);
        if (static_cast<uint64_t>(lo) + hi2 > 0xffffffffull) hi += 1;
    }
    return hi;
}

// SliceRandom::shuffle: positions 0..n-1 each swap with a uniform index in
// [0, i]; consecutive indices share one draw while their product of ranges
// fits in u32.
template <typename T>
void shuffle(ChaChaRng& rng, T* v, int64_t n) {
    if (n <= 1) return;
    uint32_t seen = 0, packed = 0;
    int left = 1;  // the first index is always 0 and costs no draw
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t range = seen + 1;
        int after;
        if (left > 0) {
            after = left - 1;
        } else {
            uint64_t prod = range;
            uint32_t next = range + 1;
            while (prod * next <= 0xffffffffull) { prod *= next; ++next; }
            packed = uniform_below(rng, static_cast<uint32_t>(prod));
            after = static_cast<int>(next - range) - 1;
        }
        uint32_t j;
        if (after == 0) {
            j = packed;
        } else {
            j = packed % range;
            packed /= range;
        }
        left = after;
        seen = range;
        T tmp = v[i];
        v[i] = v[j];
        v[j] = tmp;
    }
}

// The same shuffle in two passes, for many shuffles drawn from one stream:
// shuffle_draws consumes the stream exactly as shuffle does and records the
// packed value of each group of indices (a multiply and a compare per group,
// no division); shuffle_apply turns recorded values into the swaps (the
// divisions), in any order or thread once drawn.  Both walk the groups as
// shuffle does: index 0 is free, then each group packs consecutive ranges
// while their product fits in u32.
// The groups do not depend on n (an n-element shuffle uses the groups that
// start below n), so the chain is computed once for ranges up to kChain.
struct ShuffleChain {
    static constexpr uint32_t kChain = 1u << 17;
    std::vector<uint32_t> start, prod;  // group g covers ranges [start[g], start[g+1])
    ShuffleChain() {
        uint32_t range = 2;
        while (range < kChain) {
            uint64_t p = range;
            uint32_t next = range + 1;
            while (p * next <= 0xffffffffull) { p *= next; ++next; }
            start.push_back(range);
            prod.push_back(static_cast<uint32_t>(p));
            range = next;
        }
        start.push_back(range);
    }
    static const ShuffleChain& get() {
        static const ShuffleChain c;
        return c;
    }
};

template <typename Group>
inline void shuffle_groups(int64_t n, Group group) {
    if (n <= 1) return;
    uint32_t range = 2;  // ranges 1 (index 0) done; groups start at range 2
    int64_t i = 1;
    const ShuffleChain& ch = ShuffleChain::get();
    for (size_t g = 0; g + 1 < ch.start.size() && i < n; ++g) {
        const int64_t cnt = static_cast<int64_t>(ch.start[g + 1] - ch.start[g]);
        group(i, ch.start[g], cnt, ch.prod[g]);
        i += cnt;
        range = ch.start[g + 1];
    }
    while (i < n) {
        uint64_t prod = range;
        uint32_t next = range + 1;
        while (prod * next <= 0xffffffffull) { prod *= next; ++next; }
        const int64_t cnt = static_cast<int64_t>(next - range);  // indices in the group
        group(i, range, cnt, static_cast<uint32_t>(prod));
        i += cnt;
        range = next;
    }
}

// Appends one packed value per group of an n-element shuffle to `out`.
inline void shuffle_draws(ChaChaRng& rng, int64_t n, std::vector<uint32_t>& out) {
    shuffle_groups(n, [&](int64_t, uint32_t, int64_t, uint32_t prod) { out.push_back(uniform_below(rng, prod)); });
}

// Applies the packed values `pk` (as shuffle_draws recorded them) to v[0..n).
template <typename T>
inline void shuffle_apply(const uint32_t* pk, T* v, int64_t n) {
    shuffle_groups(n, [&](int64_t i0, uint32_t range, int64_t cnt, uint32_t) {
        uint32_t packed = *pk++;
        for (int64_t q = 0; q < cnt && i0 + q < n; ++q) {
            const uint32_t r = range + static_cast<uint32_t>(q);
            uint32_t j;
            if (q == cnt - 1) {
                j = packed;
            } else {
                j = packed % r;
                packed /= r;
            }
            T tmp = v[i0 + q];
            v[i0 + q] = v[j];
            v[j] = tmp;
        }
    });
}

// get_users: identity, or shuffled with seed ^ epoch * golden-ratio constant.
inline void user_order(int64_t n, bool shuffle_users, uint64_t seed, int64_t epoch, int64_t* out) {
    for (int64_t k = 0; k < n; ++k) out[k] = k;
    if (shuffle_users) {
        ChaChaRng rng(seed ^ (static_cast<uint64_t>(epoch) * 0x9E3779B97F4A7C15ull));
        shuffle(rng, out, n);
    }
}

}  // namespace tx
