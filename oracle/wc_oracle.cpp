// TEST INFRASTRUCTURE ONLY -- extern "C" surface of the oracle (see
// wc_oracle.hpp). Loaded by tests/ through ctypes; never by the product.
#include "wc_oracle.hpp"

#include <mutex>
#include <map>
#include <cstring>

namespace wco {

// ------------------------------------------------------------------ primes --
static u64 mulmod(u64 a, u64 b, u64 m) { return u64((unsigned __int128)a * b % m); }
static u64 powmod(u64 a, u64 e, u64 m) {
    u64 r = 1; a %= m;
    while (e) { if (e & 1) r = mulmod(r, a, m); a = mulmod(a, a, m); e >>= 1; }
    return r;
}
// Deterministic Miller-Rabin; these bases are exact for every n < 3.3e24.
bool is_prime_u64(u64 n) {
    if (n < 2) return false;
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41};
    for (u64 p : small) { if (n % p == 0) return n == p; }
    u64 d = n - 1; int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (u64 a : small) {
        u64 x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; ++r) { x = mulmod(x, x, n); if (x == n - 1) { comp = false; break; } }
        if (comp) return false;
    }
    return true;
}
static u64 next_prime(u64 x) { while (!is_prime_u64(x)) ++x; return x; }

// utils/generate_primes_hpp.py:20-43 with min=step=131072, max=17177772037
static const std::vector<u64>& prime_table() {
    static std::vector<u64> t;
    static std::once_flag once;
    std::call_once(once, [] {
        const u64 step = 131072, lo = 131072, hi = 17177772037ull;
        u64 p = next_prime(lo);
        while (p <= hi) { t.push_back(p); p = next_prime(p + step); }
    });
    return t;
}
u64 num_primes() { return prime_table().size(); }
u64 prime_at(u64 i) { return prime_table().at(i); }

// base.cuh:27-34
u64 get_valid_capacity(u64 min_capacity, u64 cg) {
    const u64 x = (min_capacity + cg - 1) / cg;
    const auto& t = prime_table();
    auto it = std::lower_bound(t.begin(), t.end(), x);
    return it == t.end() ? 0 : (*it) * cg;
}

}  // namespace wco

using namespace wco;

// ------------------------------------------------------------ dispatching --
namespace {
enum Kind { K32V32 = 0, K32V64 = 1, K64V32 = 2, K64V64 = 3 };
inline int kind(int kb, int vb) { return (kb == 64 ? 2 : 0) + (vb == 64 ? 1 : 0); }

struct Obj { int kind; void* p; };

template<template<class, class> class T, class F>
auto with(Obj* o, F f) {
    switch (o->kind) {
    case K32V32: return f(*static_cast<T<u32, u32>*>(o->p));
    case K32V64: return f(*static_cast<T<u32, u64>*>(o->p));
    case K64V32: return f(*static_cast<T<u64, u32>*>(o->p));
    default:     return f(*static_cast<T<u64, u64>*>(o->p));
    }
}
template<class T> inline T ld(const void* p, u64 i) { T v; std::memcpy(&v, (const char*)p + i * sizeof(T), sizeof(T)); return v; }
template<class T> inline void st(void* p, u64 i, T v) { std::memcpy((char*)p + i * sizeof(T), &v, sizeof(T)); }
}  // namespace

template<class F> auto with_counting(void* h, F f) {
    Obj* o = (Obj*)h;
    if (o->kind == K32V64) return f(*static_cast<Counting<u32, u64>*>(o->p));
    return f(*static_cast<Counting<u64, u64>*>(o->p));
}
template<class F> auto with_set(void* h, F f) {
    Obj* o = (Obj*)h;
    if (o->kind == 0) return f(*static_cast<HashSet<u32>*>(o->p));
    return f(*static_cast<HashSet<u64>*>(o->p));
}
extern "C" {

// ---------------------------------------------------------------- hashers --
u32 wco_fmix32(u32 x) { return fmix32(x); }
u64 wco_fmix64(u64 x) { return fmix64(x); }
u32 wco_nvidia_hash(u32 x) { return nvidia_hash(x); }
u32 wco_mueller_hash(u32 x) { return mueller_hash(x); }

// ----------------------------------------------------------------- status --
u32 wco_status_all() { return status_all(); }
u32 wco_status_error_mask() { return status_error_mask(); }
u32 wco_status_warning_mask() { return status_warning_mask(); }
u32 wco_status_not(u32 s) { return status_not(s); }
u32 wco_status_remove(u32 a, u32 b) { return status_remove(a, b); }
int wco_status_return_boolean(u32 s, u32 ignore) { return status_return_boolean(s, ignore) ? 1 : 0; }
void wco_fmix32_array(const u32* in, u32* out, u64 n) { for (u64 i = 0; i < n; ++i) out[i] = fmix32(in[i]); }
void wco_fmix64_array(const u64* in, u64* out, u64 n) { for (u64 i = 0; i < n; ++i) out[i] = fmix64(in[i]); }

// ----------------------------------------------------------------- primes --
u64 wco_num_primes() { return num_primes(); }
void wco_primes(u64* out, u64 n) { for (u64 i = 0; i < n && i < num_primes(); ++i) out[i] = prime_at(i); }
int wco_is_prime(u64 n) { return is_prime_u64(n); }
u64 wco_get_valid_capacity(u64 min_cap, u64 cg) { return get_valid_capacity(min_cap, cg); }

// ---------------------------------------------------------------- probing --
// Writes lane r's i-th slot to out[i*cg + r] (the layout of
// tests/src/probing_schemes_test.cu:44-53); returns the number of windows.
u64 wco_probe_sequence(int scheme, int key_bits, u64 key, u64 seed, u64 cap, u64 cg, u64 pl,
                       u64* out, u64 max_windows) {
    u64 w = 0;
    auto run = [&](auto k, auto s) {
        using K = decltype(k);
        Windows<K> it(scheme, cap, cg, pl);
        for (bool ok = it.begin(k, s); ok && w < max_windows; ok = it.next(), ++w)
            for (u64 r = 0; r < cg; ++r) out[w * cg + r] = it.pos[r];
    };
    if (key_bits == 32) run(u32(key), u32(seed)); else run(u64(key), u64(seed));
    return w;
}

// -------------------------------------------------------------------- SVHT --
void* wco_svht_create(int kb, int vb, u64 min_cap, u64 cg, int scheme, u64 seed) {
    Obj* o = new Obj{kind(kb, vb), nullptr};
    switch (o->kind) {
    case K32V32: o->p = new SVHT<u32, u32>(min_cap, cg, scheme, u32(seed)); break;
    case K32V64: o->p = new SVHT<u32, u64>(min_cap, cg, scheme, u32(seed)); break;
    case K64V32: o->p = new SVHT<u64, u32>(min_cap, cg, scheme, seed); break;
    default:     o->p = new SVHT<u64, u64>(min_cap, cg, scheme, seed); break;
    }
    return o;
}
// A table of exactly `capacity` slots (a multiple of cg): the geometry of
// a TableBatch table, which has no reference counterpart.
void* wco_svht_create_sized(int kb, int vb, u64 capacity, u64 cg, int scheme, u64 seed) {
    Obj* o = (Obj*)wco_svht_create(kb, vb, cg, cg, scheme, seed);
    with<SVHT>(o, [&](auto& t) {
        t.cap = capacity;
        t.init_ok = capacity != 0 && capacity % cg == 0;
        t.status = t.init_ok ? S_NONE : (S_INVALID_CONFIG | S_NOT_INIT);
        t.keys.assign(capacity, t.EMPTY);
        t.vals.assign(capacity, 0);
        return 0;
    });
    return o;
}
// TableBatch's small-table rule: the smallest prime >= max(2, SDIV(min, window)) windows
u64 wco_batch_windows(u64 min_cap, u64 window) {
    const u64 w = (min_cap + window - 1) / window;
    return next_prime(w < 2 ? 2 : w);
}
void wco_svht_destroy(void* h) {
    Obj* o = (Obj*)h;
    with<SVHT>(o, [](auto& t) { delete &t; return 0; });
    delete o;
}
u64 wco_svht_capacity(void* h) { return with<SVHT>((Obj*)h, [](auto& t) { return t.cap; }); }
u64 wco_svht_size(void* h) { return with<SVHT>((Obj*)h, [](auto& t) { return t.size(); }); }
u32 wco_svht_status(void* h) { return with<SVHT>((Obj*)h, [](auto& t) { return t.status; }); }
void wco_svht_clear_status(void* h) { with<SVHT>((Obj*)h, [](auto& t) { t.status = 0; return 0; }); }
void wco_svht_insert(void* h, const void* keys, const void* vals, u64 n, u64 pl, u32* status_out) {
    with<SVHT>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.keys[0])>; using V = std::decay_t<decltype(t.vals[0])>;
        for (u64 i = 0; i < n; ++i) { u32 s = t.insert(ld<K>(keys, i), ld<V>(vals, i), pl); if (status_out) status_out[i] = s; }
        return 0;
    });
}
// writes vals_out[i] only on success (gpu_engine.cuh:252-258)
void wco_svht_retrieve(void* h, const void* keys, u64 n, void* vals_out, u64 pl, u32* status_out) {
    with<SVHT>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.keys[0])>; using V = std::decay_t<decltype(t.vals[0])>;
        for (u64 i = 0; i < n; ++i) {
            V v{}; u32 s = t.retrieve(ld<K>(keys, i), v, pl);
            if (s == S_NONE) st<V>(vals_out, i, v);
            if (status_out) status_out[i] = s;
        }
        return 0;
    });
}
void wco_svht_erase(void* h, const void* keys, u64 n, u64 pl, u32* status_out) {
    with<SVHT>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.keys[0])>;
        for (u64 i = 0; i < n; ++i) { u32 s = t.erase(ld<K>(keys, i), pl); if (status_out) status_out[i] = s; }
        return 0;
    });
}
// raw slot dump (keys, values) for invariant checks
void wco_svht_slots(void* h, void* keys_out, void* vals_out) {
    with<SVHT>((Obj*)h, [&](auto& t) {
        std::memcpy(keys_out, t.keys.data(), t.keys.size() * sizeof(t.keys[0]));
        std::memcpy(vals_out, t.vals.data(), t.vals.size() * sizeof(t.vals[0]));
        return 0;
    });
}

// ---------------------------------------------------------------- counting --
// keys u32/u64, counters u64 (the default Value = index_t)
void* wco_counting_create(int kb, u64 min_cap, u64 cg, int scheme, u64 seed, u64 max_count) {
    Obj* o = new Obj{kind(kb, 64), nullptr};
    if (kb == 32) o->p = new Counting<u32, u64>(min_cap, cg, scheme, u32(seed), max_count);
    else o->p = new Counting<u64, u64>(min_cap, cg, scheme, seed, max_count);
    return o;
}
void wco_counting_destroy(void* h) { with_counting(h, [](auto& t) { delete &t; return 0; }); delete (Obj*)h; }
u64 wco_counting_capacity(void* h) { return with_counting(h, [](auto& t) { return t.base.cap; }); }
u64 wco_counting_size(void* h) { return with_counting(h, [](auto& t) { return t.base.size(); }); }
u32 wco_counting_status(void* h) { return with_counting(h, [](auto& t) { return t.base.status; }); }
void wco_counting_insert(void* h, const void* keys, u64 n, u64 pl, u32* status_out) {
    with_counting(h, [&](auto& t) {
        using K = std::decay_t<decltype(t.base.keys[0])>;
        for (u64 i = 0; i < n; ++i) { u32 s = t.insert(ld<K>(keys, i), pl); if (status_out) status_out[i] = s; }
        return 0;
    });
}
void wco_counting_retrieve(void* h, const void* keys, u64 n, u64* counts_out, u64 pl, u32* status_out) {
    with_counting(h, [&](auto& t) {
        using K = std::decay_t<decltype(t.base.keys[0])>;
        for (u64 i = 0; i < n; ++i) {
            u64 c = 0; u32 s = t.retrieve(ld<K>(keys, i), c, pl);
            if (s == S_NONE) counts_out[i] = c;
            if (status_out) status_out[i] = s;
        }
        return 0;
    });
}

// ---------------------------------------------------------------- hash set --
void* wco_hashset_create(int kb, u64 min_cap, u64 cg, int scheme, u64 seed) {
    Obj* o = new Obj{kb == 32 ? 0 : 2, nullptr};
    if (kb == 32) o->p = new HashSet<u32>(min_cap, cg, scheme, u32(seed));
    else o->p = new HashSet<u64>(min_cap, cg, scheme, seed);
    return o;
}
void wco_hashset_destroy(void* h) { with_set(h, [](auto& t) { delete &t; return 0; }); delete (Obj*)h; }
u64 wco_hashset_capacity(void* h) { return with_set(h, [](auto& t) { return t.t.cap; }); }
u64 wco_hashset_size(void* h) { return with_set(h, [](auto& t) { return t.t.size(); }); }
u32 wco_hashset_status(void* h) { return with_set(h, [](auto& t) { return t.t.status; }); }
void wco_hashset_insert(void* h, const void* keys, u64 n, u64 pl, u32* status_out) {
    with_set(h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>;
        for (u64 i = 0; i < n; ++i) { u32 s = t.insert(ld<K>(keys, i), pl); if (status_out) status_out[i] = s; }
        return 0;
    });
}
void wco_hashset_contains(void* h, const void* keys, u64 n, unsigned char* flags_out, u64 pl, u32* status_out) {
    with_set(h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>;
        for (u64 i = 0; i < n; ++i) {
            bool f = false; u32 s = t.contains(ld<K>(keys, i), f, pl);
            if (s == S_NONE) flags_out[i] = f;
            if (status_out) status_out[i] = s;
        }
        return 0;
    });
}
void wco_hashset_erase(void* h, const void* keys, u64 n, u64 pl, u32* status_out) {
    with_set(h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>;
        for (u64 i = 0; i < n; ++i) { u32 s = t.erase(ld<K>(keys, i), pl); if (status_out) status_out[i] = s; }
        return 0;
    });
}

// ------------------------------------------------------------ bloom filter --
// slot_bits 32 or 64
void* wco_bloom_create(int kb, int slot_bits, u64 num_bits, u64 k, u64 seed, u64 cg) {
    Obj* o = new Obj{kind(kb, slot_bits), nullptr};
    switch (o->kind) {
    case K32V32: o->p = new Bloom<u32, u32>(num_bits, k, u32(seed), cg); break;
    case K32V64: o->p = new Bloom<u32, u64>(num_bits, k, u32(seed), cg); break;
    case K64V32: o->p = new Bloom<u64, u32>(num_bits, k, seed, cg); break;
    default:     o->p = new Bloom<u64, u64>(num_bits, k, seed, cg); break;
    }
    return o;
}
void wco_bloom_destroy(void* h) { with<Bloom>((Obj*)h, [](auto& t) { delete &t; return 0; }); delete (Obj*)h; }
u64 wco_bloom_num_slots(void* h) { return with<Bloom>((Obj*)h, [](auto& t) { return t.num_slots; }); }
void wco_bloom_insert(void* h, const void* keys, u64 n) {
    with<Bloom>((Obj*)h, [&](auto& t) {
        using K = decltype(t.seed);
        for (u64 i = 0; i < n; ++i) t.insert(ld<K>(keys, i));
        return 0;
    });
}
void wco_bloom_query(void* h, const void* keys, u64 n, unsigned char* flags_out) {
    with<Bloom>((Obj*)h, [&](auto& t) {
        using K = decltype(t.seed);
        for (u64 i = 0; i < n; ++i) flags_out[i] = t.query(ld<K>(keys, i));
        return 0;
    });
}
void wco_bloom_words(void* h, void* out) {
    with<Bloom>((Obj*)h, [&](auto& t) { std::memcpy(out, t.words.data(), t.words.size() * sizeof(t.words[0])); return 0; });
}
// slot index and pattern of one key for lane `rank`
void wco_bloom_locate(void* h, u64 key, u64 rank, u64* slot_index, u64* pattern) {
    with<Bloom>((Obj*)h, [&](auto& t) {
        using K = decltype(t.seed);
        u64 s; std::decay_t<decltype(t.words[0])> p; t.locate(K(key), rank, s, p);
        *slot_index = s; *pattern = u64(p); return 0;
    });
}
double wco_bloom_fpr_stable(void* h, u64 n) { return with<Bloom>((Obj*)h, [&](auto& t) { return t.fpr_stable(n); }); }
double wco_bloom_fpr_naive(void* h, u64 n) { return with<Bloom>((Obj*)h, [&](auto& t) { return t.fpr_naive(n); }); }

// -------------------------------------------------------------------- MVHT --
void* wco_mvht_create(int kb, int vb, u64 min_cap, u64 cg, int scheme, u64 seed, u64 max_values) {
    Obj* o = new Obj{kind(kb, vb), nullptr};
    switch (o->kind) {
    case K32V32: o->p = new MVHT<u32, u32>(min_cap, cg, scheme, u32(seed), max_values); break;
    case K32V64: o->p = new MVHT<u32, u64>(min_cap, cg, scheme, u32(seed), max_values); break;
    case K64V32: o->p = new MVHT<u64, u32>(min_cap, cg, scheme, seed, max_values); break;
    default:     o->p = new MVHT<u64, u64>(min_cap, cg, scheme, seed, max_values); break;
    }
    return o;
}
void wco_mvht_destroy(void* h) { with<MVHT>((Obj*)h, [](auto& t) { delete &t; return 0; }); delete (Obj*)h; }
u64 wco_mvht_capacity(void* h) { return with<MVHT>((Obj*)h, [](auto& t) { return t.t.cap; }); }
u64 wco_mvht_size(void* h) { return with<MVHT>((Obj*)h, [](auto& t) { return t.t.size(); }); }
u64 wco_mvht_num_keys(void* h) { return with<MVHT>((Obj*)h, [](auto& t) { return t.num_keys; }); }
u32 wco_mvht_status(void* h) { return with<MVHT>((Obj*)h, [](auto& t) { return t.t.status; }); }
void wco_mvht_insert(void* h, const void* keys, const void* vals, u64 n, u64 pl, u32* status_out) {
    with<MVHT>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>; using V = std::decay_t<decltype(t.t.vals[0])>;
        for (u64 i = 0; i < n; ++i) { u32 s = t.insert(ld<K>(keys, i), ld<V>(vals, i), pl); if (status_out) status_out[i] = s; }
        return 0;
    });
}
// num_values (multi_value_hash_table.cuh:769-795 + gpu_engine.cuh:466-504)
u64 wco_mvht_num_values(void* h, const void* keys, u64 n, u64* per_key_out, u64 pl, u32* status_out) {
    return with<MVHT>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>;
        u64 total = 0;
        for (u64 i = 0; i < n; ++i) {
            u64 c; u32 s = t.for_each(ld<K>(keys, i), nullptr, c, pl);
            if (per_key_out) per_key_out[i] = c;
            total += c;
            if (status_out) status_out[i] = s;
        }
        return total;
    });
}
// two-pass retrieve (:370-510): end = inclusive scan of counts, begin = shift,
// values_out[begin[i] .. end[i]) = values of keys[i] in probe order
u64 wco_mvht_retrieve(void* h, const void* keys, u64 n, u64* begin, u64* end, void* vals_out, u64 pl,
                      u32* status_out) {
    return with<MVHT>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>; using V = std::decay_t<decltype(t.t.vals[0])>;
        u64 off = 0;
        std::vector<V> vs;
        for (u64 i = 0; i < n; ++i) {
            vs.clear(); u64 c; u32 s = t.for_each(ld<K>(keys, i), &vs, c, pl);
            begin[i] = off;
            for (u64 j = 0; j < vs.size(); ++j) st<V>(vals_out, off + j, vs[j]);
            off += c; end[i] = off;
            if (status_out) status_out[i] = s;
        }
        return off;
    });
}

// -------------------------------------------------------------- bucket list --
void* wco_bucketlist_create(int kb, int vb, u64 key_cap, u64 value_cap, u64 cg, int scheme, u64 seed,
                            float grow, u64 min_bucket, u64 max_bucket, u64 max_values) {
    Obj* o = new Obj{kind(kb, vb), nullptr};
    switch (o->kind) {
    case K32V32: o->p = new BucketList<u32, u32>(key_cap, value_cap, cg, scheme, u32(seed), grow, min_bucket, max_bucket, max_values); break;
    case K32V64: o->p = new BucketList<u32, u64>(key_cap, value_cap, cg, scheme, u32(seed), grow, min_bucket, max_bucket, max_values); break;
    case K64V32: o->p = new BucketList<u64, u32>(key_cap, value_cap, cg, scheme, seed, grow, min_bucket, max_bucket, max_values); break;
    default:     o->p = new BucketList<u64, u64>(key_cap, value_cap, cg, scheme, seed, grow, min_bucket, max_bucket, max_values); break;
    }
    return o;
}
void wco_bucketlist_destroy(void* h) { with<BucketList>((Obj*)h, [](auto& t) { delete &t; return 0; }); delete (Obj*)h; }
u64 wco_bucketlist_key_capacity(void* h) { return with<BucketList>((Obj*)h, [](auto& t) { return t.t.cap; }); }
u64 wco_bucketlist_num_keys(void* h) { return with<BucketList>((Obj*)h, [](auto& t) { return t.t.size(); }); }
u64 wco_bucketlist_next_free(void* h) { return with<BucketList>((Obj*)h, [](auto& t) { return t.next_free; }); }
u32 wco_bucketlist_status(void* h) { return with<BucketList>((Obj*)h, [](auto& t) { return t.t.status; }); }
void wco_bucketlist_insert(void* h, const void* keys, const void* vals, u64 n, u64 pl, u32* status_out) {
    with<BucketList>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>; using V = decltype(t.pool[0].value);
        for (u64 i = 0; i < n; ++i) { u32 s = t.insert(ld<K>(keys, i), ld<V>(vals, i), pl); if (status_out) status_out[i] = s; }
        return 0;
    });
}
u64 wco_bucketlist_num_values(void* h, const void* keys, u64 n, u64* per_key_out, u64 pl, u32* status_out) {
    return with<BucketList>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>;
        u64 total = 0;
        for (u64 i = 0; i < n; ++i) {
            u64 c; u32 s = t.retrieve(ld<K>(keys, i), nullptr, c, pl);
            if (per_key_out) per_key_out[i] = c;
            total += c;
            if (status_out) status_out[i] = s;
        }
        return total;
    });
}
u64 wco_bucketlist_retrieve(void* h, const void* keys, u64 n, u64* begin, u64* end, void* vals_out, u64 pl,
                            u32* status_out) {
    return with<BucketList>((Obj*)h, [&](auto& t) {
        using K = std::decay_t<decltype(t.t.keys[0])>; using V = decltype(t.pool[0].value);
        u64 off = 0;
        std::vector<V> vs;
        for (u64 i = 0; i < n; ++i) {
            vs.clear(); u64 c; u32 s = t.retrieve(ld<K>(keys, i), &vs, c, pl);
            begin[i] = off;
            for (u64 j = 0; j < vs.size(); ++j) st<V>(vals_out, off + j, vs[j]);
            off += c; end[i] = off;
            if (status_out) status_out[i] = s;
        }
        return off;
    });
}

}  // extern "C"
