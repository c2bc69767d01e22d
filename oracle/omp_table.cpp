// TEST INFRASTRUCTURE / REPORTED CPU BASELINE ONLY.
//
// An OpenMP host table with the semantics of the reference's
// SingleValueHashTable (include/warpcore/single_value_hash_table.cuh:742-823
// insert, :267-307 retrieve) and CountingHashTable
// (include/warpcore/counting_hash_table.cuh:141-172): same fmix hashers,
// cycle-free double hashing over cache-line aligned windows of cg slots, the
// same capacity rule, compare-and-swap on the key through std::atomic_ref,
// duplicate-overwrite, write-on-success and per-op status codes. It is timed
// beside the GPU path by bench.py and is never part of the product.
#include "wc_oracle.hpp"

#include <atomic>
#include <chrono>
#include <cstring>
#include <omp.h>

using namespace wco;

namespace {

template<class K, class V>
struct Pair { K key; V value; };

template<class K, class V>
struct OmpTable {
    u64 cg, cap, p; K seed; V max_count;
    Pair<K, V>* slots;
    OmpTable(u64 min_cap, u64 cg_, K seed_) : cg(cg_), cap(get_valid_capacity(min_cap, cg_)), p(cap / cg_), seed(seed_), max_count(~V(0)) {
        slots = static_cast<Pair<K, V>*>(std::aligned_alloc(128, ((cap * sizeof(Pair<K, V>) + 127) / 128) * 128));
        init();
    }
    ~OmpTable() { std::free(slots); }
    void init() {
#pragma omp parallel for schedule(static)
        for (u64 i = 0; i < cap; ++i) { slots[i].key = K(0); slots[i].value = V(0); }
    }
    static bool valid(K k) { return k != K(0) && k != K(~K(0)); }

    // returns slot index or -1; status as the reference's insert_impl
    long long insert_impl(K k, u32& stc) {
        if (!valid(k)) { stc = S_INVALID_KEY; return -1; }
        u64 w = u64(murmur<K>(K(k + seed))) % p;
        const u64 step = u64(murmur<K>(K(k + seed + K(1)))) % (p - 1) + 1;
        for (u64 win = 0; win < p; ++win) {
            Pair<K, V>* base = slots + w * cg;
            for (u64 r = 0; r < cg; ++r)
                if (std::atomic_ref<K>(base[r].key).load(std::memory_order_relaxed) == k) { stc = S_DUP; return (long long)(w * cg + r); }
            for (u64 r = 0; r < cg; ++r) {
                std::atomic_ref<K> ak(base[r].key);
                K expected = ak.load(std::memory_order_relaxed);
                while (expected == K(0)) {
                    if (ak.compare_exchange_strong(expected, k, std::memory_order_acq_rel)) { stc = S_NONE; return (long long)(w * cg + r); }
                }
                if (expected == k) { stc = S_DUP; return (long long)(w * cg + r); }
            }
            w += step; if (w >= p) w -= p;
        }
        stc = S_PLE; return -1;
    }
    u32 insert(K k, V v) {
        u32 s; long long i = insert_impl(k, s);
        if (i >= 0) std::atomic_ref<V>(slots[i].value).store(v, std::memory_order_relaxed);
        return s;
    }
    u32 count(K k) {
        u32 s; long long i = insert_impl(k, s);
        if (i >= 0) {
            std::atomic_ref<V> av(slots[i].value);
            if (av.load(std::memory_order_relaxed) < max_count) av.fetch_add(V(1), std::memory_order_relaxed);
        }
        return s;
    }
    u32 retrieve(K k, V& out) const {
        if (!valid(k)) return S_INVALID_KEY;
        u64 w = u64(murmur<K>(K(k + seed))) % p;
        const u64 step = u64(murmur<K>(K(k + seed + K(1)))) % (p - 1) + 1;
        for (u64 win = 0; win < p; ++win) {
            const Pair<K, V>* base = slots + w * cg;
            bool empty = false;
            for (u64 r = 0; r < cg; ++r) {
                const K t = base[r].key;
                if (t == k) { out = base[r].value; return S_NONE; }
                empty |= t == K(0);
            }
            if (empty) return S_NOT_FOUND;
            w += step; if (w >= p) w -= p;
        }
        return S_PLE;
    }
};

struct Obj { int kb, vb; void* p; };

template<class F> auto with(void* h, F f) {
    Obj* o = (Obj*)h;
    if (o->kb == 32 && o->vb == 32) return f(*static_cast<OmpTable<u32, u32>*>(o->p));
    if (o->kb == 32) return f(*static_cast<OmpTable<u32, u64>*>(o->p));
    if (o->vb == 32) return f(*static_cast<OmpTable<u64, u32>*>(o->p));
    return f(*static_cast<OmpTable<u64, u64>*>(o->p));
}

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

}  // namespace

extern "C" {

int wco_omp_max_threads() { return omp_get_max_threads(); }

void* wco_omp_create(int kb, int vb, u64 min_cap, u64 cg, u64 seed) {
    Obj* o = new Obj{kb, vb, nullptr};
    if (kb == 32 && vb == 32) o->p = new OmpTable<u32, u32>(min_cap, cg, u32(seed));
    else if (kb == 32) o->p = new OmpTable<u32, u64>(min_cap, cg, u32(seed));
    else if (vb == 32) o->p = new OmpTable<u64, u32>(min_cap, cg, seed);
    else o->p = new OmpTable<u64, u64>(min_cap, cg, seed);
    return o;
}
void wco_omp_destroy(void* h) { with(h, [](auto& t) { delete &t; return 0; }); delete (Obj*)h; }
u64 wco_omp_capacity(void* h) { return with(h, [](auto& t) { return t.cap; }); }
void wco_omp_init(void* h) { with(h, [](auto& t) { t.init(); return 0; }); }

// Each call returns the wall time of the parallel region in seconds.
double wco_omp_insert(void* h, const void* keys, const void* vals, u64 n, u32* status_out) {
    return with(h, [&](auto& t) {
        using K = decltype(t.seed); using V = decltype(t.max_count);
        const K* ks = (const K*)keys; const V* vs = (const V*)vals;
        const double t0 = now();
#pragma omp parallel for schedule(dynamic, 4096)
        for (u64 i = 0; i < n; ++i) { u32 s = t.insert(ks[i], vs[i]); if (status_out) status_out[i] = s; }
        return now() - t0;
    });
}
double wco_omp_count(void* h, const void* keys, u64 n, u32* status_out) {
    return with(h, [&](auto& t) {
        using K = decltype(t.seed);
        const K* ks = (const K*)keys;
        const double t0 = now();
#pragma omp parallel for schedule(dynamic, 4096)
        for (u64 i = 0; i < n; ++i) { u32 s = t.count(ks[i]); if (status_out) status_out[i] = s; }
        return now() - t0;
    });
}
double wco_omp_retrieve(void* h, const void* keys, u64 n, void* vals_out, u32* status_out) {
    return with(h, [&](auto& t) {
        using K = decltype(t.seed); using V = decltype(t.max_count);
        const K* ks = (const K*)keys; V* vo = (V*)vals_out;
        const double t0 = now();
#pragma omp parallel for schedule(dynamic, 4096)
        for (u64 i = 0; i < n; ++i) {
            V v; u32 s = t.retrieve(ks[i], v);
            if (s == S_NONE) vo[i] = v;
            if (status_out) status_out[i] = s;
        }
        return now() - t0;
    });
}

}  // extern "C"
