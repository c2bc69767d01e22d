// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// A sequential CPU restatement of the semantics of sleeepyjack/warpcore's
// hash-table hot path. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker (or, for the OpenMP
// table, as the reported CPU baseline). The product path never calls it.
//
// Every function cites the reference file:line whose behaviour it restates.
// Where the reference does not terminate (insert into a table with no EMPTY
// slot under the default probing length, include/warpcore/defaults.cuh:27,
// probing_schemes.cuh:94-100) this restatement stops after one full cycle of
// windows and reports probing_length_exceeded, exactly like the GPU build.
#pragma once

#include <cstdint>
#include <vector>
#include <cmath>
#include <algorithm>

namespace wco {

using u32 = std::uint32_t;
using u64 = std::uint64_t;

// ---------------------------------------------------------------- hashers --
// MurmurHash3 fmix32: include/warpcore/hashers.cuh:116-124
inline u32 fmix32(u32 x) {
    x ^= x >> 16; x *= 0x85ebca6bu; x ^= x >> 13; x *= 0xc2b2ae35u; x ^= x >> 16;
    return x;
}
// MurmurHash3 fmix64: include/warpcore/hashers.cuh:131-139
inline u64 fmix64(u64 x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}
// NvidiaHash: include/warpcore/hashers.cuh:34-44
inline u32 nvidia_hash(u32 x) {
    x = (x + 0x7ed55d16u) + (x << 12);
    x = (x ^ 0xc761c23cu) ^ (x >> 19);
    x = (x + 0x165667b1u) + (x << 5);
    x = (x + 0xd3a2646cu) ^ (x << 9);
    x = (x + 0xfd7046c5u) + (x << 3);
    x = (x ^ 0xb55a4f09u) ^ (x >> 16);
    return x;
}
// MuellerHash: include/warpcore/hashers.cuh:70-77
inline u32 mueller_hash(u32 x) {
    x = ((x >> 16) ^ x) * 0x45d9f3bu;
    x = ((x >> 16) ^ x) * 0x45d9f3bu;
    x = ((x >> 16) ^ x);
    return x;
}

template<class K> inline K murmur(K x);
template<> inline u32 murmur<u32>(u32 x) { return fmix32(x); }
template<> inline u64 murmur<u64>(u64 x) { return fmix64(x); }

// Hasher template argument of a probing scheme: MurmurHash (hashers.cuh:85-141),
// MuellerHash (:50-79), NvidiaHash (:14-46), IdentityMap (:147-175). The
// Mueller and Nvidia hashers exist for 32-bit keys only.
enum Hasher : int { H_MURMUR = 0, H_MUELLER = 1, H_NVIDIA = 2, H_IDENTITY = 3 };
template<class K>
inline K hash_with(int hasher, K x) {
    switch (hasher) {
    case H_MUELLER: return K(mueller_hash(u32(x)));
    case H_NVIDIA: return K(nvidia_hash(u32(x)));
    case H_IDENTITY: return x;
    default: return murmur<K>(x);
    }
}

// ----------------------------------------------------------------- status --
// bit positions: include/warpcore/status.cuh:44-68
enum : u32 {
    S_NONE = 0u,
    S_UNKNOWN = 1u << 0,
    S_PLE = 1u << 1,             // probing_length_exceeded
    S_INVALID_CONFIG = 1u << 2,
    S_INVALID_KEY = 1u << 3,
    S_DUP = 1u << 4,             // duplicate_key
    S_NOT_FOUND = 1u << 5,       // key_not_found
    S_INDEX_OVERFLOW = 1u << 6,
    S_OOM = 1u << 7,
    S_NOT_INIT = 1u << 8,
    S_DRY_RUN = 1u << 9,
    S_PHASE = 1u << 10,          // invalid_phase_overlap
    S_MAXV = 1u << 11,           // max_values_for_key_reached
    S_INVALID_VALUE = 1u << 12,
};

// Status algebra, include/warpcore/status.cuh: mask 0xFFF (:207), so
// all() = 0xFFF (:40) and invalid_value is outside it; error_mask (:71-80);
// warning_mask = all() - error_mask() with a - b = a - (a & b) (:82-86,
// :160-170); ~s = (~s) ^ mask (:146-150); the flag ReturnBoolean<Ignore>
// writes is has_any(~Status(Ignore)) (:316-333).
constexpr u32 STATUS_MASK = 0xFFFu;
inline u32 status_all() { return STATUS_MASK; }
inline u32 status_error_mask() {
    return S_UNKNOWN | S_INVALID_CONFIG | S_INDEX_OVERFLOW | S_OOM | S_NOT_INIT | S_PHASE;
}
inline u32 status_remove(u32 a, u32 b) { return a - (a & b); }
inline u32 status_warning_mask() { return status_remove(status_all(), status_error_mask()); }
inline u32 status_not(u32 s) { return (~s) ^ STATUS_MASK; }
inline bool status_return_boolean(u32 s, u32 ignore) { return (s & status_not(ignore)) != 0; }

// ----------------------------------------------------------------- primes --
// Capacity rule include/warpcore/base.cuh:27-34 over the prime table whose
// rule is utils/generate_primes_hpp.py:20-43 (first prime >= 131072, then the
// next prime >= previous + 131072, up to 17177772037).
u64 num_primes();
u64 prime_at(u64 i);
bool is_prime_u64(u64 n);
u64 get_valid_capacity(u64 min_capacity, u64 cg);

// -------------------------------------------------------- probing schemes --
enum Scheme : int {
    REF_DOUBLE = 0,     // probing_schemes.cuh:79-100 (unaligned windows)
    REF_LINEAR = 1,     // probing_schemes.cuh:173-192
    REF_QUADRATIC = 2,  // probing_schemes.cuh:264-285
    AL_DOUBLE = 3,      // same hashes, windows aligned to cg slots
    AL_LINEAR = 4,
    AL_QUADRATIC = 5,
};

// Iterates the windows of one key. slots[r] is lane r's slot in the current
// window. Window count: SDIV(pl, cg) (probing_schemes.cuh:61), capped at one
// full cycle capacity/cg.
// `scheme_` carries the Hasher in bits 4..7 (0 = MurmurHash).
template<class K>
struct Windows {
    int scheme, hasher; u64 cap, cg, p, left;
    u64 pos[64];
    u64 step;  // double: step in slots (ref) or windows (aligned); quad: counter
    u64 w;     // aligned window index

    Windows(int scheme_, u64 cap_, u64 cg_, u64 pl) : scheme(scheme_ & 15), hasher(scheme_ >> 4), cap(cap_), cg(cg_) {
        p = cap / cg;
        // The first window is always probed (begin() is not checked against
        // the probing length, single_value_hash_table.cuh:764); next() stops
        // once rank + i*cg >= SDIV(pl, cg)*cg, so max(1, SDIV(pl, cg)) windows.
        u64 windows = (pl > ~u64(0) - cg) ? (~u64(0)) / cg + 1 : (pl + cg - 1) / cg;
        if (windows == 0) windows = 1;
        left = std::min(windows, p);
    }
    bool begin(K key, K seed) {
        // probing_schemes.cuh:79-81 (double), :176 (linear), :267 (quadratic)
        const K h1 = hash_with<K>(hasher, K(key + seed));
        const K h2 = hash_with<K>(hasher, K(key + seed + K(1)));
        switch (scheme) {
        case REF_DOUBLE:
        case REF_LINEAR:
        case REF_QUADRATIC:
            // pos = (h1 + rank) computed in the hash's width, then % capacity
            for (u64 r = 0; r < cg; ++r) pos[r] = u64(K(h1 + K(r))) % cap;
            if (scheme == REF_DOUBLE) step = (u64(h2) % (p - 1) + 1) * cg;
            else if (scheme == REF_LINEAR) step = cg;
            else step = 1;
            break;
        default:
            w = u64(h1) % p;
            if (scheme == AL_DOUBLE) step = u64(h2) % (p - 1) + 1;
            else if (scheme == AL_LINEAR) step = 1;
            else step = 1;
            for (u64 r = 0; r < cg; ++r) pos[r] = w * cg + r;
        }
        return true;
    }
    bool next() {
        if (--left == 0) return false;
        switch (scheme) {
        case REF_DOUBLE:
        case REF_LINEAR:
            for (u64 r = 0; r < cg; ++r) pos[r] = (pos[r] + step) % cap;
            break;
        case REF_QUADRATIC:
            for (u64 r = 0; r < cg; ++r) pos[r] = (pos[r] + step) % cap;
            ++step;
            break;
        case AL_DOUBLE:
        case AL_LINEAR:
            w = (w + step) % p;
            for (u64 r = 0; r < cg; ++r) pos[r] = w * cg + r;
            break;
        default:
            w = (w + step) % p;
            ++step;
            for (u64 r = 0; r < cg; ++r) pos[r] = w * cg + r;
        }
        return true;
    }
};

// ------------------------------------------------------ single-value table --
// SingleValueHashTable semantics, include/warpcore/single_value_hash_table.cuh
template<class K, class V>
struct SVHT {
    static constexpr K EMPTY = K(0);        // defaults.cuh:11-13
    static constexpr K TOMB = K(~K(0));     // defaults.cuh:15-17
    int scheme; u64 cg, cap; K seed; u32 status = S_NONE; bool init_ok;
    std::vector<K> keys; std::vector<V> vals;

    SVHT(u64 min_cap, u64 cg_, int scheme_, K seed_)
        : scheme(scheme_), cg(cg_), cap(get_valid_capacity(min_cap, cg_)), seed(seed_) {
        init_ok = cap != 0;
        if (!init_ok) status = S_INVALID_CONFIG | S_NOT_INIT;  // storage.cuh:475-478
        keys.assign(cap, EMPTY); vals.assign(cap, V(0));
    }
    static bool valid(K k) { return k != EMPTY && k != TOMB; }  // :711-714

    // insert_impl :742-823; returns slot or -1
    long long insert_impl(K k, u32& st, u64 pl) {
        if (!init_ok) { st = S_NOT_INIT; return -1; }
        if (!valid(k)) { status |= S_INVALID_KEY; st = S_INVALID_KEY; return -1; }
        Windows<K> it(scheme, cap, cg, pl);
        for (bool ok = it.begin(k, seed); ok; ok = it.next()) {
            for (u64 r = 0; r < cg; ++r)
                if (keys[it.pos[r]] == k) { status |= S_DUP; st = S_DUP; return (long long)it.pos[r]; }
            for (u64 r = 0; r < cg; ++r)
                if (keys[it.pos[r]] == EMPTY) { keys[it.pos[r]] = k; st = S_NONE; return (long long)it.pos[r]; }
        }
        status |= S_PLE; st = S_PLE; return -1;
    }
    // insert :210-228
    u32 insert(K k, V v, u64 pl) {
        u32 st; long long s = insert_impl(k, st, pl);
        if (s >= 0) vals[s] = v;
        return st;
    }
    // retrieve :267-307
    u32 retrieve(K k, V& out, u64 pl) {
        if (!init_ok) return S_NOT_INIT;
        if (!valid(k)) { status |= S_INVALID_KEY; return S_INVALID_KEY; }
        Windows<K> it(scheme, cap, cg, pl);
        for (bool ok = it.begin(k, seed); ok; ok = it.next()) {
            for (u64 r = 0; r < cg; ++r)
                if (keys[it.pos[r]] == k) { out = vals[it.pos[r]]; return S_NONE; }
            for (u64 r = 0; r < cg; ++r)
                if (keys[it.pos[r]] == EMPTY) { status |= S_NOT_FOUND; return S_NOT_FOUND; }
        }
        status |= S_PLE; return S_PLE;
    }
    // erase :425-468 (misses are not joined)
    u32 erase(K k, u64 pl) {
        if (!init_ok) return S_NOT_INIT;
        if (!valid(k)) { status |= S_INVALID_KEY; return S_INVALID_KEY; }
        Windows<K> it(scheme, cap, cg, pl);
        for (bool ok = it.begin(k, seed); ok; ok = it.next()) {
            for (u64 r = 0; r < cg; ++r)
                if (keys[it.pos[r]] == k) { keys[it.pos[r]] = TOMB; return S_NONE; }
            for (u64 r = 0; r < cg; ++r)
                if (keys[it.pos[r]] == EMPTY) return S_NOT_FOUND;
        }
        return S_PLE;
    }
    u64 size() const {  // :534-592
        u64 n = 0; for (auto k : keys) n += valid(k); return n;
    }
};


// --------------------------------------------------------- counting table --
// CountingHashTable, include/warpcore/counting_hash_table.cuh
template<class K, class V>
struct Counting {
    SVHT<K, V> base; V max_count;
    Counting(u64 min_cap, u64 cg, int scheme, K seed, V max_count_)
        : base(min_cap, cg, scheme, seed), max_count(max_count_) {}  // init values := 0, :113-124
    // insert :141-172 (sequentially `*v < max_count` keeps the counter from wrapping)
    u32 insert(K k, u64 pl) {
        u32 st; long long s = base.insert_impl(k, st, pl);
        if (s >= 0 && base.vals[s] < max_count) base.vals[s] += 1;
        return st;
    }
    // retrieve :203-225
    u32 retrieve(K k, V& out, u64 pl) {
        V c; u32 st = base.retrieve(k, c, pl);
        if (st == S_NONE) out = c > max_count ? max_count : c;
        return st;
    }
};

// --------------------------------------------------------------- hash set --
// HashSet, include/warpcore/hash_set.cuh
template<class K>
struct HashSet {
    SVHT<K, u32> t;
    HashSet(u64 min_cap, u64 cg, int scheme, K seed) : t(min_cap, cg, scheme, seed) {}
    u32 insert(K k, u64 pl) { u32 st; t.insert_impl(k, st, pl); return st; }  // :192-254
    // device retrieve :291-329: hit -> true, window with EMPTY -> false, both `none`
    u32 contains(K k, bool& flag, u64 pl) {
        if (!t.init_ok) return S_NOT_INIT;
        if (!t.valid(k)) { t.status |= S_INVALID_KEY; return S_INVALID_KEY; }
        Windows<K> it(t.scheme, t.cap, t.cg, pl);
        for (bool ok = it.begin(k, t.seed); ok; ok = it.next()) {
            for (u64 r = 0; r < t.cg; ++r) if (t.keys[it.pos[r]] == k) { flag = true; return S_NONE; }
            for (u64 r = 0; r < t.cg; ++r) if (t.keys[it.pos[r]] == t.EMPTY) { flag = false; return S_NONE; }
        }
        t.status |= S_PLE; return S_PLE;
    }
    // erase :395-438 always returns none after the key check
    u32 erase(K k, u64 pl) {
        u32 st = t.erase(k, pl);
        return (st == S_INVALID_KEY || st == S_NOT_INIT) ? st : S_NONE;
    }
};

// ----------------------------------------------------------- bloom filter --
// BloomFilter, include/warpcore/bloom_filter.cuh (pattern-blocked)
template<class K, class Slot>
struct Bloom {
    u64 num_bits, num_slots, num_blocks, k, cg; K seed;
    std::vector<Slot> words;
    static constexpr u64 slot_bits = sizeof(Slot) * 8;
    Bloom(u64 num_bits_, u64 k_, K seed_, u64 cg_)
        : num_bits(num_bits_), num_slots((num_bits_ + slot_bits - 1) / slot_bits),
          num_blocks((num_slots + cg_ - 1) / cg_), k(k_), cg(cg_), seed(seed_) {  // :77-93
        words.assign(num_slots, Slot(0));
    }
    u64 block_bits() const { return slot_bits * cg; }
    // slot index :172-174, pattern :176-186
    void locate(K key, u64 rank, u64& slot_index, Slot& pattern) const {
        slot_index = ((u64(murmur<K>(K(key + seed))) % num_blocks) * cg + rank) % num_slots;
        pattern = 0;
        for (u64 j = 0; j < k; ++j) {
            const K seeded = K(u64(K(key + seed)) + j);
            const u64 h = u64(murmur<K>(seeded)) % block_bits();
            if (h / slot_bits == rank) pattern |= Slot(1) << (h % slot_bits);
        }
    }
    void insert(K key) {  // :167-192
        for (u64 r = 0; r < cg; ++r) { u64 s; Slot pat; locate(key, r, s, pat); words[s] |= pat; }
    }
    bool query(K key) const {  // :215-237
        for (u64 r = 0; r < cg; ++r) {
            u64 s; Slot pat; locate(key, r, s, pat);
            if ((words[s] & pat) != pat) return false;
        }
        return true;
    }
    // fpr model :341-408 evaluated in log space so it stays finite
    double fpr_stable(u64 n) const {
        const double b = double(num_bits / block_bits());
        const u64 m = u64(double(num_bits) / b);
        const u64 imax = 5 * n / (num_bits / block_bits());
        const double p = 1.0 / b;
        double res = 0.0;
        for (u64 i = 0; i < imax; ++i) {
            if (i > n) break;
            const double lb = std::lgamma(double(n) + 1) - std::lgamma(double(i) + 1) -
                              std::lgamma(double(n - i) + 1) + double(i) * std::log(p) +
                              double(n - i) * std::log1p(-p);
            const double base = std::pow(1.0 - std::pow(1.0 - 1.0 / double(m), double(i * k)), double(k));
            res += std::exp(lb) * base;
        }
        return res;
    }
    // the reference's own product form :341-393 (overflows for large n)
    double fpr_naive(u64 n) const {
        const double b = double(num_bits / block_bits());
        const u64 m = u64(double(num_bits) / b);
        double res = 0.0;
        for (u64 i = 0; i < 5 * n / (num_bits / block_bits()); ++i) {
            double bi = 1.0;
            for (u64 t = n - i + 1; t <= n; ++t) bi *= double(t);
            for (u64 t = 1; t <= i; ++t) bi /= double(t);
            bi *= std::pow(1.0 / b, double(i)) * std::pow(1.0 - 1.0 / b, double(n - i));
            res += bi * std::pow(1.0 - std::pow(1.0 - 1.0 / double(m), double(i * k)), double(k));
        }
        return res;
    }
};

// ------------------------------------------------------ multi-value table --
// MultiValueHashTable, include/warpcore/multi_value_hash_table.cuh
template<class K, class V>
struct MVHT {
    SVHT<K, V> t; u64 max_values; u64 num_keys = 0;
    MVHT(u64 min_cap, u64 cg, int scheme, K seed, u64 maxv)
        : t(min_cap, cg, scheme, seed), max_values(maxv) {}
    // insert :200-289 (a successful insert joins nothing)
    u32 insert(K k, V v, u64 pl) {
        if (!t.init_ok) return S_NOT_INIT;
        if (!t.valid(k)) { t.status |= S_INVALID_KEY; return S_INVALID_KEY; }
        Windows<K> it(t.scheme, t.cap, t.cg, pl);
        u64 nv = 0;
        for (bool ok = it.begin(k, t.seed); ok; ok = it.next()) {
            for (u64 r = 0; r < t.cg; ++r) nv += t.keys[it.pos[r]] == k;
            if (nv >= max_values) { t.status |= S_DUP | S_MAXV; return S_DUP | S_MAXV; }
            for (u64 r = 0; r < t.cg; ++r) {
                const u64 s = it.pos[r];
                if (t.keys[s] == t.EMPTY) {
                    t.keys[s] = k; t.vals[s] = v;
                    if (nv == 0) ++num_keys;
                    return nv > 0 ? S_DUP : S_NONE;
                }
            }
        }
        const u32 st = nv > 0 ? (S_PLE | S_DUP) : S_PLE;
        t.status |= st; return st;
    }
    // for_each :605-661 with probing capped at capacity; values in probe order
    u32 for_each(K k, std::vector<V>* out, u64& num, u64 pl) {
        num = 0;
        if (!t.init_ok) return S_NOT_INIT;
        if (!t.valid(k)) { t.status |= S_INVALID_KEY; return S_INVALID_KEY; }
        Windows<K> it(t.scheme, t.cap, t.cg, std::min(pl, t.cap));
        for (bool ok = it.begin(k, t.seed); ok; ok = it.next()) {
            bool any_empty = false;
            for (u64 r = 0; r < t.cg; ++r) {
                const u64 s = it.pos[r];
                if (t.keys[s] == k) { if (out) out->push_back(t.vals[s]); ++num; }
                any_empty |= t.keys[s] == t.EMPTY;
            }
            if (any_empty || num >= max_values) {
                if (num == 0) { t.status |= S_NOT_FOUND; return S_NOT_FOUND; }
                return S_NONE;
            }
        }
        t.status |= S_PLE; return S_PLE;
    }
};

// ------------------------------------------------------ bucket-list table --
// BucketListHashTable + BucketListStore: include/warpcore/bucket_list_hash_table.cuh,
// include/warpcore/storage.cuh:656-1454 (handle = {state, bucket index,
// value count, values in the tail bucket}; a bucket slot is a value or a
// header {previous bucket, bucket size})
template<class K, class V>
struct BucketList {
    enum State { UNINIT = 0, INIT = 1, BLOCKING = 2, FULL = 3 };
    struct Handle { int state; u64 index, count, tail; };
    struct Slot { V value; u64 prev, size; };
    static constexpr u64 HEAD = (u64(1) << 32) - 1;          // storage.cuh:957-960
    static constexpr u64 MAX_COUNT = (u64(1) << 20) - 1;     // :810-813
    static constexpr u64 MAX_BUCKET = (u64(1) << 10) - 1;    // :816-819
    SVHT<K, u32> t;   // vals[] index handles[]
    std::vector<Handle> handles;
    std::vector<Slot> pool;
    u64 value_cap, next_free = 0, min_bucket, max_bucket, max_values; float grow;
    bool pool_ok;
    BucketList(u64 key_cap, u64 value_cap_, u64 cg, int scheme, K seed, float grow_,
               u64 minb, u64 maxb, u64 maxv)
        : t(key_cap, cg, scheme, seed), value_cap(value_cap_), min_bucket(minb), max_bucket(maxb),
          max_values(std::min(maxv, MAX_COUNT)), grow(grow_) {
        // store ctor :989-1029
        pool_ok = value_cap < HEAD && grow >= 1.0f && min_bucket >= 1 && max_bucket >= min_bucket &&
                  max_bucket <= MAX_BUCKET;
        if (!pool_ok) t.status |= S_INVALID_CONFIG | S_NOT_INIT;
        handles.assign(t.cap, Handle{UNINIT, HEAD, 0, 0});
        pool.assign(pool_ok ? value_cap : 0, Slot{V(0), HEAD, min_bucket});
        for (u64 s = 0; s < t.cap; ++s) t.vals[s] = u32(s);
    }
    // append :1113-1284, run without contention
    u32 append(Handle& h, V v) {
        if (h.state == UNINIT) {
            const u64 alloc = next_free; next_free += min_bucket;
            if (alloc + min_bucket <= value_cap) {
                pool[alloc].value = v; h = Handle{INIT, alloc, 1, 1}; return S_NONE;
            }
            h = Handle{FULL, HEAD, 0, 0}; return S_OOM;
        }
        if (h.state == FULL) return S_OOM;
        if (h.count == max_values) return S_MAXV;
        const u64 cbs = (h.count <= min_bucket) ? min_bucket : pool[h.index].size;
        if (h.tail == cbs) {
            const u64 nbs = u64(std::min(float(max_bucket), std::ceil(float(cbs) * grow)));
            const u64 alloc = next_free; next_free += nbs + 1;
            if (alloc + nbs + 1 <= value_cap) {
                pool[alloc + 1].value = v;
                pool[alloc].prev = h.index; pool[alloc].size = nbs;
                h = Handle{INIT, alloc, h.count + 1, 1};
                return S_NONE;
            }
            h = Handle{FULL, h.index, h.count, h.tail}; return S_OOM;
        }
        const u64 j = (h.count + 1 <= min_bucket) ? h.tail : h.tail + 1;
        pool[h.index + j].value = v;
        h = Handle{INIT, h.index, h.count + 1, h.tail + 1};
        return S_NONE;
    }
    // device insert :184-223
    u32 insert(K k, V v, u64 pl) {
        if (!pool_ok) return S_NOT_INIT;
        u32 st; long long s = t.insert_impl(k, st, pl);
        if (s < 0) return st;
        Handle& h = handles[s];
        if (h.count >= max_values) { t.status |= S_MAXV; return st | S_MAXV; }
        const u32 a = append(h, v);
        if (a) t.status |= a;
        return st | a;
    }
    // values of one handle, tail bucket first (storage.cuh:1292-1345); a handle
    // left FULL by an out-of-memory append still yields the values it holds
    void values(const Handle& h, std::vector<V>& out) const {
        if ((h.state != INIT && h.state != FULL) || h.index == HEAD) return;
        u64 idx = h.index, off = (h.count <= min_bucket) ? 0 : 1, taken = 0;
        for (u64 i = 0; i < h.tail; ++i) out.push_back(pool[idx + off + i].value);
        taken = h.tail;
        while (taken < h.count) {
            idx = pool[idx].prev;
            const bool last = taken >= h.count - min_bucket;
            const u64 size = last ? min_bucket : pool[idx].size;
            const u64 o = last ? 0 : 1;
            for (u64 i = 0; i < size && taken < h.count; ++i, ++taken) out.push_back(pool[idx + o + i].value);
        }
    }
    // device retrieve :349-382 via the key table's retrieve
    u32 retrieve(K k, std::vector<V>* out, u64& num, u64 pl) {
        u32 slot; u32 st = t.retrieve(k, slot, pl);
        num = 0;
        if (st == S_NONE) {
            num = handles[slot].count;
            if (out) values(handles[slot], *out);
        }
        return st;
    }
};

}  // namespace wco
