// extern "C" entry points of HashSet<K> (declared in include/warpcore_c.h).
#include "capi_common.hpp"

#include <warpcore/hash_set.hpp>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {

// for_each functor of count_valid: one device counter increment per key
struct CountKeys {
    unsigned long long* counter;
    template<class K>
    __device__ void operator()(const K&) const noexcept { atomicAdd(counter, 1ull); }
};

template<class K>
struct SetApi {
    virtual ~SetApi() = default;
    virtual void init(K seed, bool reseed, hipStream_t s) = 0;
    virtual void insert(const K*, index_t, hipStream_t, index_t, Status*) = 0;
    virtual void contains(const K*, index_t, bool*, hipStream_t, index_t, Status*) = 0;
    virtual void erase(const K*, index_t, hipStream_t, index_t, Status*) = 0;
    virtual void retrieve_all(K*, index_t&, hipStream_t) = 0;
    virtual void count_valid(index_t*, hipStream_t) = 0;
    virtual index_t size(hipStream_t) = 0;
    virtual float load_factor(hipStream_t) = 0;
    virtual index_t capacity() = 0;
    virtual index_t bytes_total() = 0;
    virtual Status peek_status(hipStream_t) = 0;
    virtual Status pop_status(hipStream_t) = 0;
    virtual unsigned cg() = 0;
};

template<class K, unsigned CG>
struct SetImpl final : SetApi<K> {
    using table_t = HashSet<K, defaults::empty_key<K>(), defaults::tombstone_key<K>(), defaults::probing_scheme_t<K, CG>>;
    table_t t;
    SetImpl(index_t min_cap, K seed) : t(min_cap, seed) {}
    void init(K seed, bool reseed, hipStream_t s) override { if (reseed) t.init(seed, s); else t.init(s); }
    void insert(const K* k, index_t n, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template insert<status_handlers::ReturnStatus>(k, n, s, pl, st);
        else t.insert(k, n, s, pl);
    }
    void contains(const K* k, index_t n, bool* f, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template contains<status_handlers::ReturnStatus>(k, n, f, s, pl, st);
        else t.contains(k, n, f, s, pl);
    }
    void erase(const K* k, index_t n, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template erase<status_handlers::ReturnStatus>(k, n, s, pl, st);
        else t.erase(k, n, s, pl);
    }
    void retrieve_all(K* k, index_t& n, hipStream_t s) override { t.retrieve_all(k, n, s); }
    void count_valid(index_t* counter, hipStream_t s) override {
        t.for_each(CountKeys{reinterpret_cast<unsigned long long*>(counter)}, s);
    }
    index_t size(hipStream_t s) override { return t.size(s); }
    float load_factor(hipStream_t s) override { return t.load_factor(s); }
    index_t capacity() override { return t.capacity(); }
    index_t bytes_total() override { return t.bytes_total(); }
    Status peek_status(hipStream_t s) override { return t.peek_status(s); }
    Status pop_status(hipStream_t s) override { return t.pop_status(s); }
    unsigned cg() override { return CG; }
};

template<class K>
SetApi<K>* make(index_t min_cap, K seed, unsigned cg) {
    switch (cg) {
    case 0:
    case 16: return new SetImpl<K, 16>(min_cap, seed);
    case 1: return new SetImpl<K, 1>(min_cap, seed);
    case 2: return new SetImpl<K, 2>(min_cap, seed);
    case 4: return new SetImpl<K, 4>(min_cap, seed);
    case 8: return new SetImpl<K, 8>(min_cap, seed);
    case 32: return new SetImpl<K, 32>(min_cap, seed);
    case 64: return new SetImpl<K, 64>(min_cap, seed);
    default: return nullptr;
    }
}

}  // namespace

#define WC_SET_API(NAME, K)                                                                                       \
    extern "C" void* wc_hashset_##NAME##_create(uint64_t min_capacity, K seed, uint32_t cg_size) {                \
        try { return make<K>(min_capacity, seed, cg_size); } catch (...) { return nullptr; }                      \
    }                                                                                                             \
    extern "C" void wc_hashset_##NAME##_destroy(void* h) { delete static_cast<SetApi<K>*>(h); }                  \
    extern "C" int wc_hashset_##NAME##_init(void* h, void* stream) {                                              \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<SetApi<K>*>(h)->init(K(0), false, as_stream(stream));                                        \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_hashset_##NAME##_init_seed(void* h, K seed, void* stream) {                                 \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<SetApi<K>*>(h)->init(seed, true, as_stream(stream));                                         \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_hashset_##NAME##_insert(void* h, const K* keys, uint64_t n, void* stream, uint64_t pl,      \
                                              uint32_t* status_out) {                                             \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<SetApi<K>*>(h)->insert(keys, n, as_stream(stream), pl, reinterpret_cast<Status*>(status_out)); \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_hashset_##NAME##_contains(void* h, const K* keys, uint64_t n, bool* flags_out,              \
                                                void* stream, uint64_t pl, uint32_t* status_out) {                \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<SetApi<K>*>(h)->contains(keys, n, flags_out, as_stream(stream), pl,                          \
                                             reinterpret_cast<Status*>(status_out));                              \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_hashset_##NAME##_erase(void* h, const K* keys, uint64_t n, void* stream, uint64_t pl,       \
                                             uint32_t* status_out) {                                              \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<SetApi<K>*>(h)->erase(keys, n, as_stream(stream), pl, reinterpret_cast<Status*>(status_out)); \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_hashset_##NAME##_for_each_count(void* h, uint64_t* counter_d, void* stream) {              \
        if (!h || !counter_d) return wc_capi::kBadHandle;                                                          \
        static_cast<SetApi<K>*>(h)->count_valid(counter_d, as_stream(stream));                                    \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_hashset_##NAME##_retrieve_all(void* h, K* keys_out, uint64_t* num_out, void* stream) {      \
        if (!h || !num_out) return wc_capi::kBadHandle;                                                           \
        index_t n = 0;                                                                                            \
        static_cast<SetApi<K>*>(h)->retrieve_all(keys_out, n, as_stream(stream));                                \
        *num_out = n;                                                                                             \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" uint64_t wc_hashset_##NAME##_size(void* h, void* stream) {                                         \
        return h ? static_cast<SetApi<K>*>(h)->size(as_stream(stream)) : 0;                                       \
    }                                                                                                             \
    extern "C" float wc_hashset_##NAME##_load_factor(void* h, void* stream) {                                     \
        return h ? static_cast<SetApi<K>*>(h)->load_factor(as_stream(stream)) : 0.0f;                             \
    }                                                                                                             \
    extern "C" uint64_t wc_hashset_##NAME##_capacity(void* h) { return h ? static_cast<SetApi<K>*>(h)->capacity() : 0; } \
    extern "C" uint64_t wc_hashset_##NAME##_bytes_total(void* h) {                                                \
        return h ? static_cast<SetApi<K>*>(h)->bytes_total() : 0;                                                 \
    }                                                                                                             \
    extern "C" uint32_t wc_hashset_##NAME##_peek_status(void* h, void* stream) {                                  \
        return h ? static_cast<SetApi<K>*>(h)->peek_status(as_stream(stream)).base() : Status::not_initialized().base(); \
    }                                                                                                             \
    extern "C" uint32_t wc_hashset_##NAME##_pop_status(void* h, void* stream) {                                   \
        return h ? static_cast<SetApi<K>*>(h)->pop_status(as_stream(stream)).base() : Status::not_initialized().base(); \
    }                                                                                                             \
    extern "C" uint32_t wc_hashset_##NAME##_cg_size(void* h) { return h ? static_cast<SetApi<K>*>(h)->cg() : 0; }

WC_SET_API(u32, uint32_t)
WC_SET_API(u64, uint64_t)
