// Shared part of the SingleValueHashTable extern "C" layer: the virtual
// handle interface every wc_svht_* entry point dispatches through, its
// implementation over any SingleValueHashTable instantiation, and the macro
// that defines one (key, value) family of entry points (declared in
// include/warpcore_c.h).
#pragma once

#include "capi_common.hpp"

namespace wc_capi {

using namespace warpcore;

template<class K, class V>
struct SvhtApi {
    virtual ~SvhtApi() = default;
    virtual void init(K seed, bool reseed, hipStream_t s) = 0;
    virtual void insert(const K*, const V*, index_t, hipStream_t, index_t, Status*) = 0;
    virtual void insert_bool(const K*, const V*, index_t, hipStream_t, index_t, bool*) = 0;
    virtual void retrieve(const K*, index_t, V*, hipStream_t, index_t, Status*) = 0;
    virtual void retrieve_bool(const K*, index_t, V*, hipStream_t, index_t, bool*) = 0;
    virtual void erase(const K*, index_t, hipStream_t, index_t, Status*) = 0;
    virtual void retrieve_all(K*, V*, index_t&, hipStream_t) = 0;
    virtual void count_valid(index_t*, hipStream_t) = 0;
    virtual index_t size(hipStream_t) = 0;
    virtual float load_factor(hipStream_t) = 0;
    virtual index_t capacity() = 0;
    virtual index_t bytes_total() = 0;
    virtual Status peek_status(hipStream_t) = 0;
    virtual Status pop_status(hipStream_t) = 0;
    virtual bool is_initialized() = 0;
    virtual unsigned cg() = 0;
};

// for_each functor of count_valid: one device counter increment per pair
struct CountPairs {
    unsigned long long* counter;
    template<class K, class V>
    __device__ void operator()(const K&, const V&) const noexcept { atomicAdd(counter, 1ull); }
};

template<class Table>
struct SvhtImpl final : SvhtApi<typename Table::key_type, typename Table::value_type> {
    using K = typename Table::key_type;
    using V = typename Table::value_type;
    Table t;
    SvhtImpl(index_t min_cap, K seed, bool no_init) : t(min_cap, seed, no_init) {}
    void init(K seed, bool reseed, hipStream_t s) override { if (reseed) t.init(seed, s); else t.init(s); }
    void insert(const K* k, const V* v, index_t n, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template insert<status_handlers::ReturnStatus>(k, v, n, s, pl, st);
        else t.insert(k, v, n, s, pl);
    }
    void insert_bool(const K* k, const V* v, index_t n, hipStream_t s, index_t pl, bool* f) override {
        t.template insert<status_handlers::ReturnBoolean<>>(k, v, n, s, pl, f);
    }
    void retrieve(const K* k, index_t n, V* out, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template retrieve<status_handlers::ReturnStatus>(k, n, out, s, pl, st);
        else t.retrieve(k, n, out, s, pl);
    }
    void retrieve_bool(const K* k, index_t n, V* out, hipStream_t s, index_t pl, bool* f) override {
        t.template retrieve<status_handlers::ReturnBoolean<>>(k, n, out, s, pl, f);
    }
    void erase(const K* k, index_t n, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template erase<status_handlers::ReturnStatus>(k, n, s, pl, st);
        else t.erase(k, n, s, pl);
    }
    void retrieve_all(K* k, V* v, index_t& n, hipStream_t s) override { t.retrieve_all(k, v, n, s); }
    void count_valid(index_t* counter, hipStream_t s) override {
        t.for_each(CountPairs{reinterpret_cast<unsigned long long*>(counter)}, s);
    }
    index_t size(hipStream_t s) override { return t.size(s); }
    float load_factor(hipStream_t s) override { return t.load_factor(s); }
    index_t capacity() override { return t.capacity(); }
    index_t bytes_total() override { return t.bytes_total(); }
    Status peek_status(hipStream_t s) override { return t.peek_status(s); }
    Status pop_status(hipStream_t s) override { return t.pop_status(s); }
    bool is_initialized() override { return t.is_initialized(); }
    unsigned cg() override { return unsigned(Table::cg_size()); }
};

template<class K, class V, unsigned CG>
using DefaultSvht = SingleValueHashTable<K, V, defaults::empty_key<K>(), defaults::tombstone_key<K>(),
                                         defaults::probing_scheme_t<K, CG>>;

}  // namespace wc_capi

#define WC_SVHT_API(NAME, K, V, MAKE)                                                                              \
    extern "C" void* wc_svht_##NAME##_create(uint64_t min_capacity, K seed, int no_init, uint32_t cg) {            \
        try { return MAKE<K, V>(min_capacity, seed, no_init != 0, cg); } catch (...) { return nullptr; }           \
    }                                                                                                              \
    extern "C" void wc_svht_##NAME##_destroy(void* h) { delete static_cast<wc_capi::SvhtApi<K, V>*>(h); }                  \
    extern "C" int wc_svht_##NAME##_init(void* h, void* stream) {                                                  \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->init(K(0), false, wc_capi::as_stream(stream));                                     \
        return wc_capi::last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_init_seed(void* h, K seed, void* stream) {                                     \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->init(seed, true, wc_capi::as_stream(stream));                                      \
        return wc_capi::last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_insert(void* h, const K* keys, const V* values, uint64_t n, void* stream,      \
                                           uint64_t probing_length, uint32_t* status_out) {                        \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        return wc_capi::guarded([&] {                                                                                       \
            static_cast<wc_capi::SvhtApi<K, V>*>(h)->insert(keys, values, n, wc_capi::as_stream(stream), probing_length,            \
                                                   reinterpret_cast<warpcore::Status*>(status_out));                         \
            return wc_capi::last_error();                                                                                   \
        });                                                                                                        \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_insert_bool(void* h, const K* keys, const V* values, uint64_t n,              \
                                                void* stream, uint64_t probing_length, bool* flags_out) {          \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->insert_bool(keys, values, n, wc_capi::as_stream(stream), probing_length,           \
                                                    flags_out);                                                    \
        return wc_capi::last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_retrieve(void* h, const K* keys, uint64_t n, V* values_out, void* stream,      \
                                             uint64_t probing_length, uint32_t* status_out) {                      \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->retrieve(keys, n, values_out, wc_capi::as_stream(stream), probing_length,          \
                                                 reinterpret_cast<warpcore::Status*>(status_out));                           \
        return wc_capi::last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_retrieve_bool(void* h, const K* keys, uint64_t n, V* values_out,               \
                                                  void* stream, uint64_t probing_length, bool* flags_out) {        \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->retrieve_bool(keys, n, values_out, wc_capi::as_stream(stream), probing_length,     \
                                                      flags_out);                                                  \
        return wc_capi::last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_erase(void* h, const K* keys, uint64_t n, void* stream,                        \
                                          uint64_t probing_length, uint32_t* status_out) {                         \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->erase(keys, n, wc_capi::as_stream(stream), probing_length,                         \
                                              reinterpret_cast<warpcore::Status*>(status_out));                              \
        return wc_capi::last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_retrieve_all(void* h, K* keys_out, V* values_out, uint64_t* num_out,           \
                                                 void* stream) {                                                   \
        if (!h || !num_out) return wc_capi::kBadHandle;                                                            \
        index_t n = 0;                                                                                             \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->retrieve_all(keys_out, values_out, n, wc_capi::as_stream(stream));                 \
        *num_out = n;                                                                                              \
        return wc_capi::last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_for_each_count(void* h, uint64_t* counter_d, void* stream) {                  \
        if (!h || !counter_d) return wc_capi::kBadHandle;                                                          \
        static_cast<wc_capi::SvhtApi<K, V>*>(h)->count_valid(counter_d, wc_capi::as_stream(stream));              \
        return wc_capi::last_error();                                                                              \
    }                                                                                                              \
    extern "C" uint64_t wc_svht_##NAME##_size(void* h, void* stream) {                                             \
        return h ? static_cast<wc_capi::SvhtApi<K, V>*>(h)->size(wc_capi::as_stream(stream)) : 0;                                    \
    }                                                                                                              \
    extern "C" float wc_svht_##NAME##_load_factor(void* h, void* stream) {                                         \
        return h ? static_cast<wc_capi::SvhtApi<K, V>*>(h)->load_factor(wc_capi::as_stream(stream)) : 0.0f;                          \
    }                                                                                                              \
    extern "C" uint64_t wc_svht_##NAME##_capacity(void* h) {                                                       \
        return h ? static_cast<wc_capi::SvhtApi<K, V>*>(h)->capacity() : 0;                                                 \
    }                                                                                                              \
    extern "C" uint64_t wc_svht_##NAME##_bytes_total(void* h) {                                                    \
        return h ? static_cast<wc_capi::SvhtApi<K, V>*>(h)->bytes_total() : 0;                                              \
    }                                                                                                              \
    extern "C" uint32_t wc_svht_##NAME##_peek_status(void* h, void* stream) {                                      \
        return h ? static_cast<wc_capi::SvhtApi<K, V>*>(h)->peek_status(wc_capi::as_stream(stream)).base()                           \
                 : warpcore::Status::not_initialized().base();                                                               \
    }                                                                                                              \
    extern "C" uint32_t wc_svht_##NAME##_pop_status(void* h, void* stream) {                                       \
        return h ? static_cast<wc_capi::SvhtApi<K, V>*>(h)->pop_status(wc_capi::as_stream(stream)).base()                            \
                 : warpcore::Status::not_initialized().base();                                                               \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_is_initialized(void* h) {                                                      \
        return h ? int(static_cast<wc_capi::SvhtApi<K, V>*>(h)->is_initialized()) : 0;                                      \
    }                                                                                                              \
    extern "C" uint32_t wc_svht_##NAME##_cg_size(void* h) {                                                        \
        return h ? static_cast<wc_capi::SvhtApi<K, V>*>(h)->cg() : 0;                                                       \
    }
