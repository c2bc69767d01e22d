// extern "C" entry points of MultiValueHashTable<K, V> (declared in
// include/warpcore_c.h).
#include "capi_common.hpp"

#include <warpcore/multi_value_hash_table.hpp>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {

template<class K, class V>
struct MvApi {
    virtual ~MvApi() = default;
    virtual void init(hipStream_t s) = 0;
    virtual void insert(const K*, const V*, index_t, hipStream_t, index_t, Status*) = 0;
    virtual void retrieve(const K*, index_t, index_t*, index_t*, V*, index_t&, hipStream_t, index_t, Status*) = 0;
    virtual void num_values(const K*, index_t, index_t&, index_t*, hipStream_t, index_t, Status*) = 0;
    virtual void retrieve_all_keys(K*, index_t&, hipStream_t) = 0;
    virtual void retrieve_all(K*, index_t&, index_t*, index_t*, V*, index_t&, hipStream_t) = 0;
    virtual index_t num_keys(hipStream_t) = 0;
    virtual index_t size(hipStream_t) = 0;
    virtual float load_factor(hipStream_t) = 0;
    virtual float storage_density(hipStream_t) = 0;
    virtual index_t capacity() = 0;
    virtual index_t bytes_total() = 0;
    virtual index_t max_values_per_key() = 0;
    virtual bool is_initialized() = 0;
    virtual Status peek_status(hipStream_t) = 0;
    virtual Status pop_status(hipStream_t) = 0;
    virtual unsigned cg() = 0;
};

template<class K, class V, unsigned CG>
struct MvImpl final : MvApi<K, V> {
    using table_t = MultiValueHashTable<K, V, defaults::empty_key<K>(), defaults::tombstone_key<K>(),
                                        defaults::probing_scheme_t<K, CG>>;
    table_t t;
    MvImpl(index_t min_cap, K seed, index_t max_values, bool no_init) : t(min_cap, seed, max_values, no_init) {}
    void init(hipStream_t s) override { t.init(s); }
    void insert(const K* k, const V* v, index_t n, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template insert<status_handlers::ReturnStatus>(k, v, n, s, pl, st);
        else t.insert(k, v, n, s, pl);
    }
    void retrieve(const K* k, index_t n, index_t* b, index_t* e, V* v, index_t& num_out, hipStream_t s, index_t pl,
                  Status* st) override {
        if (st) t.template retrieve<status_handlers::ReturnStatus>(k, n, b, e, v, num_out, s, pl, st);
        else t.retrieve(k, n, b, e, v, num_out, s, pl);
    }
    void num_values(const K* k, index_t n, index_t& num_out, index_t* per_key, hipStream_t s, index_t pl,
                    Status* st) override {
        if (st) t.template num_values<status_handlers::ReturnStatus>(k, n, num_out, per_key, s, pl, st);
        else t.num_values(k, n, num_out, per_key, s, pl);
    }
    void retrieve_all_keys(K* k, index_t& n, hipStream_t s) override { t.retrieve_all_keys(k, n, s); }
    void retrieve_all(K* k, index_t& nk, index_t* b, index_t* e, V* v, index_t& nv, hipStream_t s) override {
        t.retrieve_all(k, nk, b, e, v, nv, s);
    }
    index_t num_keys(hipStream_t s) override { return t.num_keys(s); }
    index_t size(hipStream_t s) override { return t.size(s); }
    float load_factor(hipStream_t s) override { return t.load_factor(s); }
    float storage_density(hipStream_t s) override { return t.storage_density(s); }
    index_t capacity() override { return t.capacity(); }
    index_t bytes_total() override { return t.bytes_total(); }
    index_t max_values_per_key() override { return t.max_values_per_key(); }
    bool is_initialized() override { return t.is_initialized(); }
    Status peek_status(hipStream_t s) override { return t.peek_status(s); }
    Status pop_status(hipStream_t s) override { return t.pop_status(s); }
    unsigned cg() override { return CG; }
};

template<class K, class V>
MvApi<K, V>* make(index_t min_cap, K seed, index_t max_values, bool no_init, unsigned cg) {
    switch (cg) {
    case 0:
    case 8: return new MvImpl<K, V, 8>(min_cap, seed, max_values, no_init);
    case 1: return new MvImpl<K, V, 1>(min_cap, seed, max_values, no_init);
    case 2: return new MvImpl<K, V, 2>(min_cap, seed, max_values, no_init);
    case 4: return new MvImpl<K, V, 4>(min_cap, seed, max_values, no_init);
    case 16: return new MvImpl<K, V, 16>(min_cap, seed, max_values, no_init);
    case 32: return new MvImpl<K, V, 32>(min_cap, seed, max_values, no_init);
    case 64: return new MvImpl<K, V, 64>(min_cap, seed, max_values, no_init);
    default: return nullptr;
    }
}

}  // namespace

#define WC_MVHT_API(NAME, K, V)                                                                                   \
    using Mv_##NAME = MvApi<K, V>;                                                                                \
    extern "C" void* wc_mvht_##NAME##_create(uint64_t min_capacity, K seed, uint64_t max_values_per_key,          \
                                             int no_init, uint32_t cg_size) {                                     \
        try { return make<K, V>(min_capacity, seed, max_values_per_key, no_init != 0, cg_size); }                 \
        catch (...) { return nullptr; }                                                                           \
    }                                                                                                             \
    extern "C" void wc_mvht_##NAME##_destroy(void* h) { delete static_cast<Mv_##NAME*>(h); }                     \
    extern "C" int wc_mvht_##NAME##_init(void* h, void* stream) {                                                 \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<Mv_##NAME*>(h)->init(as_stream(stream));                                                      \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_mvht_##NAME##_insert(void* h, const K* keys, const V* values, uint64_t n, void* stream,     \
                                           uint64_t pl, uint32_t* status_out) {                                   \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<Mv_##NAME*>(h)->insert(keys, values, n, as_stream(stream), pl,                                \
                                           reinterpret_cast<Status*>(status_out));                                \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_mvht_##NAME##_retrieve(void* h, const K* keys, uint64_t n, uint64_t* begin_offsets_out,    \
                                             uint64_t* end_offsets_out, V* values_out, uint64_t* num_out,         \
                                             void* stream, uint64_t pl, uint32_t* status_out) {                   \
        if (!h || !num_out) return wc_capi::kBadHandle;                                                           \
        if (n && !end_offsets_out) return wc_capi::kBadArgument;                                                  \
        if (n && values_out && !begin_offsets_out) return wc_capi::kBadArgument;                                  \
        index_t num = 0;                                                                                          \
        static_cast<Mv_##NAME*>(h)->retrieve(keys, n, begin_offsets_out, end_offsets_out, values_out, num,        \
                                             as_stream(stream), pl, reinterpret_cast<Status*>(status_out));       \
        *num_out = num;                                                                                           \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_mvht_##NAME##_num_values(void* h, const K* keys, uint64_t n, uint64_t* num_out,             \
                                               uint64_t* num_per_key_out, void* stream, uint64_t pl,              \
                                               uint32_t* status_out) {                                            \
        if (!h || !num_out) return wc_capi::kBadHandle;                                                           \
        index_t num = 0;                                                                                          \
        static_cast<Mv_##NAME*>(h)->num_values(keys, n, num, num_per_key_out, as_stream(stream), pl,              \
                                               reinterpret_cast<Status*>(status_out));                            \
        *num_out = num;                                                                                           \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_mvht_##NAME##_retrieve_all_keys(void* h, K* keys_out, uint64_t* num_out, void* stream) {    \
        if (!h || !num_out) return wc_capi::kBadHandle;                                                           \
        index_t num = 0;                                                                                          \
        static_cast<Mv_##NAME*>(h)->retrieve_all_keys(keys_out, num, as_stream(stream));                          \
        *num_out = num;                                                                                           \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_mvht_##NAME##_retrieve_all(void* h, K* keys_out, uint64_t* num_keys_out,                    \
                                                 uint64_t* begin_offsets_out, uint64_t* end_offsets_out,          \
                                                 V* values_out, uint64_t* num_values_out, void* stream) {         \
        if (!h || !num_keys_out || !num_values_out) return wc_capi::kBadHandle;                                   \
        index_t nk = 0, nv = 0;                                                                                   \
        static_cast<Mv_##NAME*>(h)->retrieve_all(keys_out, nk, begin_offsets_out, end_offsets_out, values_out,    \
                                                 nv, as_stream(stream));                                          \
        *num_keys_out = nk;                                                                                       \
        *num_values_out = nv;                                                                                     \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" uint64_t wc_mvht_##NAME##_num_keys(void* h, void* stream) {                                        \
        return h ? static_cast<Mv_##NAME*>(h)->num_keys(as_stream(stream)) : 0;                                   \
    }                                                                                                             \
    extern "C" uint64_t wc_mvht_##NAME##_size(void* h, void* stream) {                                            \
        return h ? static_cast<Mv_##NAME*>(h)->size(as_stream(stream)) : 0;                                       \
    }                                                                                                             \
    extern "C" float wc_mvht_##NAME##_load_factor(void* h, void* stream) {                                        \
        return h ? static_cast<Mv_##NAME*>(h)->load_factor(as_stream(stream)) : 0.0f;                             \
    }                                                                                                             \
    extern "C" float wc_mvht_##NAME##_storage_density(void* h, void* stream) {                                    \
        return h ? static_cast<Mv_##NAME*>(h)->storage_density(as_stream(stream)) : 0.0f;                         \
    }                                                                                                             \
    extern "C" uint64_t wc_mvht_##NAME##_capacity(void* h) { return h ? static_cast<Mv_##NAME*>(h)->capacity() : 0; } \
    extern "C" uint64_t wc_mvht_##NAME##_bytes_total(void* h) {                                                   \
        return h ? static_cast<Mv_##NAME*>(h)->bytes_total() : 0;                                                 \
    }                                                                                                             \
    extern "C" uint64_t wc_mvht_##NAME##_max_values_per_key(void* h) {                                            \
        return h ? static_cast<Mv_##NAME*>(h)->max_values_per_key() : 0;                                          \
    }                                                                                                             \
    extern "C" int wc_mvht_##NAME##_is_initialized(void* h) {                                                     \
        return h ? int(static_cast<Mv_##NAME*>(h)->is_initialized()) : 0;                                         \
    }                                                                                                             \
    extern "C" uint32_t wc_mvht_##NAME##_peek_status(void* h, void* stream) {                                     \
        return h ? static_cast<Mv_##NAME*>(h)->peek_status(as_stream(stream)).base()                              \
                 : Status::not_initialized().base();                                                              \
    }                                                                                                             \
    extern "C" uint32_t wc_mvht_##NAME##_pop_status(void* h, void* stream) {                                      \
        return h ? static_cast<Mv_##NAME*>(h)->pop_status(as_stream(stream)).base()                               \
                 : Status::not_initialized().base();                                                              \
    }                                                                                                             \
    extern "C" uint32_t wc_mvht_##NAME##_cg_size(void* h) { return h ? static_cast<Mv_##NAME*>(h)->cg() : 0; }

WC_MVHT_API(u32u32, uint32_t, uint32_t)
WC_MVHT_API(u64u64, uint64_t, uint64_t)
