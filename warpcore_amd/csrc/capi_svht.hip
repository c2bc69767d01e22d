// extern "C" entry points of SingleValueHashTable (declared in
// include/warpcore_c.h). One handle type per (key, value) pair; the
// cooperative group size is chosen at creation among the compiled sizes.
#include "capi_common.hpp"

using namespace warpcore;

namespace {

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
    virtual index_t size(hipStream_t) = 0;
    virtual float load_factor(hipStream_t) = 0;
    virtual index_t capacity() = 0;
    virtual index_t bytes_total() = 0;
    virtual Status peek_status(hipStream_t) = 0;
    virtual Status pop_status(hipStream_t) = 0;
    virtual bool is_initialized() = 0;
    virtual unsigned cg() = 0;
};

template<class K, class V, unsigned CG>
struct SvhtImpl final : SvhtApi<K, V> {
    using table_t = SingleValueHashTable<K, V, defaults::empty_key<K>(), defaults::tombstone_key<K>(),
                                         defaults::probing_scheme_t<K, CG>>;
    table_t t;
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
    index_t size(hipStream_t s) override { return t.size(s); }
    float load_factor(hipStream_t s) override { return t.load_factor(s); }
    index_t capacity() override { return t.capacity(); }
    index_t bytes_total() override { return t.bytes_total(); }
    Status peek_status(hipStream_t s) override { return t.peek_status(s); }
    Status pop_status(hipStream_t s) override { return t.pop_status(s); }
    bool is_initialized() override { return t.is_initialized(); }
    unsigned cg() override { return CG; }
};

template<class K, class V>
SvhtApi<K, V>* make_svht(index_t min_cap, K seed, bool no_init, unsigned cg) {
    switch (cg) {
    case 0:
    case 8: return new SvhtImpl<K, V, 8>(min_cap, seed, no_init);
    case 1: return new SvhtImpl<K, V, 1>(min_cap, seed, no_init);
    case 2: return new SvhtImpl<K, V, 2>(min_cap, seed, no_init);
    case 4: return new SvhtImpl<K, V, 4>(min_cap, seed, no_init);
    case 16: return new SvhtImpl<K, V, 16>(min_cap, seed, no_init);
    case 32: return new SvhtImpl<K, V, 32>(min_cap, seed, no_init);
    case 64: return new SvhtImpl<K, V, 64>(min_cap, seed, no_init);
    default: return nullptr;
    }
}

template<class K, class V>
SvhtApi<K, V>* make_svht_default_cg(index_t min_cap, K seed, bool no_init, unsigned cg) {
    if (cg != 0 && cg != 8) return nullptr;
    return new SvhtImpl<K, V, 8>(min_cap, seed, no_init);
}

}  // namespace

using wc_capi::as_stream;
using wc_capi::guarded;
using wc_capi::last_error;

#define WC_SVHT_API(NAME, K, V, MAKE)                                                                              \
    extern "C" void* wc_svht_##NAME##_create(uint64_t min_capacity, K seed, int no_init, uint32_t cg) {            \
        try { return MAKE<K, V>(min_capacity, seed, no_init != 0, cg); } catch (...) { return nullptr; }           \
    }                                                                                                              \
    extern "C" void wc_svht_##NAME##_destroy(void* h) { delete static_cast<SvhtApi<K, V>*>(h); }                  \
    extern "C" int wc_svht_##NAME##_init(void* h, void* stream) {                                                  \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<SvhtApi<K, V>*>(h)->init(K(0), false, as_stream(stream));                                     \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_init_seed(void* h, K seed, void* stream) {                                     \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<SvhtApi<K, V>*>(h)->init(seed, true, as_stream(stream));                                      \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_insert(void* h, const K* keys, const V* values, uint64_t n, void* stream,      \
                                           uint64_t probing_length, uint32_t* status_out) {                        \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        return guarded([&] {                                                                                       \
            static_cast<SvhtApi<K, V>*>(h)->insert(keys, values, n, as_stream(stream), probing_length,            \
                                                   reinterpret_cast<Status*>(status_out));                         \
            return last_error();                                                                                   \
        });                                                                                                        \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_insert_bool(void* h, const K* keys, const V* values, uint64_t n,              \
                                                void* stream, uint64_t probing_length, bool* flags_out) {          \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<SvhtApi<K, V>*>(h)->insert_bool(keys, values, n, as_stream(stream), probing_length,           \
                                                    flags_out);                                                    \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_retrieve(void* h, const K* keys, uint64_t n, V* values_out, void* stream,      \
                                             uint64_t probing_length, uint32_t* status_out) {                      \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<SvhtApi<K, V>*>(h)->retrieve(keys, n, values_out, as_stream(stream), probing_length,          \
                                                 reinterpret_cast<Status*>(status_out));                           \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_retrieve_bool(void* h, const K* keys, uint64_t n, V* values_out,               \
                                                  void* stream, uint64_t probing_length, bool* flags_out) {        \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<SvhtApi<K, V>*>(h)->retrieve_bool(keys, n, values_out, as_stream(stream), probing_length,     \
                                                      flags_out);                                                  \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_erase(void* h, const K* keys, uint64_t n, void* stream,                        \
                                          uint64_t probing_length, uint32_t* status_out) {                         \
        if (!h) return wc_capi::kBadHandle;                                                                        \
        static_cast<SvhtApi<K, V>*>(h)->erase(keys, n, as_stream(stream), probing_length,                         \
                                              reinterpret_cast<Status*>(status_out));                              \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_retrieve_all(void* h, K* keys_out, V* values_out, uint64_t* num_out,           \
                                                 void* stream) {                                                   \
        if (!h || !num_out) return wc_capi::kBadHandle;                                                            \
        index_t n = 0;                                                                                             \
        static_cast<SvhtApi<K, V>*>(h)->retrieve_all(keys_out, values_out, n, as_stream(stream));                 \
        *num_out = n;                                                                                              \
        return last_error();                                                                                       \
    }                                                                                                              \
    extern "C" uint64_t wc_svht_##NAME##_size(void* h, void* stream) {                                             \
        return h ? static_cast<SvhtApi<K, V>*>(h)->size(as_stream(stream)) : 0;                                    \
    }                                                                                                              \
    extern "C" float wc_svht_##NAME##_load_factor(void* h, void* stream) {                                         \
        return h ? static_cast<SvhtApi<K, V>*>(h)->load_factor(as_stream(stream)) : 0.0f;                          \
    }                                                                                                              \
    extern "C" uint64_t wc_svht_##NAME##_capacity(void* h) {                                                       \
        return h ? static_cast<SvhtApi<K, V>*>(h)->capacity() : 0;                                                 \
    }                                                                                                              \
    extern "C" uint64_t wc_svht_##NAME##_bytes_total(void* h) {                                                    \
        return h ? static_cast<SvhtApi<K, V>*>(h)->bytes_total() : 0;                                              \
    }                                                                                                              \
    extern "C" uint32_t wc_svht_##NAME##_peek_status(void* h, void* stream) {                                      \
        return h ? static_cast<SvhtApi<K, V>*>(h)->peek_status(as_stream(stream)).base()                           \
                 : Status::not_initialized().base();                                                               \
    }                                                                                                              \
    extern "C" uint32_t wc_svht_##NAME##_pop_status(void* h, void* stream) {                                       \
        return h ? static_cast<SvhtApi<K, V>*>(h)->pop_status(as_stream(stream)).base()                            \
                 : Status::not_initialized().base();                                                               \
    }                                                                                                              \
    extern "C" int wc_svht_##NAME##_is_initialized(void* h) {                                                      \
        return h ? int(static_cast<SvhtApi<K, V>*>(h)->is_initialized()) : 0;                                      \
    }                                                                                                              \
    extern "C" uint32_t wc_svht_##NAME##_cg_size(void* h) {                                                        \
        return h ? static_cast<SvhtApi<K, V>*>(h)->cg() : 0;                                                       \
    }

WC_SVHT_API(u32u32, uint32_t, uint32_t, make_svht)
WC_SVHT_API(u64u64, uint64_t, uint64_t, make_svht)
WC_SVHT_API(u32u64, uint32_t, uint64_t, make_svht_default_cg)
WC_SVHT_API(u64u32, uint64_t, uint32_t, make_svht_default_cg)
