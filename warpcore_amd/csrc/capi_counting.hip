// extern "C" entry points of CountingHashTable<K, uint64_t> (declared in
// include/warpcore_c.h); cooperative group size chosen at creation.
#include "capi_common.hpp"

#include <warpcore/counting_hash_table.hpp>

#include <mutex>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {

template<class K>
struct CountingApi {
    virtual ~CountingApi() = default;
    virtual void init(K seed, bool reseed, hipStream_t s) = 0;
    virtual void insert(const K*, index_t, hipStream_t, index_t, Status*) = 0;
    virtual void retrieve(const K*, index_t, uint64_t*, hipStream_t, index_t, Status*) = 0;
    virtual void erase(const K*, index_t, hipStream_t, index_t, Status*) = 0;
    virtual void retrieve_all(K*, uint64_t*, index_t&, hipStream_t) = 0;
    virtual index_t size(hipStream_t) = 0;
    virtual float load_factor(hipStream_t) = 0;
    virtual index_t capacity() = 0;
    virtual index_t bytes_total() = 0;
    virtual Status peek_status(hipStream_t) = 0;
    virtual Status pop_status(hipStream_t) = 0;
    virtual unsigned cg() = 0;
    // one combining pass in caller-provided buffers (bytes_for(cap) bytes at
    // base), no allocation and no host round trip
    virtual void insert_async(const K*, index_t, hipStream_t, index_t, Status*, void* base, index_t cap) = 0;
};

template<class K, unsigned CG>
struct CountingImpl final : CountingApi<K> {
    using table_t = CountingHashTable<K, uint64_t, defaults::empty_key<K>(), defaults::tombstone_key<K>(),
                                      defaults::probing_scheme_t<K, CG>>;
    table_t t;
    CountingImpl(index_t min_cap, K seed, uint64_t max_count, bool no_init) : t(min_cap, seed, max_count, no_init) {}
    void init(K seed, bool reseed, hipStream_t s) override { if (reseed) t.init(seed, s); else t.init(s); }
    void insert(const K* k, index_t n, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template insert<status_handlers::ReturnStatus>(k, n, s, pl, st);
        else t.insert(k, n, s, pl);
    }
    void retrieve(const K* k, index_t n, uint64_t* out, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template retrieve<status_handlers::ReturnStatus>(k, n, out, s, pl, st);
        else t.retrieve(k, n, out, s, pl);
    }
    void erase(const K* k, index_t n, hipStream_t s, index_t pl, Status* st) override {
        if (st) t.template erase<status_handlers::ReturnStatus>(k, n, s, pl, st);
        else t.erase(k, n, s, pl);
    }
    void retrieve_all(K* k, uint64_t* v, index_t& n, hipStream_t s) override { t.retrieve_all(k, v, n, s); }
    index_t size(hipStream_t s) override { return t.size(s); }
    float load_factor(hipStream_t s) override { return t.load_factor(s); }
    index_t capacity() override { return t.capacity(); }
    index_t bytes_total() override { return t.bytes_total(); }
    Status peek_status(hipStream_t s) override { return t.peek_status(s); }
    Status pop_status(hipStream_t s) override { return t.pop_status(s); }
    unsigned cg() override { return CG; }
    void insert_async(const K* k, index_t n, hipStream_t s, index_t pl, Status* st, void* base, index_t cap) override {
        const auto buf = table_t::CombineBuffers::carve(base, cap);
        if (st) t.template insert_combined_async<status_handlers::ReturnStatus>(k, n, buf, s, pl, st);
        else t.insert_combined_async(k, n, buf, s, pl);
    }
};

template<class K>
CountingApi<K>* make(index_t min_cap, K seed, uint64_t max_count, bool no_init, unsigned cg) {
    switch (cg) {
    case 0:
    case 4: return new CountingImpl<K, 4>(min_cap, seed, max_count, no_init);
    case 1: return new CountingImpl<K, 1>(min_cap, seed, max_count, no_init);
    case 2: return new CountingImpl<K, 2>(min_cap, seed, max_count, no_init);
    case 8: return new CountingImpl<K, 8>(min_cap, seed, max_count, no_init);
    case 16: return new CountingImpl<K, 16>(min_cap, seed, max_count, no_init);
    case 32: return new CountingImpl<K, 32>(min_cap, seed, max_count, no_init);
    case 64: return new CountingImpl<K, 64>(min_cap, seed, max_count, no_init);
    default: return nullptr;
    }
}

}  // namespace

#define WC_COUNTING_API(NAME, K)                                                                                  \
    extern "C" void* wc_counting_##NAME##_create(uint64_t min_capacity, K seed, uint64_t max_count, int no_init,  \
                                                 uint32_t cg_size) {                                              \
        try { return make<K>(min_capacity, seed, max_count, no_init != 0, cg_size); } catch (...) { return nullptr; } \
    }                                                                                                             \
    extern "C" void wc_counting_##NAME##_destroy(void* h) { delete static_cast<CountingApi<K>*>(h); }            \
    extern "C" int wc_counting_##NAME##_init(void* h, void* stream) {                                             \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<CountingApi<K>*>(h)->init(K(0), false, as_stream(stream));                                   \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_counting_##NAME##_init_seed(void* h, K seed, void* stream) {                                \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<CountingApi<K>*>(h)->init(seed, true, as_stream(stream));                                    \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_counting_##NAME##_insert(void* h, const K* keys, uint64_t n, void* stream, uint64_t pl,     \
                                               uint32_t* status_out) {                                            \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<CountingApi<K>*>(h)->insert(keys, n, as_stream(stream), pl, reinterpret_cast<Status*>(status_out)); \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_counting_##NAME##_retrieve(void* h, const K* keys, uint64_t n, uint64_t* counts_out,       \
                                                 void* stream, uint64_t pl, uint32_t* status_out) {               \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<CountingApi<K>*>(h)->retrieve(keys, n, counts_out, as_stream(stream), pl,                    \
                                                  reinterpret_cast<Status*>(status_out));                         \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_counting_##NAME##_erase(void* h, const K* keys, uint64_t n, void* stream, uint64_t pl,      \
                                              uint32_t* status_out) {                                             \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<CountingApi<K>*>(h)->erase(keys, n, as_stream(stream), pl, reinterpret_cast<Status*>(status_out)); \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_counting_##NAME##_retrieve_all(void* h, K* keys_out, uint64_t* counts_out, uint64_t* num_out, \
                                                     void* stream) {                                              \
        if (!h || !num_out) return wc_capi::kBadHandle;                                                           \
        index_t n = 0;                                                                                            \
        static_cast<CountingApi<K>*>(h)->retrieve_all(keys_out, counts_out, n, as_stream(stream));               \
        *num_out = n;                                                                                             \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" uint64_t wc_counting_##NAME##_size(void* h, void* stream) {                                        \
        return h ? static_cast<CountingApi<K>*>(h)->size(as_stream(stream)) : 0;                                  \
    }                                                                                                             \
    extern "C" float wc_counting_##NAME##_load_factor(void* h, void* stream) {                                    \
        return h ? static_cast<CountingApi<K>*>(h)->load_factor(as_stream(stream)) : 0.0f;                        \
    }                                                                                                             \
    extern "C" uint64_t wc_counting_##NAME##_capacity(void* h) {                                                  \
        return h ? static_cast<CountingApi<K>*>(h)->capacity() : 0;                                               \
    }                                                                                                             \
    extern "C" uint64_t wc_counting_##NAME##_bytes_total(void* h) {                                               \
        return h ? static_cast<CountingApi<K>*>(h)->bytes_total() : 0;                                            \
    }                                                                                                             \
    extern "C" uint32_t wc_counting_##NAME##_peek_status(void* h, void* stream) {                                 \
        return h ? static_cast<CountingApi<K>*>(h)->peek_status(as_stream(stream)).base()                         \
                 : Status::not_initialized().base();                                                              \
    }                                                                                                             \
    extern "C" uint32_t wc_counting_##NAME##_pop_status(void* h, void* stream) {                                  \
        return h ? static_cast<CountingApi<K>*>(h)->pop_status(as_stream(stream)).base()                          \
                 : Status::not_initialized().base();                                                              \
    }                                                                                                             \
    extern "C" uint32_t wc_counting_##NAME##_cg_size(void* h) {                                                   \
        return h ? static_cast<CountingApi<K>*>(h)->cg() : 0;                                                     \
    }

WC_COUNTING_API(u32, uint32_t)
WC_COUNTING_API(u64, uint64_t)

// ---------------------------------------------------------------------------
// Insert without allocation (no counterpart in the reference): the combining
// buffers live in `scratch`, a file-scope device array of fixed size that
// every GPU the library runs on holds until the process ends. Its device
// address is taken once per device with hipGetSymbolAddress (host code must
// never pass the array's name to a kernel: on the host it is a stand-in, not
// a device address). Calls on one GPU share the array, so calls that overlap
// on one GPU (different streams, or other host threads) race; calls issued
// one after another on one stream do not. Everything the passes need zeroed
// (the entry count) is cleared by a fill kernel at the start of each call.

namespace {

constexpr index_t kNoAllocOccurrences = index_t(1) << 20;
using noalloc_buffers = CountingHashTable<std::uint64_t, std::uint64_t>::CombineBuffers;
constexpr index_t kNoAllocBytes = noalloc_buffers::bytes_for(kNoAllocOccurrences);
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Device address of `scratch` on the current device (nullptr on failure).
void* scratch_address() noexcept {
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    if (addr[dev] == nullptr) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return nullptr;
        addr[dev] = p;
    }
    return addr[dev];
}

}  // namespace

extern "C" int wc_counting_u64_insert_noalloc(void* h, const uint64_t* keys, uint64_t n, void* stream, uint64_t pl,
                                              uint32_t* status_out) {
    if (!h) return wc_capi::kBadHandle;
    if (n > uint64_t(kNoAllocOccurrences)) return wc_capi::kTooLarge;
    auto* t = static_cast<CountingApi<uint64_t>*>(h);
    if (n < 4096 || detail::counting_combine() != 1) {   // as wc_counting_u64_insert: no combining buffers
        t->insert(keys, n, as_stream(stream), pl, reinterpret_cast<Status*>(status_out));
        return last_error();
    }
    void* base = scratch_address();
    if (base == nullptr) return wc_capi::kNoScratch;
    t->insert_async(keys, n, as_stream(stream), pl, reinterpret_cast<Status*>(status_out), base, kNoAllocOccurrences);
    return last_error();
}

extern "C" uint64_t wc_counting_u64_insert_noalloc_max(void) { return kNoAllocOccurrences; }
