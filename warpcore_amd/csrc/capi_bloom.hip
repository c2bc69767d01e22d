// extern "C" entry points of BloomFilter<K, MurmurHash<K>, Slot, CG>
// (declared in include/warpcore_c.h).
#include "capi_common.hpp"

#include <warpcore/bloom_filter.hpp>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {

template<class K>
struct BloomApi {
    virtual ~BloomApi() = default;
    virtual bool ok() = 0;
    virtual void init(K seed, bool reseed, hipStream_t s) = 0;
    virtual void insert(const K*, index_t, hipStream_t) = 0;
    virtual void retrieve(const K*, index_t, bool*, hipStream_t) = 0;
    virtual bool insert_and_query(const K*, index_t, bool*, hipStream_t) = 0;
    virtual index_t num_bits() = 0;
    virtual index_t num_slots() = 0;
    virtual index_t num_blocks() = 0;
    virtual index_t k() = 0;
    virtual K seed() = 0;
    virtual unsigned slot_bits() = 0;
    virtual unsigned cg() = 0;
    virtual index_t bytes_total() = 0;
    virtual double fpr(index_t n) = 0;
    virtual int copy_words(void* out, hipStream_t s) = 0;
};

template<class K, class Slot, unsigned CG>
struct BloomImpl final : BloomApi<K> {
    using filter_t = BloomFilter<K, defaults::hasher_t<K>, Slot, CG>;
    filter_t f;
    BloomImpl(index_t num_bits, index_t k, K seed) : f(num_bits, k, seed) {}
    bool ok() override { return f.data() != nullptr; }
    void init(K seed, bool reseed, hipStream_t s) override { if (reseed) f.init(seed, s); else f.init(s); }
    void insert(const K* k, index_t n, hipStream_t s) override { f.insert(k, n, s); }
    void retrieve(const K* k, index_t n, bool* o, hipStream_t s) override { f.retrieve(k, n, o, s); }
    bool insert_and_query(const K* k, index_t n, bool* o, hipStream_t s) override {
        if constexpr (CG == 1) {
            if (n == 0) return true;
            kernels::bloom_insert_and_query<typename filter_t::view_type>
                <<<detail::grid_for(n, 1), block_size(), 0, s>>>(k, n, o, f.view());
            return true;
        } else {
            return false;
        }
    }
    index_t num_bits() override { return f.num_bits(); }
    index_t num_slots() override { return f.num_slots(); }
    index_t num_blocks() override { return f.num_blocks(); }
    index_t k() override { return f.k(); }
    K seed() override { return f.seed(); }
    unsigned slot_bits() override { return unsigned(filter_t::slot_bits()); }
    unsigned cg() override { return CG; }
    index_t bytes_total() override { return f.bytes_total(); }
    double fpr(index_t n) override { return f.fpr(n); }
    int copy_words(void* out, hipStream_t s) override {
        if (f.bytes_total() == 0) return 0;
        return int(hipMemcpyAsync(out, f.data(), f.bytes_total(), hipMemcpyDeviceToDevice, s));
    }
};

template<class K, class Slot>
BloomApi<K>* make_slot(index_t num_bits, index_t k, K seed, unsigned cg) {
    switch (cg) {
    case 0:
    case 1: return new BloomImpl<K, Slot, 1>(num_bits, k, seed);
    case 2: return new BloomImpl<K, Slot, 2>(num_bits, k, seed);
    case 4: return new BloomImpl<K, Slot, 4>(num_bits, k, seed);
    case 8: return new BloomImpl<K, Slot, 8>(num_bits, k, seed);
    case 16: return new BloomImpl<K, Slot, 16>(num_bits, k, seed);
    case 32: return new BloomImpl<K, Slot, 32>(num_bits, k, seed);
    case 64: return new BloomImpl<K, Slot, 64>(num_bits, k, seed);
    default: return nullptr;
    }
}

template<class K>
BloomApi<K>* make(index_t num_bits, index_t k, K seed, unsigned slot_bits, unsigned cg) {
    if (num_bits == 0 || k == 0) return nullptr;
    BloomApi<K>* f = nullptr;
    if (slot_bits == 0 || slot_bits == 64) f = make_slot<K, std::uint64_t>(num_bits, k, seed, cg);
    else if (slot_bits == 32) f = make_slot<K, std::uint32_t>(num_bits, k, seed, cg);
    if (f && !f->ok()) {
        delete f;
        f = nullptr;
    }
    return f;
}

}  // namespace

#define WC_BLOOM_API(NAME, K)                                                                                   \
    extern "C" void* wc_bloom_##NAME##_create(uint64_t num_bits, uint64_t k, K seed, uint32_t slot_bits,        \
                                              uint32_t cg_size) {                                               \
        try { return make<K>(num_bits, k, seed, slot_bits, cg_size); } catch (...) { return nullptr; }          \
    }                                                                                                           \
    extern "C" void wc_bloom_##NAME##_destroy(void* h) { delete static_cast<BloomApi<K>*>(h); }                \
    extern "C" int wc_bloom_##NAME##_init(void* h, void* stream) {                                              \
        if (!h) return wc_capi::kBadHandle;                                                                     \
        static_cast<BloomApi<K>*>(h)->init(K(0), false, as_stream(stream));                                    \
        return last_error();                                                                                    \
    }                                                                                                           \
    extern "C" int wc_bloom_##NAME##_init_seed(void* h, K seed, void* stream) {                                 \
        if (!h) return wc_capi::kBadHandle;                                                                     \
        static_cast<BloomApi<K>*>(h)->init(seed, true, as_stream(stream));                                     \
        return last_error();                                                                                    \
    }                                                                                                           \
    extern "C" int wc_bloom_##NAME##_insert(void* h, const K* keys, uint64_t n, void* stream) {                 \
        if (!h) return wc_capi::kBadHandle;                                                                     \
        static_cast<BloomApi<K>*>(h)->insert(keys, n, as_stream(stream));                                      \
        return last_error();                                                                                    \
    }                                                                                                           \
    extern "C" int wc_bloom_##NAME##_retrieve(void* h, const K* keys, uint64_t n, bool* flags_out,              \
                                              void* stream) {                                                   \
        if (!h) return wc_capi::kBadHandle;                                                                     \
        static_cast<BloomApi<K>*>(h)->retrieve(keys, n, flags_out, as_stream(stream));                         \
        return last_error();                                                                                    \
    }                                                                                                           \
    extern "C" int wc_bloom_##NAME##_insert_and_query(void* h, const K* keys, uint64_t n, bool* flags_out,      \
                                                      void* stream) {                                           \
        if (!h) return wc_capi::kBadHandle;                                                                     \
        if (!static_cast<BloomApi<K>*>(h)->insert_and_query(keys, n, flags_out, as_stream(stream)))            \
            return wc_capi::kBadArgument;                                                                       \
        return last_error();                                                                                    \
    }                                                                                                           \
    extern "C" int wc_bloom_##NAME##_copy_words(void* h, void* words_out, void* stream) {                       \
        if (!h) return wc_capi::kBadHandle;                                                                     \
        return static_cast<BloomApi<K>*>(h)->copy_words(words_out, as_stream(stream));                         \
    }                                                                                                           \
    extern "C" uint64_t wc_bloom_##NAME##_num_bits(void* h) { return h ? static_cast<BloomApi<K>*>(h)->num_bits() : 0; } \
    extern "C" uint64_t wc_bloom_##NAME##_num_slots(void* h) {                                                  \
        return h ? static_cast<BloomApi<K>*>(h)->num_slots() : 0;                                               \
    }                                                                                                           \
    extern "C" uint64_t wc_bloom_##NAME##_num_blocks(void* h) {                                                 \
        return h ? static_cast<BloomApi<K>*>(h)->num_blocks() : 0;                                              \
    }                                                                                                           \
    extern "C" uint64_t wc_bloom_##NAME##_k(void* h) { return h ? static_cast<BloomApi<K>*>(h)->k() : 0; }      \
    extern "C" K wc_bloom_##NAME##_seed(void* h) { return h ? static_cast<BloomApi<K>*>(h)->seed() : K(0); }    \
    extern "C" uint32_t wc_bloom_##NAME##_slot_bits(void* h) {                                                  \
        return h ? static_cast<BloomApi<K>*>(h)->slot_bits() : 0;                                               \
    }                                                                                                           \
    extern "C" uint32_t wc_bloom_##NAME##_cg_size(void* h) { return h ? static_cast<BloomApi<K>*>(h)->cg() : 0; } \
    extern "C" uint64_t wc_bloom_##NAME##_bytes_total(void* h) {                                                \
        return h ? static_cast<BloomApi<K>*>(h)->bytes_total() : 0;                                             \
    }                                                                                                           \
    extern "C" double wc_bloom_##NAME##_fpr(void* h, uint64_t n) {                                              \
        return h ? static_cast<BloomApi<K>*>(h)->fpr(n) : 1.0;                                                  \
    }

WC_BLOOM_API(u32, uint32_t)
WC_BLOOM_API(u64, uint64_t)
