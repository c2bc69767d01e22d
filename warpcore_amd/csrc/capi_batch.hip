// extern "C" entry points of TableBatch (declared in include/warpcore_c.h),
// including the launch-parameter autotuner.
#include "capi_common.hpp"

#include <warpcore/table_batch.hpp>

#include <vector>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {

template<class K, class V>
struct BatchApi {
    detail::BatchLaunch launch{8, 64, 8};
    virtual ~BatchApi() = default;
    virtual bool ok() = 0;
    virtual void init(hipStream_t) = 0;
    virtual int insert(const K*, const V*, const index_t*, index_t, hipStream_t, index_t, Status*,
                       const detail::BatchLaunch&) = 0;
    virtual int retrieve(const K*, const index_t*, index_t, V*, hipStream_t, index_t, Status*,
                         const detail::BatchLaunch&) = 0;
    virtual index_t num_tables() = 0;
    virtual index_t capacity(index_t) = 0;
    virtual index_t total_slots() = 0;
    virtual index_t bytes_total() = 0;
    virtual const Status* status_words() = 0;
    virtual unsigned cg() = 0;
};

template<class K, class V, unsigned CG>
struct BatchImpl final : BatchApi<K, V> {
    TableBatch<K, V, CG> b;
    BatchImpl(const index_t* caps, index_t n, K seed) : b(caps, n, seed) {}
    bool ok() override { return b.ok(); }
    void init(hipStream_t s) override { b.init(s); }
    int insert(const K* k, const V* v, const index_t* off, index_t n, hipStream_t s, index_t pl, Status* st,
               const detail::BatchLaunch& l) override {
        if (st) b.template insert<status_handlers::ReturnStatus>(k, v, off, n, s, pl, st, &l);
        else b.insert(k, v, off, n, s, pl, nullptr, &l);
        return b.last_error();
    }
    int retrieve(const K* k, const index_t* off, index_t n, V* out, hipStream_t s, index_t pl, Status* st,
                 const detail::BatchLaunch& l) override {
        if (st) b.template retrieve<status_handlers::ReturnStatus>(k, off, n, out, s, pl, st, &l);
        else b.retrieve(k, off, n, out, s, pl, nullptr, &l);
        return b.last_error();
    }
    index_t num_tables() override { return b.num_tables(); }
    index_t capacity(index_t t) override { return t < b.num_tables() ? b.capacity(t) : 0; }
    index_t total_slots() override { return b.total_slots(); }
    index_t bytes_total() override { return b.bytes_total(); }
    const Status* status_words() override { return b.status_words(); }
    unsigned cg() override { return CG; }
};

template<class K, class V>
BatchApi<K, V>* make(const index_t* caps, index_t n, K seed, unsigned cg) {
    BatchApi<K, V>* b = nullptr;
    switch (cg) {
    case 4: b = new BatchImpl<K, V, 4>(caps, n, seed); break;
    case 8: b = new BatchImpl<K, V, 8>(caps, n, seed); break;
    case 16: b = new BatchImpl<K, V, 16>(caps, n, seed); break;
    default: return nullptr;
    }
    if (!b->ok()) {
        delete b;
        return nullptr;
    }
    return b;
}

// Times every candidate launch on the given workload (init, insert, then
// retrieve of the queries) and keeps the fastest in b->launch.
template<class K, class V>
int tune(BatchApi<K, V>* b, const K* keys, const V* values, const index_t* offsets, index_t n, const K* queries,
         const index_t* qoffsets, index_t nq, V* out, hipStream_t s, std::uint64_t* launch_out, double* ms_out) {
    const int kpb[] = {4, 8};
    const index_t items[] = {32, 64, 128, 256};
    const int bpc[] = {4, 8, 16};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return int(hipErrorOutOfMemory);
    float best = -1.0f;
    detail::BatchLaunch best_launch = b->launch;
    for (int k : kpb)
        for (index_t it : items)
            for (int bc : bpc) {
                const detail::BatchLaunch l{k, it, bc};
                float total = 0.0f;
                for (int rep = 0; rep < 2; ++rep) {   // first repetition warms up
                    b->init(s);
                    (void)hipEventRecord(e0, s);
                    int rc = b->insert(keys, values, offsets, n, s, defaults::probing_length(), nullptr, l);
                    if (rc == 0 && nq) rc = b->retrieve(queries, qoffsets, nq, out, s, defaults::probing_length(),
                                                        nullptr, l);
                    (void)hipEventRecord(e1, s);
                    if (rc != 0) {
                        (void)hipEventDestroy(e0);
                        (void)hipEventDestroy(e1);
                        return rc;
                    }
                    (void)hipEventSynchronize(e1);
                    float ms = 0.0f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (rep == 1) total = ms;
                }
                if (best < 0.0f || total < best) {
                    best = total;
                    best_launch = l;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->launch = best_launch;
    b->init(s);
    if (launch_out) {
        launch_out[0] = std::uint64_t(best_launch.keys_per_batch);
        launch_out[1] = best_launch.item_keys;
        launch_out[2] = std::uint64_t(best_launch.blocks_per_cu);
    }
    if (ms_out) *ms_out = best;
    return int(hipStreamSynchronize(s));
}

}  // namespace

#define WC_BATCH_API(NAME, K, V)                                                                                  \
    using Batch_##NAME = BatchApi<K, V>;                                                                          \
    extern "C" void* wc_batch_##NAME##_create(const uint64_t* min_capacities, uint64_t num_tables, K seed,        \
                                              uint32_t cg_size) {                                                 \
        if (!min_capacities || num_tables == 0 || num_tables >= (uint64_t(1) << 32)) return nullptr;              \
        try { return make<K, V>(min_capacities, num_tables, seed, cg_size ? cg_size : 16); }                     \
        catch (...) { return nullptr; }                                                                           \
    }                                                                                                             \
    extern "C" void wc_batch_##NAME##_destroy(void* h) { delete static_cast<Batch_##NAME*>(h); }                 \
    extern "C" int wc_batch_##NAME##_init(void* h, void* stream) {                                                \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<Batch_##NAME*>(h)->init(as_stream(stream));                                                   \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_insert(void* h, const K* keys_in, const V* values_in,                        \
                                            const uint64_t* offsets, uint64_t num_in, void* stream,               \
                                            uint64_t probing_length, uint32_t* status_out) {                      \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto* b = static_cast<Batch_##NAME*>(h);                                                                  \
        return b->insert(keys_in, values_in, offsets, num_in, as_stream(stream), probing_length,                  \
                         reinterpret_cast<Status*>(status_out), b->launch);                                       \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_retrieve(void* h, const K* keys_in, const uint64_t* offsets,                 \
                                              uint64_t num_in, V* values_out, void* stream,                       \
                                              uint64_t probing_length, uint32_t* status_out) {                    \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto* b = static_cast<Batch_##NAME*>(h);                                                                  \
        return b->retrieve(keys_in, offsets, num_in, values_out, as_stream(stream), probing_length,               \
                           reinterpret_cast<Status*>(status_out), b->launch);                                     \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_set_launch(void* h, uint32_t keys_per_batch, uint64_t item_keys,             \
                                                uint32_t blocks_per_cu) {                                         \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        if (item_keys == 0 || blocks_per_cu == 0 || (keys_per_batch != 4 && keys_per_batch != 8))                 \
            return wc_capi::kBadArgument;                                                                         \
        static_cast<Batch_##NAME*>(h)->launch = detail::BatchLaunch{int(keys_per_batch), item_keys,               \
                                                                    int(blocks_per_cu)};                          \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_get_launch(void* h, uint64_t* launch_out) {                                  \
        if (!h || !launch_out) return wc_capi::kBadHandle;                                                        \
        const auto& l = static_cast<Batch_##NAME*>(h)->launch;                                                    \
        launch_out[0] = uint64_t(l.keys_per_batch);                                                               \
        launch_out[1] = l.item_keys;                                                                              \
        launch_out[2] = uint64_t(l.blocks_per_cu);                                                                \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_tune(void* h, const K* keys_in, const V* values_in, const uint64_t* offsets, \
                                          uint64_t num_in, const K* queries, const uint64_t* query_offsets,       \
                                          uint64_t num_queries, V* values_out, void* stream,                      \
                                          uint64_t* launch_out, double* ms_out) {                                 \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        return tune<K, V>(static_cast<Batch_##NAME*>(h), keys_in, values_in, offsets, num_in, queries,            \
                          query_offsets, num_queries, values_out, as_stream(stream), launch_out, ms_out);         \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_num_tables(void* h) {                                                   \
        return h ? static_cast<Batch_##NAME*>(h)->num_tables() : 0;                                               \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_capacity(void* h, uint64_t table) {                                     \
        return h ? static_cast<Batch_##NAME*>(h)->capacity(table) : 0;                                            \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_total_slots(void* h) {                                                  \
        return h ? static_cast<Batch_##NAME*>(h)->total_slots() : 0;                                              \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_bytes_total(void* h) {                                                  \
        return h ? static_cast<Batch_##NAME*>(h)->bytes_total() : 0;                                              \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_copy_status_words(void* h, uint32_t* words_out, void* stream) {              \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto* b = static_cast<Batch_##NAME*>(h);                                                                  \
        return int(hipMemcpyAsync(words_out, b->status_words(), sizeof(uint32_t) * b->num_tables(),               \
                                  hipMemcpyDeviceToDevice, as_stream(stream)));                                   \
    }                                                                                                             \
    extern "C" uint32_t wc_batch_##NAME##_cg_size(void* h) { return h ? static_cast<Batch_##NAME*>(h)->cg() : 0; }

WC_BATCH_API(u32u32, uint32_t, uint32_t)
WC_BATCH_API(u64u64, uint64_t, uint64_t)
