// extern "C" entry points of TableBatch (declared in include/warpcore_c.h),
// including the per-class autotuner.
#include "capi_common.hpp"

#include <warpcore/table_batch.hpp>

#include <vector>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {

// Fields of one row of wc_batch_*_tune_report.
constexpr int kReportFields = 10;

template<class K, class V>
struct Batch {
    TableBatch<K, V> b;
    detail::BatchTuneReport report;
    bool tuned = false;
    Batch(const index_t* caps, index_t n, K seed, int config) : b(caps, n, seed, config) {}
};

template<class K, class V>
Batch<K, V>* make(const index_t* caps, index_t n, K seed, unsigned cg) {
    const int config = TableBatch<K, V>::config_index(cg, cg);
    if (config < 0) return nullptr;
    auto* b = new Batch<K, V>(caps, n, seed, config);
    if (!b->b.ok()) {
        delete b;
        return nullptr;
    }
    return b;
}

template<class K, class V>
int write_report(const Batch<K, V>* b, uint64_t* rows, uint64_t max_rows, double* totals) {
    if (!b->tuned) return 0;
    const auto& r = b->report;
    if (totals) {
        totals[0] = r.fixed_ms;
        totals[1] = r.tuned_ms;
    }
    const uint64_t n = r.classes.size();
    for (uint64_t i = 0; rows && i < n && i < max_rows; ++i) {
        const auto& c = r.classes[i];
        uint64_t* o = rows + i * kReportFields;
        o[0] = c.id;
        o[1] = c.tables;
        o[2] = c.keys;
        o[3] = c.queries;
        o[4] = uint64_t(c.fixed_config);
        o[5] = uint64_t(c.chosen_config);
        o[6] = uint64_t(c.fixed_ms * 1e6 + 0.5);   // ns
        o[7] = uint64_t(c.chosen_ms * 1e6 + 0.5);
        o[8] = c.fixed_slots;
        o[9] = c.chosen_slots;
    }
    return int(n);
}

}  // namespace

#define WC_BATCH_API(NAME, K, V)                                                                                  \
    using Batch_##NAME = Batch<K, V>;                                                                             \
    using Table_##NAME = TableBatch<K, V>;                                                                        \
    extern "C" void* wc_batch_##NAME##_create(const uint64_t* min_capacities, uint64_t num_tables, K seed,        \
                                              uint32_t cg_size) {                                                 \
        if (!min_capacities || num_tables == 0 || num_tables >= (uint64_t(1) << 32)) return nullptr;              \
        try { return make<K, V>(min_capacities, num_tables, seed, cg_size ? cg_size : 16); }                     \
        catch (...) { return nullptr; }                                                                           \
    }                                                                                                             \
    extern "C" void wc_batch_##NAME##_destroy(void* h) { delete static_cast<Batch_##NAME*>(h); }                 \
    extern "C" int wc_batch_##NAME##_init(void* h, void* stream) {                                                \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<Batch_##NAME*>(h)->b.init(as_stream(stream));                                                 \
        return last_error();                                                                                      \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_insert(void* h, const K* keys_in, const V* values_in,                        \
                                            const uint64_t* offsets, uint64_t num_in, void* stream,               \
                                            uint64_t probing_length, uint32_t* status_out) {                      \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto& b = static_cast<Batch_##NAME*>(h)->b;                                                               \
        if (status_out)                                                                                           \
            b.template insert<status_handlers::ReturnStatus>(keys_in, values_in, offsets, num_in,                 \
                                                             as_stream(stream), probing_length,                   \
                                                             reinterpret_cast<Status*>(status_out));              \
        else                                                                                                      \
            b.insert(keys_in, values_in, offsets, num_in, as_stream(stream), probing_length);                     \
        return b.last_error();                                                                                    \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_retrieve(void* h, const K* keys_in, const uint64_t* offsets,                 \
                                              uint64_t num_in, V* values_out, void* stream,                       \
                                              uint64_t probing_length, uint32_t* status_out) {                    \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto& b = static_cast<Batch_##NAME*>(h)->b;                                                               \
        if (status_out)                                                                                           \
            b.template retrieve<status_handlers::ReturnStatus>(keys_in, offsets, num_in, values_out,              \
                                                               as_stream(stream), probing_length,                 \
                                                               reinterpret_cast<Status*>(status_out));            \
        else                                                                                                      \
            b.retrieve(keys_in, offsets, num_in, values_out, as_stream(stream), probing_length);                  \
        return b.last_error();                                                                                    \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_set_launch(void* h, uint64_t item_keys, uint32_t blocks_per_cu,              \
                                                uint32_t block_size) {                                            \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        const detail::BatchLaunch l{item_keys, int(blocks_per_cu), int(block_size)};                             \
        return static_cast<Batch_##NAME*>(h)->b.set_launch(l) ? 0 : wc_capi::kBadArgument;                        \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_set_lds_insert(void* h, int on) {                                            \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<Batch_##NAME*>(h)->b.set_lds_insert(on != 0);                                                 \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_lds_tables(void* h) {                                                   \
        return h ? static_cast<Batch_##NAME*>(h)->b.lds_tables() : 0;                                             \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_get_launch(void* h, uint64_t* launch_out) {                                  \
        if (!h || !launch_out) return wc_capi::kBadHandle;                                                        \
        const auto& l = static_cast<Batch_##NAME*>(h)->b.launch();                                                \
        launch_out[0] = l.item_keys;                                                                              \
        launch_out[1] = uint64_t(l.blocks_per_cu);                                                                \
        launch_out[2] = uint64_t(l.block_size);                                                                   \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_tune(void* h, const K* keys_in, const V* values_in, const uint64_t* offsets, \
                                          uint64_t num_in, const K* queries, const uint64_t* query_offsets,       \
                                          uint64_t num_queries, V* values_out, void* stream,                      \
                                          uint64_t* launch_out, double* ms_out) {                                 \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        if (!offsets || (num_queries && !query_offsets)) return wc_capi::kBadArgument;                            \
        auto* bb = static_cast<Batch_##NAME*>(h);                                                                 \
        bb->tuned = false;                                                                                        \
        const int rc = wc_capi::guarded([&] {                                                                     \
            return bb->b.tune(keys_in, values_in, offsets, num_in, queries, query_offsets, num_queries,           \
                              values_out, as_stream(stream), &bb->report);                                        \
        });                                                                                                       \
        if (rc != 0) return rc;                                                                                   \
        bb->tuned = true;                                                                                         \
        if (launch_out) {                                                                                         \
            launch_out[0] = bb->report.launch.item_keys;                                                          \
            launch_out[1] = uint64_t(bb->report.launch.blocks_per_cu);                                            \
            launch_out[2] = uint64_t(bb->report.launch.block_size);                                               \
        }                                                                                                         \
        if (ms_out) *ms_out = bb->report.tuned_ms;                                                                \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_tune_report(void* h, uint64_t* rows, uint64_t max_rows, double* totals) {    \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        return write_report<K, V>(static_cast<Batch_##NAME*>(h), rows, max_rows, totals);                         \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_set_configs(void* h, const uint8_t* configs, uint64_t num_tables) {          \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto& b = static_cast<Batch_##NAME*>(h)->b;                                                               \
        if (!configs || num_tables != b.num_tables()) return wc_capi::kBadArgument;                               \
        for (uint64_t t = 0; t < num_tables; ++t)                                                                 \
            if (configs[t] >= Table_##NAME::num_configs()) return wc_capi::kBadArgument;                          \
        return wc_capi::guarded([&] { return b.set_configs(configs) ? 0 : wc_capi::kNoMemory; });                \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_table_config(void* h, uint64_t table) {                                      \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto& b = static_cast<Batch_##NAME*>(h)->b;                                                               \
        return table < b.num_tables() ? b.table_config(table) : wc_capi::kBadArgument;                            \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_num_configs(void) { return Table_##NAME::num_configs(); }                   \
    extern "C" int wc_batch_##NAME##_config_shape(uint32_t config, uint32_t* lanes_window) {                      \
        if (!lanes_window || config >= uint32_t(Table_##NAME::num_configs())) return wc_capi::kBadArgument;       \
        lanes_window[0] = Table_##NAME::config_lanes(int(config));                                                \
        lanes_window[1] = Table_##NAME::config_window(int(config));                                               \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_num_tables(void* h) {                                                   \
        return h ? static_cast<Batch_##NAME*>(h)->b.num_tables() : 0;                                             \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_capacity(void* h, uint64_t table) {                                     \
        if (!h) return 0;                                                                                         \
        auto& b = static_cast<Batch_##NAME*>(h)->b;                                                               \
        return table < b.num_tables() ? b.capacity(table) : 0;                                                    \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_total_slots(void* h) {                                                  \
        return h ? static_cast<Batch_##NAME*>(h)->b.total_slots() : 0;                                            \
    }                                                                                                             \
    extern "C" uint64_t wc_batch_##NAME##_bytes_total(void* h) {                                                  \
        return h ? static_cast<Batch_##NAME*>(h)->b.bytes_total() : 0;                                            \
    }                                                                                                             \
    extern "C" int wc_batch_##NAME##_copy_status_words(void* h, uint32_t* words_out, void* stream) {              \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        auto& b = static_cast<Batch_##NAME*>(h)->b;                                                               \
        return int(hipMemcpyAsync(words_out, b.status_words(), sizeof(uint32_t) * b.num_tables(),                 \
                                  hipMemcpyDeviceToDevice, as_stream(stream)));                                   \
    }                                                                                                             \
    extern "C" uint32_t wc_batch_##NAME##_cg_size(void* h) {                                                      \
        return h ? Table_##NAME::config_lanes(static_cast<Batch_##NAME*>(h)->b.fixed_config()) : 0;               \
    }

WC_BATCH_API(u32u32, uint32_t, uint32_t)
WC_BATCH_API(u64u64, uint64_t, uint64_t)
