// extern "C" entry points of the serving loop (declared in
// include/warpcore_c.h): a Server over SingleValueHashTable<u64,u64> (cg 8)
// or <u32,u32> (cg 16), and an open-loop driver that replays a request
// list against the server or one launch per request.
#include "capi_common.hpp"

#include <warpcore/serving.hpp>

#include <chrono>
#include <thread>
#include <vector>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {

template<class K, class V, index_t CG>
using server_t = serving::Server<K, V, CG>;

template<class S>
S* make(const std::uint64_t* caps, std::uint64_t n, std::uint64_t max_requests, std::uint64_t item_keys,
        std::uint64_t max_step_keys, float max_load, int use_graph, int blocks_per_cu, std::uint64_t direct_keys) {
    typename S::Config cfg;
    if (max_requests) cfg.max_requests = max_requests;
    if (item_keys) cfg.item_keys = item_keys;
    if (max_step_keys) cfg.max_step_keys = max_step_keys;
    if (max_load > 0.0f) cfg.max_load = max_load;
    if (blocks_per_cu > 0) cfg.blocks_per_cu = blocks_per_cu;
    if (direct_keys) cfg.direct_keys = direct_keys == ~std::uint64_t(0) ? 0 : direct_keys;
    cfg.use_graph = use_graph != 0;
    S* s = new S(caps, n, cfg);
    if (!s->ok()) {
        delete s;
        return nullptr;
    }
    return s;
}

// Replays n requests at their arrival times (microseconds from the start).
// mode 0: submit to the server; mode 1: one plain table call per request on
// the server's stream, in arrival order, completion seen through events.
// latency_out[i] = completion - arrival (us); *makespan = last completion.
template<class S, class K, class V>
int drive(S* s, int mode, const std::uint32_t* ops, const std::uint32_t* tables, const std::uint64_t* sizes,
          const double* arrival_us, std::uint64_t n, const K* keys, const V* values, V* values_out,
          std::uint32_t* status_out, double* latency_out, double* makespan_us) {
    using clock = std::chrono::steady_clock;
    std::vector<std::uint64_t> off(n + 1, 0);
    for (std::uint64_t i = 0; i < n; ++i) off[i + 1] = off[i] + sizes[i];
    const auto t0 = clock::now();
    auto wait_until = [&](double us) {
        const auto at = t0 + std::chrono::nanoseconds(std::int64_t(us * 1e3));
        while (clock::now() < at) {
            const auto left = at - clock::now();
            if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left - std::chrono::microseconds(100));
        }
    };
    std::vector<double> arrival(n);
    if (mode == 0) {
        std::vector<index_t> ids(n);
        for (std::uint64_t i = 0; i < n; ++i) {
            wait_until(arrival_us[i]);
            arrival[i] = std::chrono::duration<double, std::micro>(clock::now() - t0).count();
            ids[i] = s->submit(serving::Op(ops[i]), tables[i], keys + off[i], values + off[i], values_out + off[i],
                               status_out ? status_out + off[i] : nullptr, sizes[i]);
        }
        s->drain();
        double last = 0.0;
        for (std::uint64_t i = 0; i < n; ++i) {
            latency_out[i] = s->latency_us(ids[i]);
            last = std::max(last, arrival[i] + latency_out[i]);
        }
        *makespan_us = last;
        return 0;
    }
    // one launch per request; a poller thread stamps completions in order
    hipStream_t st = s->stream();
    std::vector<hipEvent_t> ev(n, nullptr);
    std::vector<double> done(n, 0.0);
    std::atomic<std::uint64_t> issued{0};
    std::thread poller([&] {
        for (std::uint64_t i = 0; i < n; ++i) {
            while (issued.load() <= i) std::this_thread::yield();
            while (hipEventQuery(ev[i]) == hipErrorNotReady) std::this_thread::yield();
            done[i] = std::chrono::duration<double, std::micro>(clock::now() - t0).count();
        }
    });
    int rc = 0;
    for (std::uint64_t i = 0; i < n; ++i) {
        wait_until(arrival_us[i]);
        arrival[i] = std::chrono::duration<double, std::micro>(clock::now() - t0).count();
        auto& t = s->table(tables[i]);
        using H = status_handlers::ReturnStatus;
        auto* so = status_out ? reinterpret_cast<Status*>(status_out + off[i]) : nullptr;
        if (ops[i] == std::uint32_t(serving::Op::insert))
            t.template insert<H>(keys + off[i], values + off[i], sizes[i], st, defaults::probing_length(), so);
        else if (ops[i] == std::uint32_t(serving::Op::retrieve))
            t.template retrieve<H>(keys + off[i], sizes[i], values_out + off[i], st, defaults::probing_length(), so);
        else
            t.template erase<H>(keys + off[i], sizes[i], st, defaults::probing_length(), so);
        (void)hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
        (void)hipEventRecord(ev[i], st);
        issued.store(i + 1);
    }
    poller.join();
    double last = 0.0;
    for (std::uint64_t i = 0; i < n; ++i) {
        latency_out[i] = done[i] - arrival[i];
        last = std::max(last, done[i]);
        (void)hipEventDestroy(ev[i]);
    }
    *makespan_us = last;
    rc = int(hipGetLastError());
    return rc;
}

}  // namespace

#define WC_SERVER_API(NAME, K, V, CG)                                                                             \
    using Srv_##NAME = server_t<K, V, CG>;                                                                        \
    extern "C" void* wc_server_##NAME##_create(const uint64_t* table_capacities, uint64_t num_tables,             \
                                               uint64_t max_requests, uint64_t item_keys,                         \
                                               uint64_t max_step_keys, float max_load, int use_graph,             \
                                               int blocks_per_cu, uint64_t direct_keys) {                         \
        if (!table_capacities || num_tables == 0) return nullptr;                                                 \
        try {                                                                                                     \
            return make<Srv_##NAME>(table_capacities, num_tables, max_requests, item_keys, max_step_keys,         \
                                    max_load, use_graph, blocks_per_cu, direct_keys);                             \
        } catch (...) { return nullptr; }                                                                         \
    }                                                                                                             \
    extern "C" void wc_server_##NAME##_destroy(void* h) { delete static_cast<Srv_##NAME*>(h); }                  \
    extern "C" int64_t wc_server_##NAME##_submit(void* h, uint32_t op, uint32_t table, const K* keys_in,          \
                                                 const V* values_in, V* values_out, uint32_t* status_out,         \
                                                 uint64_t num_in) {                                               \
        if (!h || op > 2) return -1;                                                                              \
        try {                                                                                                     \
            return int64_t(static_cast<Srv_##NAME*>(h)->submit(serving::Op(op), table, keys_in, values_in,       \
                                                                values_out, status_out, num_in));                \
        } catch (...) { return -1; }                                                                              \
    }                                                                                                             \
    extern "C" int wc_server_##NAME##_wait(void* h, int64_t id) {                                                 \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        if (id < 0) return serving::kUnknownId;                                                                   \
        return static_cast<Srv_##NAME*>(h)->wait(index_t(id));                                                    \
    }                                                                                                             \
    extern "C" int wc_server_##NAME##_drain(void* h) {                                                            \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<Srv_##NAME*>(h)->drain();                                                                     \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" double wc_server_##NAME##_latency_us(void* h, int64_t id) {                                        \
        return h && id >= 0 ? static_cast<Srv_##NAME*>(h)->latency_us(index_t(id)) : -1.0;                       \
    }                                                                                                             \
    extern "C" uint64_t wc_server_##NAME##_step_of(void* h, int64_t id) {                                         \
        return h && id >= 0 ? static_cast<Srv_##NAME*>(h)->step_of(index_t(id)) : 0;                             \
    }                                                                                                             \
    extern "C" int wc_server_##NAME##_stats(void* h, uint64_t* out3) {                                            \
        if (!h || !out3) return wc_capi::kBadHandle;                                                              \
        const auto s = static_cast<Srv_##NAME*>(h)->stats();                                                      \
        out3[0] = s.requests;                                                                                     \
        out3[1] = s.steps;                                                                                        \
        out3[2] = s.keys;                                                                                         \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" uint64_t wc_server_##NAME##_direct_steps(void* h) {                                                \
        return h ? static_cast<Srv_##NAME*>(h)->stats().direct_steps : 0;                                        \
    }                                                                                                             \
    extern "C" int wc_server_##NAME##_counters(void* h, uint64_t* out4) {                                         \
        if (!h || !out4) return wc_capi::kBadHandle;                                                              \
        const auto s = static_cast<Srv_##NAME*>(h)->stats();                                                      \
        out4[0] = s.direct_steps;                                                                                 \
        out4[1] = s.compactions;                                                                                  \
        out4[2] = s.failed;                                                                                       \
        out4[3] = 0;                                                                                              \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" uint64_t wc_server_##NAME##_table_reserved(void* h, uint32_t table) {                              \
        if (!h || table >= static_cast<Srv_##NAME*>(h)->num_tables()) return 0;                                   \
        return static_cast<Srv_##NAME*>(h)->reserved(table);                                                      \
    }                                                                                                             \
    extern "C" int wc_server_##NAME##_fail_step(void* h, uint64_t step) {                                         \
        if (!h) return wc_capi::kBadHandle;                                                                       \
        static_cast<Srv_##NAME*>(h)->fail_step(step);                                                             \
        return 0;                                                                                                 \
    }                                                                                                             \
    extern "C" uint64_t wc_server_##NAME##_table_budget(void* h, uint32_t table) {                                \
        if (!h || table >= static_cast<Srv_##NAME*>(h)->num_tables()) return 0;                                   \
        return static_cast<Srv_##NAME*>(h)->budget(table);                                                        \
    }                                                                                                             \
    extern "C" uint64_t wc_server_##NAME##_table_size(void* h, uint32_t table) {                                  \
        if (!h || table >= static_cast<Srv_##NAME*>(h)->num_tables()) return 0;                                   \
        auto* s = static_cast<Srv_##NAME*>(h);                                                                    \
        return s->table(table).size(s->stream());                                                                 \
    }                                                                                                             \
    extern "C" uint64_t wc_server_##NAME##_table_capacity(void* h, uint32_t table) {                              \
        if (!h || table >= static_cast<Srv_##NAME*>(h)->num_tables()) return 0;                                   \
        return static_cast<Srv_##NAME*>(h)->table(table).capacity();                                              \
    }                                                                                                             \
    extern "C" uint32_t wc_server_##NAME##_table_status(void* h, uint32_t table) {                                \
        if (!h || table >= static_cast<Srv_##NAME*>(h)->num_tables()) return Status::not_initialized().base();    \
        auto* s = static_cast<Srv_##NAME*>(h);                                                                    \
        return s->table(table).peek_status(s->stream()).base();                                                   \
    }                                                                                                             \
    extern "C" int wc_server_##NAME##_drive(void* h, int mode, const uint32_t* ops, const uint32_t* tables,       \
                                            const uint64_t* sizes, const double* arrival_us, uint64_t num_requests,\
                                            const K* keys_in, const V* values_in, V* values_out,                  \
                                            uint32_t* status_out, double* latency_us_out, double* makespan_us) {  \
        if (!h || !ops || !tables || !sizes || !arrival_us || !latency_us_out || !makespan_us)                    \
            return wc_capi::kBadHandle;                                                                           \
        for (uint64_t i = 0; i < num_requests; ++i)                                                               \
            if (ops[i] > 2 || tables[i] >= static_cast<Srv_##NAME*>(h)->num_tables()) return wc_capi::kBadArgument; \
        try {                                                                                                     \
            return drive<Srv_##NAME, K, V>(static_cast<Srv_##NAME*>(h), mode, ops, tables, sizes, arrival_us,     \
                                          num_requests, keys_in, values_in, values_out, status_out,               \
                                          latency_us_out, makespan_us);                                           \
        } catch (...) { return wc_capi::kBadArgument; }                                                          \
    }

WC_SERVER_API(u64u64, uint64_t, uint64_t, 8)
WC_SERVER_API(u32u32, uint32_t, uint32_t, 16)
