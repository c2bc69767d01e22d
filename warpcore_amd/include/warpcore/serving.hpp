// Serving loop: device-resident SingleValueHashTables serve insert /
// retrieve / erase requests from many clients through continuous batching.
//
// No reference counterpart (the reference is a library of one-call-per-
// batch operations). Clients submit requests (device buffers of keys,
// values, outputs); a scheduler thread admits them, coalesces the admitted
// requests into steps, and runs a step as one fixed-shape sequence of
// launches (work-item listing and one persistent kernel over small work
// items of every request in the step), built once as a HIP graph and
// replayed; only the request descriptors are uploaded per step. Steps are
// issued back to back on one stream while the host already builds the next
// one, so requests that arrive while a step runs join the following step; a
// request completes when the event recorded after its (last) step has fired.
//
// A step holds at most max_step_keys keys and max_requests requests: a
// request larger than what is left of a step runs in chunks over several
// steps, so a large request does not hold up the small ones that arrive
// behind it for longer than one bounded step.
//
// Latency mode: while the backlog is short (fewer than direct_backlog keys
// pending), a request of at least direct_keys keys is a step of its own and
// runs on its table's bulk kernels, the kernels of one plain call, which
// keep more wavefronts per SIMD than the fused step kernel; a step that
// holds one whole request runs there too. Batching only pays for several
// small requests at once, and a long backlog (throughput) goes back to full
// steps.
//
// Admission against the HBM budget: every table has a slot budget of
// max_load * capacity, counted in claimed slots (stored keys and the
// tombstones erase leaves, which inserts never reuse). An insert of n keys
// reserves n slots of its table when it is admitted (a key may already be
// stored, so this is an upper bound). An insert that does not fit waits
// until the steps in flight have run; then the table's reservation is
// recounted as its claimed slots (one scan of the table: what the earlier
// inserts really claimed); if it still does not fit, the table is compacted
// when it holds tombstones (its live pairs are read out, the table
// re-initialised and the pairs inserted again, which drops the tombstones;
// slot placement changes, results do not: with the default probing length
// they depend only on the stored pairs while the table has an empty slot,
// which the budget keeps); if it still does not fit, its keys are looked
// up, and the ones not stored (an upper bound of the slots it can claim,
// exact up to repeats inside the request) are its reservation instead of n;
// an insert that does not fit even then is refused (status out_of_memory
// for every key).
//
// Failures: a step whose launch or completion event reports an error fails
// every request it holds (wait() returns an error code; outputs and statuses
// of such a request are undefined).
//
// Order: per table, requests run in arrival order. A step takes, per
// table, either any run of consecutive lookups, or one insert or erase (or
// one chunk of it), so no two requests of one step conflict; steps run in
// order on the stream. Each request therefore sees exactly the table state
// it would see if all requests of its table ran one at a time in arrival
// order, which is what the tests replay. (An insert that runs in chunks
// behaves like the same keys inserted by consecutive calls: for duplicate
// keys across chunks the first chunk's copy reports `none`.)
#pragma once

#include "base.hpp"
#include "engine.hpp"
#include "lockstep_probing.hpp"
#include "refill_probing.hpp"
#include "scan.hpp"
#include "single_value_hash_table.hpp"
#include "status.hpp"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace warpcore {
namespace serving {

enum class Op : std::uint32_t { insert = 0, retrieve = 1, erase = 2 };

// Per-request descriptor read by the step kernel.
template<class K, class V>
struct RequestDesc {
    const K* keys;
    const V* values;        // insert
    V* values_out;          // retrieve
    std::uint32_t* status_out;   // may be null
    index_t n;
    std::uint32_t table;
    std::uint32_t op;
};

}  // namespace serving

namespace kernels {

// Step preparation (one workgroup): copies the step's live descriptors from
// the host-pinned staging buffer the scheduler wrote, counts each request's
// work items and writes their exclusive scan (first item of each request).
template<class Desc>
__global__ __launch_bounds__(256) void serve_prepare(const Desc* __restrict__ staged,
                                                     const std::uint32_t* __restrict__ staged_num, index_t max_reqs,
                                                     index_t item_keys, Desc* __restrict__ reqs,
                                                     index_t* __restrict__ first, index_t* __restrict__ state) {
    const index_t num_raw = reinterpret_cast<const volatile std::uint32_t*>(staged_num)[0];
    const index_t num = num_raw < max_reqs ? num_raw : max_reqs;
    const index_t step_item_keys = reinterpret_cast<const volatile std::uint32_t*>(staged_num)[1];
    if (step_item_keys != 0 && step_item_keys < item_keys) item_keys = step_item_keys;
    index_t carry = 0;
    for (index_t base = 0; base < num; base += blockDim.x) {
        const index_t r = base + threadIdx.x;
        index_t k = 0;
        if (r < num) {
            const Desc d = staged[r];
            reqs[r] = d;
            k = SDIV(d.n, item_keys);
        }
        index_t total;
        const index_t incl = detail::block_inclusive_scan(k, total);
        if (r < num) first[r] = carry + incl - k;
        carry += total;
    }
    if (threadIdx.x == 0) {
        state[0] = num;         // live requests
        state[1] = carry;       // work items
        state[2] = item_keys;   // keys per work item of this step
    }
}

// Wavefronts per SIMD the step kernel is compiled for (0: compiler's choice).
#ifndef WARPCORE_SERVE_WAVES
#define WARPCORE_SERVE_WAVES 0
#endif
#if WARPCORE_SERVE_WAVES > 0
#define WC_SERVE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WARPCORE_SERVE_WAVES)))
#else
#define WC_SERVE_OCCUPANCY
#endif

// One step: wavefront v of the (resident) grid takes work items v, v + V,
// ... (items hold item_keys keys, so a static assignment balances and
// needs no shared counter). An item's request is found by binary search in
// an LDS copy of the requests' first items; the wavefront then runs the
// lane-owner range of the request's operation (owner_probing.hpp) on the
// request's table, 64 keys in flight.
template<class View>
__global__ __launch_bounds__(256) WC_SERVE_OCCUPANCY void serve_step(const serving::RequestDesc<typename View::key_type,
                                                                             typename View::value_type>* __restrict__ reqs,
                                                  const index_t* __restrict__ first_item,
                                                  index_t* __restrict__ state, index_t max_reqs, index_t item_keys,
                                                  const View* __restrict__ views, index_t probing_length) {
    using K = typename View::key_type;
    using W = std::uint32_t;
    static_assert(detail::owner_capable<View>(), "the step kernel runs the lane-owner ranges");
    __shared__ detail::OwnerEntry<K, W> lds[256];
    extern __shared__ index_t first_lds[];
    auto* my = lds + (threadIdx.x & ~63u);
    const index_t num = state[0] < max_reqs ? state[0] : max_reqs;
    const index_t total = state[1];
    item_keys = state[2];
    for (index_t r = threadIdx.x; r < num; r += blockDim.x) first_lds[r] = first_item[r];
    __syncthreads();
    using H = status_handlers::ReturnStatus;
    const index_t wave = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const index_t waves = (index_t(gridDim.x) * blockDim.x) / 64;
    for (index_t item = wave; item < total; item += waves) {
        index_t lo = 0, hi = num;   // last request whose first item is <= item
        while (hi - lo > 1) {
            const index_t mid = (lo + hi) / 2;
            if (first_lds[mid] <= item) lo = mid;
            else hi = mid;
        }
        const std::uint32_t r = std::uint32_t(lo);
        const auto d = reqs[r];
        const index_t begin = (item - first_lds[r]) * item_keys;
        const index_t end = begin + item_keys < d.n ? begin + item_keys : d.n;
        const View view = views[d.table];
        const index_t budget = probing_schemes::detail::window_budget(probing_length, View::cg_size(),
                                                                      view.geo.windows);
        detail::DeferredJoin join;
        auto* st = reinterpret_cast<Status*>(d.status_out);
        if (d.op == std::uint32_t(serving::Op::insert))
            detail::insert_owner_range<View, H, W, true>(d.keys, d.values, begin, end, view, budget, st, my, join);
        else if (d.op == std::uint32_t(serving::Op::retrieve))
            detail::retrieve_owner_range<View, H, W>(d.keys, begin, end, d.values_out, view, budget, st, my, join);
        else
            detail::retrieve_owner_range<View, H, W, 1, true>(d.keys, begin, end, nullptr, view, budget, st, my,
                                                              join);
        join.flush(view.status);
    }
}

}  // namespace kernels

namespace serving {

struct Stats {
    index_t requests = 0, steps = 0, keys = 0;
    index_t direct_steps = 0;   // steps run on a table's own bulk kernels
    index_t compactions = 0;    // tables rebuilt to drop their tombstones
    index_t failed = 0;         // requests that failed with their step
};

// wait() results
constexpr int kServed = 1;
constexpr int kRefused = 0;        // out of budget: every key reported out_of_memory
constexpr int kFailed = -6;        // a step of the request failed to launch or to complete
constexpr int kUnknownId = -7;     // never issued, or collected and dropped from the history

// K/V: u32 or u64; CG: probing group (8 for 16-byte slots, 16 for 8-byte).
template<class K, class V, index_t CG>
class Server {
public:
    using table_type = SingleValueHashTable<K, V, defaults::empty_key<K>(), defaults::tombstone_key<K>(),
                                            defaults::probing_scheme_t<K, CG>>;
    using view_type = typename table_type::view_type;
    using desc_type = RequestDesc<K, V>;
    using clock = std::chrono::steady_clock;

    struct Config {
        index_t max_requests = 1024;                 // requests (chunks) per step: descriptor array
        index_t item_keys = 2048;                    // keys per work item (at most)
        index_t max_step_keys = index_t(1) << 24;    // keys per step
        float max_load = 0.9f;                       // admission: slot budget per table
        int blocks_per_cu = 8;                       // persistent grid
        bool use_graph = true;                       // replay the step as a HIP graph
        // Latency mode: while fewer than direct_backlog keys are pending, a
        // request of at least direct_keys keys runs as a step of its own on
        // its table's bulk kernels (0: every step runs the step kernel).
        index_t direct_keys = index_t(1) << 18;
        index_t direct_backlog = index_t(1) << 24;
        // completed requests whose outcome stays available to wait() /
        // latency_us() / step_of() (older ones are dropped)
        index_t history = index_t(1) << 20;
    };

    struct Request {
        desc_type d;
        index_t id;
        index_t issued = 0;              // keys handed to steps so far
        clock::time_point arrival, done;
        std::uint64_t step = 0;          // step that finished it (0 = not yet)
        std::atomic<int> state{0};       // 0 queued, 1 running, 2 done
        bool refused = false;
        bool failed = false;
        bool probed = false;    // its keys were looked up for admission
        index_t need = 0;       // probed: keys of it not stored
    };

    Server(const index_t* table_capacities, index_t num_tables, Config cfg = Config{}) : cfg_(cfg) {
        if (cfg_.item_keys == 0 || cfg_.max_requests == 0 || cfg_.max_step_keys == 0) {
            ok_ = false;
            return;
        }
        for (index_t t = 0; t < num_tables; ++t) {
            tables_.emplace_back(new table_type(table_capacities[t]));
            // the step kernel keeps window indices in 32 bits
            if (!tables_.back()->is_initialized() || tables_.back()->capacity() / CG >= (index_t(1) << 31)) {
                ok_ = false;
                return;
            }
            budget_.push_back(index_t(double(tables_.back()->capacity()) * cfg_.max_load));
            reserved_.push_back(0);
            recounted_.push_back(1);
            compacted_.push_back(0);
        }
        std::vector<view_type> views;
        for (auto& t : tables_) views.push_back(t->view());
        ok_ = alloc(&views_d_, views.size()) && alloc(&first_, cfg_.max_requests) && alloc(&state_, 3);
        for (int b = 0; b < 2 && ok_; ++b) {
            ok_ = alloc(&desc_d_[b], cfg_.max_requests) &&
                  hipHostMalloc(reinterpret_cast<void**>(&desc_h_[b]), sizeof(desc_type) * cfg_.max_requests) ==
                      hipSuccess &&
                  hipHostMalloc(reinterpret_cast<void**>(&num_reqs_h_[b]), 4 * sizeof(std::uint32_t)) == hipSuccess;
        }
        if (!ok_) return;
        // blocking streams: work the client queued on the default stream
        // (filling its buffers) is ordered before the steps
        ok_ = hipStreamCreate(&stream_) == hipSuccess && hipStreamCreate(&side_) == hipSuccess &&
              hipMemcpy(views_d_, views.data(), sizeof(view_type) * views.size(), hipMemcpyHostToDevice) == hipSuccess;
        if (!ok_) return;
        // persistent grid: as many blocks as can be resident (occupancy),
        // at most blocks_per_cu per CU
        int resident = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, step_kernel(), int(block_size()),
                                                         sizeof(index_t) * cfg_.max_requests) != hipSuccess ||
            resident < 1)
            resident = 1;
        step_grid_ = unsigned(detail::num_compute_units()) * unsigned(std::min(resident, cfg_.blocks_per_cu));
        worker_ = std::thread([this] { loop(); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            stop_flag_.store(true);
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (auto& g : graphs_) if (g.exec) (void)hipGraphExecDestroy(g.exec);
        for (void* p : {(void*)views_d_, (void*)first_, (void*)state_, (void*)desc_d_[0], (void*)desc_d_[1]})
            if (p) (void)hipFree(p);
        for (int b = 0; b < 2; ++b) {
            if (desc_h_[b]) (void)hipHostFree(desc_h_[b]);
            if (num_reqs_h_[b]) (void)hipHostFree(num_reqs_h_[b]);
        }
        if (stream_) (void)hipStreamDestroy(stream_);
        if (side_) (void)hipStreamDestroy(side_);
    }

    bool ok() const noexcept { return ok_; }
    table_type& table(index_t t) { return *tables_[t]; }
    index_t num_tables() const noexcept { return tables_.size(); }
    hipStream_t stream() const noexcept { return stream_; }

    // Submit a request; the buffers must stay valid until it is done.
    // Returns the request id. A request naming no table, or missing the
    // keys, the values of an insert or the output of a lookup, is refused.
    index_t submit(Op op, std::uint32_t table, const K* keys, const V* values, V* values_out,
                   std::uint32_t* status_out, index_t n) {
        auto r = std::make_shared<Request>();
        r->d = desc_type{keys, values, values_out, status_out, n, table, std::uint32_t(op)};
        r->arrival = clock::now();
        const bool bad = table >= tables_.size() || (n != 0 && keys == nullptr) ||
                         (op == Op::insert && n != 0 && values == nullptr) ||
                         (op == Op::retrieve && n != 0 && values_out == nullptr);
        index_t id;
        {
            std::lock_guard<std::mutex> g(mu_);
            id = next_id_++;
            r->id = id;
            all_.emplace(id, r);
            ++outstanding_;
            if (bad) {
                finish_refused(*r);
            } else {
                pending_.push_back(r);
                waiting_.store(pending_.size(), std::memory_order_release);
            }
        }
        cv_.notify_all();
        return id;
    }

    // Block until request `id` is done: kServed, kRefused, kFailed, or
    // kUnknownId for an id never issued or dropped from the history.
    int wait(index_t id) {
        std::unique_lock<std::mutex> g(mu_);
        auto r = find(id);
        if (!r) return kUnknownId;
        done_cv_.wait(g, [&] { return r->state.load() == 2; });
        return r->failed ? kFailed : (r->refused ? kRefused : kServed);
    }
    void drain() {
        std::unique_lock<std::mutex> g(mu_);
        done_cv_.wait(g, [&] { return outstanding_ == 0; });
    }
    // latency of request id in microseconds (arrival to completion seen);
    // negative for an unknown id or a request not done yet
    double latency_us(index_t id) {
        std::lock_guard<std::mutex> g(mu_);
        auto r = find(id);
        if (!r || r->state.load() != 2) return -1.0;
        return std::chrono::duration<double, std::micro>(r->done - r->arrival).count();
    }
    std::uint64_t step_of(index_t id) {
        std::lock_guard<std::mutex> g(mu_);
        auto r = find(id);
        return r ? r->step : 0;
    }
    Stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        return stats_;
    }
    // slots of table t counted against its budget (claimed slots as of the
    // last recount, plus the reservations of the inserts admitted since)
    index_t reserved(index_t t) {
        std::lock_guard<std::mutex> g(mu_);
        return t < reserved_.size() ? reserved_[t] : 0;
    }
    index_t budget(index_t t) const noexcept { return t < budget_.size() ? budget_[t] : 0; }
    // Failure handling from the host: step number `step` (1, 2, ... in
    // issue order) is not launched and reports a launch failure, so its
    // requests fail (wait() returns kFailed) while later steps run.
    void fail_step(std::uint64_t step) noexcept { fail_step_.store(step); }

private:
    struct Chunk {
        std::shared_ptr<Request> r;
        bool last;      // the request's final chunk
        index_t keys;   // keys of this chunk
    };
    using StepList = std::vector<Chunk>;

    template<class T>
    static bool alloc(T** p, index_t n) {
        *p = nullptr;
        return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1)) == hipSuccess;
    }

    // under mu_
    std::shared_ptr<Request> find(index_t id) const {
        const auto it = all_.find(id);
        return it == all_.end() ? nullptr : it->second;
    }

    // under mu_: request r is done; keep its outcome for the last
    // cfg_.history completions
    void complete(Request& r) {
        r.done = clock::now();
        r.state.store(2);
        --outstanding_;
        history_.push_back(r.id);
        while (history_.size() > cfg_.history) {
            all_.erase(history_.front());
            history_.pop_front();
        }
        done_cv_.notify_all();
    }

    void finish_refused(Request& r) {
        // every key of a refused request reports out_of_memory; on a side
        // stream, as the step stream may be busy
        r.refused = true;
        if (r.d.status_out && r.d.n)
            storage::kernels::fill<std::uint32_t>
                <<<detail::grid_for(r.d.n, 1, 8), block_size(), 0, side_>>>(r.d.status_out, r.d.n,
                                                                            Status::out_of_memory().base());
        if (hipStreamSynchronize(side_) != hipSuccess) r.failed = true;
        complete(r);
    }

    // Requests (chunks) of the next step, chosen under mu_: per table in
    // arrival order, a run of lookups or one mutation; an insert is admitted
    // against its table's budget when its first chunk is taken. Writes the
    // chunk descriptors to `descs`. In latency mode (short backlog) a large
    // request is taken alone, whole, and `direct` is set: the step runs on
    // the table's own bulk kernels; so does a step that holds one whole
    // request. An insert that does not fit waits while steps are in flight
    // (`idle` false: their unused reservations come back when they retire),
    // and no request after it is taken until they have; with none in flight
    // it asks for its table's claimed slots to be recounted (`recount`),
    // then for the table to be compacted (`compact`) unless that was done
    // since the table's last erase, then for its keys to be looked up
    // (`probe`: the ones not stored are its reservation), and is refused
    // when it does not fit even then.
    void pick(StepList& step, desc_type* descs, bool& direct, bool idle, index_t& recount, index_t& compact,
              std::shared_ptr<Request>& probe) {
        std::vector<int> mode(tables_.size(), -1);   // -1 untouched, 0 lookups, 1 mutation, 2 closed
        std::deque<std::shared_ptr<Request>> keep;
        index_t keys = 0;
        direct = false;
        recount = ~index_t(0);
        compact = ~index_t(0);
        probe = nullptr;
        bool hold = false;   // an insert waits for the device to go idle: take nothing more
        index_t backlog = 0;
        if (cfg_.direct_keys != 0)
            for (const auto& r : pending_) backlog += r->d.n - r->issued;
        const bool latency = cfg_.direct_keys != 0 && backlog < cfg_.direct_backlog;
        while (!pending_.empty()) {
            auto r = pending_.front();
            pending_.pop_front();
            const std::uint32_t t = r->d.table;
            const Op op = Op(r->d.op);
            bool take = !direct && !hold && recount == ~index_t(0) && compact == ~index_t(0) && !probe &&
                        step.size() < cfg_.max_requests &&
                        keys < cfg_.max_step_keys && mode[t] != 2 && mode[t] != 1 &&
                        (op == Op::retrieve || mode[t] == -1);
            const bool large = latency && r->d.n - r->issued >= cfg_.direct_keys;
            if (take && large && !step.empty()) take = false;   // runs alone in a later step
            const index_t want = r->probed ? r->need : r->d.n;
            if (take && op == Op::insert && r->issued == 0 && reserved_[t] + want > budget_[t]) {
                if (!idle || !step.empty()) {
                    // wait for the steps in flight (and this one) to retire;
                    // no new step starts meanwhile, so the device goes idle
                    take = false;
                    hold = true;
                } else if (!recounted_[t]) {
                    take = false;
                    recount = t;
                } else if (!compacted_[t]) {
                    take = false;
                    compact = t;
                } else if (!r->probed) {
                    take = false;
                    probe = r;
                } else {
                    finish_refused(*r);
                    continue;
                }
            }
            if (take && op == Op::insert && r->issued == 0) {
                reserved_[t] += want;
                recounted_[t] = 0;
            }
            if (!take) {
                mode[t] = 2;   // later requests of this table wait behind it
                keep.push_back(r);
                continue;
            }
            const index_t left = r->d.n - r->issued;
            const index_t room = cfg_.max_step_keys - keys;
            const index_t c = large || left < room ? left : room;
            direct = large;
            desc_type d = r->d;
            d.keys += r->issued;
            if (d.values) d.values += r->issued;
            if (d.values_out) d.values_out += r->issued;
            if (d.status_out) d.status_out += r->issued;
            d.n = c;
            descs[step.size()] = d;
            r->issued += c;
            const bool last = r->issued == r->d.n;
            step.push_back(Chunk{r, last, c});
            keys += c;
            mode[t] = op == Op::retrieve ? 0 : 1;
            if (!last) {   // the rest runs in a later step, ahead of this table's later requests
                mode[t] = 2;
                keep.push_back(r);
            }
        }
        // latency mode: a step of one whole request gains nothing from the
        // step kernel, so it runs on the bulk kernels too (grid sized to it)
        if (latency && step.size() == 1 && step[0].keys == step[0].r->d.n) direct = true;
        pending_.swap(keep);
        waiting_.store(pending_.size(), std::memory_order_release);
    }

    // Table t's reservation becomes its claimed slots (no step in flight).
    bool recount_table(index_t t) {
        (void)hipGetLastError();   // drop the not-ready of earlier event polls
        const index_t claimed = tables_[t]->claimed_slots(stream_);
        const bool ok = hipGetLastError() == hipSuccess;
        std::lock_guard<std::mutex> g(mu_);
        if (ok) reserved_[t] = claimed;
        recounted_[t] = 1;
        return ok;
    }

    // Rebuilds table t without its tombstones (no step in flight): its live
    // pairs are read out, the table is re-initialised and they are inserted
    // again. Its claimed slots are then its live pairs. Returns false on a
    // device error.
    bool compact_table(index_t t) {
        table_type& tab = *tables_[t];
        (void)hipGetLastError();   // drop the not-ready of earlier event polls
        const index_t live = tab.size(stream_);
        bool ok = hipGetLastError() == hipSuccess;
        if (ok && live < reserved_[t]) {
            K* keys = nullptr;
            V* vals = nullptr;
            ok = alloc(&keys, live) && alloc(&vals, live);
            if (ok) {
                index_t got = 0;
                tab.retrieve_all(keys, vals, got, stream_);
                ok = got == live;
                if (ok) {
                    tab.init(stream_);
                    tab.insert(keys, vals, got, stream_);
                    ok = hipStreamSynchronize(stream_) == hipSuccess;
                }
            }
            if (keys) (void)hipFree(keys);
            if (vals) (void)hipFree(vals);
            if (ok) {
                std::lock_guard<std::mutex> g(mu_);
                reserved_[t] = live;
                stats_.compactions++;
            }
        }
        std::lock_guard<std::mutex> g(mu_);
        compacted_[t] = 1;
        return ok;
    }

    // Admission probe (no step in flight): the keys of insert r not stored
    // in its table, counted from a lookup of all of them. Returns false on a
    // device error (r is then refused).
    bool probe_request(Request& r) {
        table_type& tab = *tables_[r.d.table];
        const index_t n = r.d.n;
        std::uint32_t* st = nullptr;
        V* out = nullptr;
        std::vector<std::uint32_t> host(n);
        bool ok = alloc(&st, n) && alloc(&out, n);
        if (ok) {
            tab.template retrieve<status_handlers::ReturnStatus>(r.d.keys, n, out, stream_, defaults::probing_length(),
                                                                reinterpret_cast<Status*>(st));
            ok = hipMemcpyAsync(host.data(), st, sizeof(std::uint32_t) * n, hipMemcpyDeviceToHost, stream_) ==
                     hipSuccess &&
                 hipStreamSynchronize(stream_) == hipSuccess;
        }
        if (st) (void)hipFree(st);
        if (out) (void)hipFree(out);
        index_t missing = 0;
        for (const std::uint32_t x : host) missing += x != Status::none().base();
        std::lock_guard<std::mutex> g(mu_);
        r.probed = true;
        r.need = ok ? missing : r.d.n;
        return ok;
    }

    struct Graph {
        hipGraphExec_t exec = nullptr;
    };

    // One step: preparation, work-item listing and the persistent kernel,
    // reading staging buffer b (fixed shape: max_requests descriptors, of
    // which the first *num_reqs_h_[b] are live).
    hipError_t enqueue_step(int b, hipStream_t stream) {
        (void)hipGetLastError();   // drop the not-ready of earlier event polls: report only these launches
        const index_t R = cfg_.max_requests;
        kernels::serve_prepare<desc_type><<<1, 256, 0, stream>>>(desc_h_[b], num_reqs_h_[b], R, cfg_.item_keys,
                                                                 desc_d_[b], first_, state_);
        step_kernel()<<<step_grid_, block_size(), sizeof(index_t) * R, stream>>>(
            desc_d_[b], first_, state_, R, cfg_.item_keys, views_d_, defaults::probing_length());
        return hipGetLastError();
    }

    WC_H static auto step_kernel() noexcept { return &kernels::serve_step<view_type>; }

    // The step as a HIP graph built node by node (no stream capture, so no
    // other thread's device calls can meet a capture in progress): the
    // prepare kernel, then the step kernel after it.
    bool build_graph(int b) {
        hipGraph_t g = nullptr;
        if (hipGraphCreate(&g, 0) != hipSuccess) return false;
        const desc_type* staged = desc_h_[b];
        const std::uint32_t* staged_num = num_reqs_h_[b];
        index_t R = cfg_.max_requests, item_keys = cfg_.item_keys, probing_length = defaults::probing_length();
        desc_type* reqs = desc_d_[b];
        const desc_type* reqs_c = desc_d_[b];
        index_t* first = first_;
        const index_t* first_c = first_;
        index_t* state = state_;
        const view_type* views = views_d_;
        void* prep_args[] = {&staged, &staged_num, &R, &item_keys, &reqs, &first, &state};
        void* step_args[] = {&reqs_c, &first_c, &state, &R, &item_keys, &views, &probing_length};
        hipKernelNodeParams prep{};
        prep.func = reinterpret_cast<void*>(kernels::serve_prepare<desc_type>);
        prep.gridDim = dim3(1);
        prep.blockDim = dim3(256);
        prep.sharedMemBytes = 0;
        prep.kernelParams = prep_args;
        prep.extra = nullptr;
        hipKernelNodeParams step{};
        step.func = reinterpret_cast<void*>(step_kernel());
        step.gridDim = dim3(step_grid_);
        step.blockDim = dim3(unsigned(block_size()));
        step.sharedMemBytes = unsigned(sizeof(index_t) * cfg_.max_requests);
        step.kernelParams = step_args;
        step.extra = nullptr;
        hipGraphNode_t n_prep = nullptr, n_step = nullptr;
        bool ok = hipGraphAddKernelNode(&n_prep, g, nullptr, 0, &prep) == hipSuccess &&
                  hipGraphAddKernelNode(&n_step, g, &n_prep, 1, &step) == hipSuccess &&
                  hipGraphInstantiate(&graphs_[b].exec, g, nullptr, nullptr, 0) == hipSuccess;
        (void)hipGraphDestroy(g);
        if (!ok) graphs_[b].exec = nullptr;
        return ok;
    }

    // A step of one request on its table's own bulk kernels (the kernels of
    // one plain call), on the step stream.
    hipError_t launch_direct(const desc_type& d) {
        (void)hipGetLastError();   // drop the not-ready of earlier event polls: report only these launches
        table_type& t = *tables_[d.table];
        const index_t pl = defaults::probing_length();
        auto* st = reinterpret_cast<Status*>(d.status_out);
        using H = status_handlers::ReturnStatus;
        switch (Op(d.op)) {
        case Op::insert:
            if (st) t.template insert<H>(d.keys, d.values, d.n, stream_, pl, st);
            else t.insert(d.keys, d.values, d.n, stream_, pl);
            break;
        case Op::retrieve:
            if (st) t.template retrieve<H>(d.keys, d.n, d.values_out, stream_, pl, st);
            else t.retrieve(d.keys, d.n, d.values_out, stream_, pl);
            break;
        case Op::erase:
            if (st) t.template erase<H>(d.keys, d.n, stream_, pl, st);
            else t.erase(d.keys, d.n, stream_, pl);
            break;
        }
        return hipGetLastError();
    }

    hipError_t launch(int b) {
        // the step reads its request count and descriptors from the pinned
        // staging buffer b, written by this thread before the launch
        std::atomic_thread_fence(std::memory_order_release);
        if (!cfg_.use_graph) return enqueue_step(b, stream_);
        if (!graphs_[b].exec && !build_graph(b)) {
            cfg_.use_graph = false;
            return enqueue_step(b, stream_);
        }
        return hipGraphLaunch(graphs_[b].exec, stream_);
    }

    // Keys per work item of a step: the smallest power of two from 64 (one
    // key per lane) that still gives every wavefront about four items (at
    // most cfg_.item_keys). A wavefront takes ceil(items / wavefronts)
    // items, so small items keep the busiest wavefront close to the average
    // and a small step ends sooner; large steps keep large items, whose
    // per-item setup and drain are amortised.
    index_t step_item_keys(const StepList& step) const noexcept {
        index_t keys = 0;
        for (const auto& c : step) keys += c.keys;
        const index_t waves = index_t(step_grid_) * block_size() / 64;
        index_t ik = 64;
        while (ik < cfg_.item_keys && ik * waves * 4 < keys) ik *= 2;
        return ik < cfg_.item_keys ? ik : cfg_.item_keys;
    }

    struct InFlight {
        hipEvent_t ev;
        std::uint64_t step_no;
        StepList chunks;
        bool failed;           // launch or completion error
    };

    // The oldest step is done (or failed): reopen compaction of tables it
    // erased from, and complete the requests whose last chunk it ran.
    void retire(std::deque<InFlight>& inflight) {
        InFlight& f = inflight.front();
        {
            std::lock_guard<std::mutex> g(mu_);
            for (auto& c : f.chunks) {
                const std::uint32_t t = c.r->d.table;
                if (c.r->d.op == std::uint32_t(Op::erase)) compacted_[t] = 0;
                if (f.failed) c.r->failed = true;
                if (!c.last) continue;
                c.r->step = f.step_no;
                stats_.requests++;
                stats_.keys += c.r->d.n;
                if (c.r->failed) stats_.failed++;
                complete(*c.r);
            }
        }
        (void)hipEventDestroy(f.ev);
        inflight.pop_front();
    }

    void loop() {
        std::deque<InFlight> inflight;
        int b = 0;
        std::uint64_t step_no = 0;
        while (true) {
            StepList step;
            bool direct = false;
            index_t recount = ~index_t(0), compact = ~index_t(0);
            std::shared_ptr<Request> probe;
            {
                std::unique_lock<std::mutex> g(mu_);
                if (inflight.empty()) cv_.wait(g, [&] { return stop_ || !pending_.empty(); });
                if (stop_ && pending_.empty() && inflight.empty()) break;
                // at most two steps in flight: descriptors are double-buffered
                if (inflight.size() < 2) pick(step, desc_h_[b], direct, inflight.empty(), recount, compact, probe);
                for (auto& c : step) c.r->state.store(1);
                if (!step.empty()) stats_.steps++;
                if (direct) stats_.direct_steps++;
            }
            if (recount != ~index_t(0)) {   // no step in flight: recount, then pick again
                (void)recount_table(recount);
                continue;
            }
            if (compact != ~index_t(0)) {   // no step in flight: rebuild, then pick again
                (void)compact_table(compact);
                continue;
            }
            if (probe) {   // no step in flight: count the insert's keys not stored, then pick again
                (void)probe_request(*probe);
                continue;
            }
            if (!step.empty()) {
                ++step_no;
                hipError_t err;
                if (step_no == fail_step_.load()) {
                    err = hipErrorLaunchFailure;
                } else if (direct) {
                    err = launch_direct(desc_h_[b][0]);
                } else {
                    num_reqs_h_[b][0] = std::uint32_t(step.size());
                    num_reqs_h_[b][1] = std::uint32_t(step_item_keys(step));
                    err = launch(b);
                }
                hipEvent_t ev = nullptr;
                if (err == hipSuccess) err = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                if (err == hipSuccess) err = hipEventRecord(ev, stream_);
                if (err != hipSuccess && ev != nullptr) {
                    (void)hipEventDestroy(ev);
                    ev = nullptr;
                }
                inflight.push_back(InFlight{ev, step_no, std::move(step), err != hipSuccess});
                b ^= 1;
            }
            // retire finished steps; with a free staging buffer, go back and
            // build the next step as soon as requests are pending (spinning:
            // the scheduler thread owns its core)
            while (!inflight.empty()) {
                InFlight& f = inflight.front();
                const hipError_t q = f.ev ? hipEventQuery(f.ev) : hipSuccess;
                if (q != hipErrorNotReady) {
                    if (q != hipSuccess) f.failed = true;
                    retire(inflight);
                    continue;
                }
                if (inflight.size() < 2) {   // poll: a new request or the step's end, whichever comes first
                    if (waiting_.load(std::memory_order_acquire) != 0 || stop_flag_.load()) break;
                    std::this_thread::yield();
                    continue;
                }
                if (hipEventSynchronize(f.ev) != hipSuccess) f.failed = true;
                retire(inflight);
            }
        }
    }

    Config cfg_;
    bool ok_ = true;
    std::vector<std::unique_ptr<table_type>> tables_;
    std::vector<index_t> budget_, reserved_;
    std::vector<char> recounted_;   // reserved_ is the claimed slots (no insert admitted since the recount)
    std::vector<char> compacted_;   // compacted since the table's last erase
    view_type* views_d_ = nullptr;
    desc_type* desc_d_[2] = {nullptr, nullptr};
    desc_type* desc_h_[2] = {nullptr, nullptr};
    std::uint32_t* num_reqs_h_[2] = {nullptr, nullptr};
    index_t *first_ = nullptr, *state_ = nullptr;   // state: live requests, work items, keys per item
    unsigned step_grid_ = 0;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;
    Graph graphs_[2];
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::shared_ptr<Request>> pending_;
    std::unordered_map<index_t, std::shared_ptr<Request>> all_;   // issued, not dropped from the history
    std::deque<index_t> history_;                                 // completed ids, oldest first
    index_t outstanding_ = 0;                                     // issued, not completed
    index_t next_id_ = 0;
    bool stop_ = false;
    std::atomic<bool> stop_flag_{false};   // stop_, readable without mu_
    std::atomic<std::size_t> waiting_{0};  // pending_.size(), readable without mu_
    std::atomic<std::uint64_t> fail_step_{0};
    Stats stats_;
    std::thread worker_;
};

}  // namespace serving
}  // namespace warpcore
