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
// max_load * capacity. An insert of n keys reserves n slots of its table
// (a key may be a duplicate, so this is an upper bound); slots are never
// given back, because erase leaves tombstones that keep their slots, so an
// insert that does not fit its table's remaining budget is refused (status
// out_of_memory for every key).
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

// One step: group g of the (resident) grid takes work items g, g + G, ...
// (items hold item_keys keys, so a static assignment balances and needs no
// shared counter). An item's request is found by binary search in an LDS
// copy of the requests' first items. Lookups and inserts use the refill
// kernels' range code, erases the lockstep one.
template<class View, int UR, int UI, int UE>
__global__ __launch_bounds__(256) WC_SERVE_OCCUPANCY void serve_step(const serving::RequestDesc<typename View::key_type,
                                                                             typename View::value_type>* __restrict__ reqs,
                                                  const index_t* __restrict__ first_item,
                                                  index_t* __restrict__ state, index_t max_reqs, index_t item_keys,
                                                  const View* __restrict__ views, index_t probing_length) {
    using K = typename View::key_type;
    using W = std::uint32_t;
    constexpr unsigned CG = unsigned(View::cg_size());
    __shared__ detail::StagedKey<K, W> stage[256 / CG * UE];
    __shared__ detail::RefillEntry<K, W> refill[256];
    __shared__ detail::RefillInsertEntry<K, W, typename View::value_type> refill_insert[256];
    extern __shared__ index_t first_lds[];
    auto* my = stage + (threadIdx.x / CG) * UE;
    auto* my_refill = refill + (threadIdx.x & ~(CG - 1));
    auto* my_refill_insert = refill_insert + (threadIdx.x & ~(CG - 1));
    const unsigned rank = lane_id() & (CG - 1);
    const index_t num = state[0] < max_reqs ? state[0] : max_reqs;
    const index_t total = state[1];
    item_keys = state[2];
    for (index_t r = threadIdx.x; r < num; r += blockDim.x) first_lds[r] = first_item[r];
    __syncthreads();
    using H = status_handlers::ReturnStatus;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    for (index_t item = gid; item < total; item += groups) {
        {
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
            const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, view.geo.windows);
            detail::DeferredJoin join;
            auto* st = reinterpret_cast<Status*>(d.status_out);
            if (d.op == std::uint32_t(serving::Op::insert))
                detail::insert_refill_range<View, H, UI, W, true>(d.keys, d.values, begin, end, view, budget, st,
                                                                  my_refill_insert, join);
            else if (d.op == std::uint32_t(serving::Op::retrieve))
                detail::retrieve_refill_range<View, H, UR, W>(d.keys, begin, end, d.values_out, view, budget, st,
                                                              my_refill, join);
            else
                detail::erase_lockstep_range<View, H, UE, W>(d.keys, begin, end, UE, view, budget, st, my, join);
            const status_base_t all = tile<CG>::reduce_or(join.acc);
            if (rank == 0 && all != 0) view.status->atomic_join(Status::from_base(all));
        }
    }
}

}  // namespace kernels

namespace serving {

struct Stats {
    index_t requests = 0, steps = 0, keys = 0;
    index_t direct_steps = 0;   // steps run on a table's own bulk kernels
};

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
        index_t item_keys = 64;                      // keys per work item
        index_t max_step_keys = index_t(1) << 24;    // keys per step
        float max_load = 0.9f;                       // admission: slot budget per table
        int blocks_per_cu = 8;                       // persistent grid
        bool use_graph = true;                       // replay the step as a HIP graph
        // Latency mode: while fewer than direct_backlog keys are pending, a
        // request of at least direct_keys keys runs as a step of its own on
        // its table's bulk kernels (0: every step runs the step kernel).
        index_t direct_keys = index_t(1) << 18;
        index_t direct_backlog = index_t(1) << 24;
    };

    struct Request {
        desc_type d;
        index_t id;
        index_t issued = 0;              // keys handed to steps so far
        clock::time_point arrival, done;
        std::uint64_t step = 0;          // step that finished it (0 = not yet)
        std::atomic<int> state{0};       // 0 queued, 1 running, 2 done
        bool refused = false;
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
            all_.push_back(r);
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

    // Block until request `id` is done; returns true unless it was refused.
    bool wait(index_t id) {
        std::unique_lock<std::mutex> g(mu_);
        done_cv_.wait(g, [&] { return all_[id]->state.load() == 2; });
        return !all_[id]->refused;
    }
    void drain() {
        std::unique_lock<std::mutex> g(mu_);
        done_cv_.wait(g, [&] {
            for (auto& r : all_) if (r->state.load() != 2) return false;
            return true;
        });
    }
    // latency of request id in microseconds (arrival to completion seen)
    double latency_us(index_t id) {
        std::lock_guard<std::mutex> g(mu_);
        const auto& r = *all_[id];
        return std::chrono::duration<double, std::micro>(r.done - r.arrival).count();
    }
    std::uint64_t step_of(index_t id) {
        std::lock_guard<std::mutex> g(mu_);
        return all_[id]->step;
    }
    Stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        return stats_;
    }

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

    void finish_refused(Request& r) {
        // every key of a refused request reports out_of_memory; on a side
        // stream, as the step stream may be capturing a graph
        r.refused = true;
        if (r.d.status_out && r.d.n)
            storage::kernels::fill<std::uint32_t>
                <<<detail::grid_for(r.d.n, 1, 8), block_size(), 0, side_>>>(r.d.status_out, r.d.n,
                                                                            Status::out_of_memory().base());
        (void)hipStreamSynchronize(side_);
        r.done = clock::now();
        r.state.store(2);
        done_cv_.notify_all();
    }

    // Requests (chunks) of the next step, chosen under mu_: per table in
    // arrival order, a run of lookups or one mutation; an insert is admitted
    // against its table's budget when its first chunk is taken. Writes the
    // chunk descriptors to `descs`. In latency mode (short backlog) a large
    // request is taken alone, whole, and `direct` is set: the step runs on
    // the table's own bulk kernels; so does a step that holds one whole
    // request.
    void pick(StepList& step, desc_type* descs, bool& direct) {
        std::vector<int> mode(tables_.size(), -1);   // -1 untouched, 0 lookups, 1 mutation, 2 closed
        std::deque<std::shared_ptr<Request>> keep;
        index_t keys = 0;
        direct = false;
        index_t backlog = 0;
        if (cfg_.direct_keys != 0)
            for (const auto& r : pending_) backlog += r->d.n - r->issued;
        const bool latency = cfg_.direct_keys != 0 && backlog < cfg_.direct_backlog;
        while (!pending_.empty()) {
            auto r = pending_.front();
            pending_.pop_front();
            const std::uint32_t t = r->d.table;
            const Op op = Op(r->d.op);
            bool take = !direct && step.size() < cfg_.max_requests && keys < cfg_.max_step_keys && mode[t] != 2 &&
                        mode[t] != 1 && (op == Op::retrieve || mode[t] == -1);
            const bool large = latency && r->d.n - r->issued >= cfg_.direct_keys;
            if (take && large && !step.empty()) take = false;   // runs alone in a later step
            if (take && op == Op::insert && r->issued == 0) {
                if (reserved_[t] + r->d.n > budget_[t]) {
                    finish_refused(*r);
                    continue;
                }
                reserved_[t] += r->d.n;
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

    struct Graph {
        hipGraphExec_t exec = nullptr;
    };

    // One step: preparation, work-item listing and the persistent kernel,
    // reading staging buffer b (fixed shape: max_requests descriptors, of
    // which the first *num_reqs_h_[b] are live).
    void enqueue_step(int b, hipStream_t stream) {
        const index_t R = cfg_.max_requests;
        kernels::serve_prepare<desc_type><<<1, 256, 0, stream>>>(desc_h_[b], num_reqs_h_[b], R, cfg_.item_keys,
                                                                 desc_d_[b], first_, state_);
        step_kernel()<<<step_grid_, block_size(), sizeof(index_t) * R, stream>>>(
            desc_d_[b], first_, state_, R, cfg_.item_keys, views_d_, defaults::probing_length());
    }

    WC_H static auto step_kernel() noexcept {
        return &kernels::serve_step<view_type, detail::refill_keys<unsigned(CG)>(),
                                    detail::refill_insert_keys<unsigned(CG)>(), detail::lockstep_keys<unsigned(CG)>()>;
    }

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
    void launch_direct(const desc_type& d) {
        table_type& t = *tables_[d.table];
        const index_t pl = defaults::probing_length();
        auto* st = reinterpret_cast<Status*>(d.status_out);
        switch (Op(d.op)) {
        case Op::insert:
            if (st) t.template insert<status_handlers::ReturnStatus>(d.keys, d.values, d.n, stream_, pl, st);
            else t.insert(d.keys, d.values, d.n, stream_, pl);
            break;
        case Op::retrieve:
            if (st) t.template retrieve<status_handlers::ReturnStatus>(d.keys, d.n, d.values_out, stream_, pl, st);
            else t.retrieve(d.keys, d.n, d.values_out, stream_, pl);
            break;
        case Op::erase:
            if (st) t.template erase<status_handlers::ReturnStatus>(d.keys, d.n, stream_, pl, st);
            else t.erase(d.keys, d.n, stream_, pl);
            break;
        }
    }

    void launch(int b) {
        // the step reads its request count and descriptors from the pinned
        // staging buffer b, written by this thread before the launch
        std::atomic_thread_fence(std::memory_order_release);
        if (!cfg_.use_graph) {
            enqueue_step(b, stream_);
            return;
        }
        if (!graphs_[b].exec && !build_graph(b)) {
            cfg_.use_graph = false;
            enqueue_step(b, stream_);
            return;
        }
        (void)hipGraphLaunch(graphs_[b].exec, stream_);
    }

    // Keys per work item of a step: the smallest power of two from 8 that
    // still gives every group about four items (at most cfg_.item_keys).
    // A group takes ceil(items / groups) items, so small items keep the
    // busiest group close to the average and a small step ends sooner;
    // large steps keep large items, whose per-item setup is amortised.
    index_t step_item_keys(const StepList& step) const noexcept {
        index_t keys = 0;
        for (const auto& c : step) keys += c.keys;
        const index_t groups = index_t(step_grid_) * block_size() / CG;
        index_t ik = 8;
        while (ik < cfg_.item_keys && ik * groups * 4 < keys) ik *= 2;
        return ik < cfg_.item_keys ? ik : cfg_.item_keys;
    }

    struct InFlight {
        hipEvent_t ev;
        std::uint64_t step_no;
        StepList chunks;
    };

    void retire(std::deque<InFlight>& inflight) {
        const auto now = clock::now();
        {
            std::lock_guard<std::mutex> g(mu_);
            for (auto& c : inflight.front().chunks) {
                if (!c.last) continue;
                c.r->done = now;
                c.r->step = inflight.front().step_no;
                c.r->state.store(2);
                stats_.requests++;
                stats_.keys += c.r->d.n;
            }
        }
        done_cv_.notify_all();
        (void)hipEventDestroy(inflight.front().ev);
        inflight.pop_front();
    }

    void loop() {
        std::deque<InFlight> inflight;
        int b = 0;
        std::uint64_t step_no = 0;
        while (true) {
            StepList step;
            bool direct = false;
            {
                std::unique_lock<std::mutex> g(mu_);
                if (inflight.empty()) cv_.wait(g, [&] { return stop_ || !pending_.empty(); });
                if (stop_ && pending_.empty() && inflight.empty()) break;
                // at most two steps in flight: descriptors are double-buffered
                if (inflight.size() < 2) pick(step, desc_h_[b], direct);
                for (auto& c : step) c.r->state.store(1);
                if (!step.empty()) stats_.steps++;
                if (direct) stats_.direct_steps++;
            }
            if (!step.empty()) {
                ++step_no;
                if (direct) {
                    launch_direct(desc_h_[b][0]);
                } else {
                    num_reqs_h_[b][0] = std::uint32_t(step.size());
                    num_reqs_h_[b][1] = std::uint32_t(step_item_keys(step));
                    launch(b);
                }
                hipEvent_t ev = nullptr;
                (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                (void)hipEventRecord(ev, stream_);
                inflight.push_back(InFlight{ev, step_no, std::move(step)});
                b ^= 1;
            }
            // retire finished steps; with a free staging buffer, go back and
            // build the next step as soon as requests are pending (spinning:
            // the scheduler thread owns its core)
            while (!inflight.empty()) {
                if (hipEventQuery(inflight.front().ev) == hipSuccess) {
                    retire(inflight);
                    continue;
                }
                if (inflight.size() < 2) {   // poll: a new request or the step's end, whichever comes first
                    if (waiting_.load(std::memory_order_acquire) != 0 || stop_flag_.load()) break;
                    std::this_thread::yield();
                    continue;
                }
                (void)hipEventSynchronize(inflight.front().ev);
                retire(inflight);
            }
        }
    }

    Config cfg_;
    bool ok_ = true;
    std::vector<std::unique_ptr<table_type>> tables_;
    std::vector<index_t> budget_, reserved_;
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
    std::vector<std::shared_ptr<Request>> all_;
    index_t next_id_ = 0;
    bool stop_ = false;
    std::atomic<bool> stop_flag_{false};   // stop_, readable without mu_
    std::atomic<std::size_t> waiting_{0};  // pending_.size(), readable without mu_
    Stats stats_;
    std::thread worker_;
};

}  // namespace serving
}  // namespace warpcore
