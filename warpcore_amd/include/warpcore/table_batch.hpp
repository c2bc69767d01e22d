// TableBatch: many small single-value hash tables in one slot arena, all
// inserted into or looked up in one persistent launch.
//
// No reference counterpart: warpcore builds one table per object, and its
// capacity rule (first prime >= 131072 windows) makes every table at least
// 131101 * cg slots. A batch of thousands of small tables therefore has its
// own rule: table t gets p_t = the smallest prime >= max(2, SDIV(min_t, cg))
// windows of cg slots, so a 64-slot request is a few windows, not a
// megabyte. Each table is otherwise a SingleValueHashTable<Key, Value> with
// double hashing over aligned windows: the same hashes, insert protocol,
// statuses and per-table status word, so per-key results equal those of a
// plain table fed the same keys (only slot placement differs).
//
// A bulk call takes the keys of all tables concatenated, table t owning
// keys [offsets[t], offsets[t+1]). The call cuts every table's range into
// work items of `item_keys` keys, lists them on the device (count, scan,
// fill), and one persistent kernel lets each cooperative group take items
// from an atomic counter (work stealing): small and large tables mix
// freely and no group idles while items remain. Inside an item the group
// runs the refill kernels' per-range code against that table's view.
// Launch parameters (keys in flight per group, item size, blocks per CU)
// come from a BatchLaunch; the C ABI's tuner times candidates on the
// caller's workload and keeps the fastest.
#pragma once

#include "base.hpp"
#include "capacity.hpp"
#include "defaults.hpp"
#include "engine.hpp"
#include "lockstep_probing.hpp"
#include "refill_probing.hpp"
#include "scan.hpp"
#include "single_value_hash_table.hpp"
#include "status.hpp"
#include "storage.hpp"

#include <vector>

namespace warpcore {

namespace detail {

// Per-table geometry in device memory.
struct BatchTableDesc {
    index_t slot_offset;
    probing_schemes::Geometry geo;
};

// small-table capacity rule (see the file comment)
WC_H index_t batch_table_windows(index_t min_capacity, index_t cg) noexcept {
    const index_t w = SDIV(min_capacity, cg);
    return next_prime_at_least(w < 2 ? 2 : w);
}

struct BatchLaunch {
    int keys_per_batch;   // keys in flight per group (U of the refill code): 4 or 8
    index_t item_keys;    // keys per work item
    int blocks_per_cu;    // persistent grid: CUs * blocks_per_cu blocks of 256
};

}  // namespace detail

namespace kernels {

// items[t] = SDIV(n_t, item_keys)
template<class T>
__global__ __launch_bounds__(256) void batch_count_items(const T* __restrict__ offsets, index_t num_tables,
                                                         index_t item_keys, T* __restrict__ items) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t t = index_t(blockIdx.x) * blockDim.x + threadIdx.x; t < num_tables; t += stride)
        items[t] = SDIV(offsets[t + 1] - offsets[t], item_keys);
}

// item_table[first_t + j] = t for j < items[t]; first = exclusive scan of items
template<class T>
__global__ __launch_bounds__(256) void batch_fill_items(const T* __restrict__ first, const T* __restrict__ offsets,
                                                        index_t num_tables, index_t item_keys,
                                                        std::uint32_t* __restrict__ item_table) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t t = index_t(blockIdx.x) * blockDim.x + threadIdx.x; t < num_tables; t += stride) {
        const index_t n = SDIV(offsets[t + 1] - offsets[t], item_keys);
        for (index_t j = 0; j < n; ++j) item_table[first[t] + j] = std::uint32_t(t);
    }
}

template<class View, class StatusHandler, int U, bool IsInsert>
__global__ __launch_bounds__(256) void batch_persistent(
    const typename View::key_type* __restrict__ keys_in, const typename View::value_type* __restrict__ values_in,
    typename View::value_type* __restrict__ values_out, const index_t* __restrict__ offsets, index_t num_tables,
    const index_t* __restrict__ first_item, const std::uint32_t* __restrict__ item_table, index_t item_keys,
    index_t* __restrict__ counter, const detail::BatchTableDesc* __restrict__ tables,
    typename View::slot_type* slots, Status* __restrict__ statuses, typename View::key_type seed,
    index_t probing_length, typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(View::cg_size());
    using W = std::uint32_t;
    using Stage = typename std::conditional<IsInsert,
                                            detail::RefillInsertEntry<typename View::key_type, W,
                                                                      typename View::value_type>,
                                            detail::RefillEntry<typename View::key_type, W>>::type;
    __shared__ Stage stage[256];
    auto* my = stage + (threadIdx.x & ~(CG - 1));
    const unsigned rank = lane_id() & (CG - 1);
    const index_t total = num_tables ? first_item[num_tables - 1] +
                                           SDIV(offsets[num_tables] - offsets[num_tables - 1], item_keys)
                                     : 0;
    while (true) {
        index_t item = 0;
        if (rank == 0) item = detail::atomic_add(counter, index_t(1));
        item = tile<CG>::shfl(item, 0);
        if (item >= total) break;
        const index_t t = item_table[item];
        const index_t j = item - first_item[t];
        const index_t begin = offsets[t] + j * item_keys;
        const index_t end_t = offsets[t + 1];
        const index_t end = begin + item_keys < end_t ? begin + item_keys : end_t;
        const detail::BatchTableDesc d = tables[t];
        const View view{slots + d.slot_offset, statuses + t, d.geo, seed, true};
        const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, d.geo.windows);
        detail::DeferredJoin join;
        if constexpr (IsInsert)
            detail::insert_refill_range<View, StatusHandler, U, W, true>(keys_in, values_in, begin, end, view, budget,
                                                                         status_out, my, join);
        else
            detail::retrieve_refill_range<View, StatusHandler, U, W>(keys_in, begin, end, values_out, view, budget,
                                                                     status_out, my, join);
        const status_base_t all = tile<CG>::reduce_or(join.acc);
        if (rank == 0 && all != 0) statuses[t].atomic_join(Status::from_base(all));
    }
}

}  // namespace kernels

template<class Key, class Value, index_t CGSize = 16, Key EmptyKey = defaults::empty_key<Key>(),
         Key TombstoneKey = defaults::tombstone_key<Key>()>
class TableBatch {
    static_assert(checks::is_valid_key_type<Key>(), "invalid key type");
    static_assert(checks::is_valid_cg_size(CGSize), "invalid cooperative group size");

public:
    using key_type = Key;
    using value_type = Value;
    using index_type = index_t;
    using status_type = Status;
    using probing_scheme_type = defaults::probing_scheme_t<Key, CGSize>;
    using slot_type = storage::key_value::pair_t<Key, Value>;
    using view_type = detail::SingleValueView<Key, Value, EmptyKey, TombstoneKey, probing_scheme_type, slot_type>;
    using launch_type = detail::BatchLaunch;

    WC_HD static constexpr index_type cg_size() noexcept { return CGSize; }

    // min_capacities: host array of num_tables requested capacities (slots)
    WC_H TableBatch(const index_type* min_capacities, index_type num_tables,
                    key_type seed = defaults::seed<key_type>()) noexcept
        : num_tables_(num_tables), seed_(seed) {
        std::vector<detail::BatchTableDesc> h(num_tables);
        index_type slots = 0;
        bool small = true;   // the persistent kernel keeps window indices in 32 bits
        for (index_type t = 0; t < num_tables; ++t) {
            const index_type p = detail::batch_table_windows(min_capacities[t], CGSize);
            small = small && p < (index_type(1) << 31);
            h[t].slot_offset = slots;
            h[t].geo = probing_schemes::Geometry::make(p * CGSize, CGSize);
            slots += p * CGSize;
        }
        total_slots_ = slots;
        capacities_.resize(num_tables);
        for (index_type t = 0; t < num_tables; ++t) capacities_[t] = h[t].geo.capacity;
        ok_ = small && num_tables > 0 && malloc_(&tables_, sizeof(detail::BatchTableDesc) * num_tables) &&
              malloc_(&slots_, sizeof(slot_type) * total_slots_) && malloc_(&statuses_, sizeof(Status) * num_tables) &&
              malloc_(&parts_, sizeof(index_type) * detail::kScanParts) && malloc_(&counter_, sizeof(index_type)) &&
              malloc_(&items_, sizeof(index_type) * num_tables) && malloc_(&first_item_, sizeof(index_type) * num_tables);
        if (ok_)
            ok_ = hipMemcpy(tables_, h.data(), sizeof(detail::BatchTableDesc) * num_tables, hipMemcpyHostToDevice) ==
                  hipSuccess;
        if (ok_) init();
    }
    TableBatch(const TableBatch&) = delete;
    TableBatch& operator=(const TableBatch&) = delete;
    WC_H ~TableBatch() noexcept {
        for (void* p : {(void*)tables_, (void*)slots_, (void*)statuses_, (void*)parts_, (void*)counter_,
                        (void*)items_, (void*)first_item_, (void*)item_table_})
            if (p) (void)hipFree(p);
    }

    WC_H bool ok() const noexcept { return ok_; }
    WC_H index_type num_tables() const noexcept { return num_tables_; }
    WC_H index_type capacity(index_type t) const noexcept { return capacities_[t]; }
    WC_H index_type total_slots() const noexcept { return total_slots_; }
    WC_H index_type bytes_total() const noexcept {
        return total_slots_ * sizeof(slot_type) + num_tables_ * (sizeof(detail::BatchTableDesc) + sizeof(Status));
    }

    // every slot EMPTY, every status word none
    WC_H void init(hipStream_t stream = 0) noexcept {
        if (!ok_) return;
        storage::kernels::fill_pairs<<<detail::grid_for(total_slots_, 1, 8), block_size(), 0, stream>>>(slots_, total_slots_,
                                                                                              EmptyKey, Value{});
        storage::kernels::fill<std::uint32_t><<<detail::grid_for(num_tables_, 1, 8), block_size(), 0, stream>>>(
            reinterpret_cast<std::uint32_t*>(statuses_), num_tables_, 0u);
    }

    // keys_in / values_in: all tables' pairs, table t at [offsets[t], offsets[t+1])
    // (offsets: device array of num_tables + 1); num_in = offsets[num_tables]
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert(const key_type* keys_in, const value_type* values_in, const index_type* offsets,
                     index_type num_in, hipStream_t stream = 0,
                     index_type probing_length = defaults::probing_length(),
                     typename StatusHandler::base_type* status_out = nullptr,
                     const launch_type* launch = nullptr) noexcept {
        run<StatusHandler, true>(keys_in, values_in, nullptr, offsets, num_in, stream, probing_length, status_out,
                                 launch ? *launch : default_launch());
    }
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void retrieve(const key_type* keys_in, const index_type* offsets, index_type num_in,
                       value_type* values_out, hipStream_t stream = 0,
                       index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr,
                       const launch_type* launch = nullptr) const noexcept {
        const_cast<TableBatch*>(this)->template run<StatusHandler, false>(
            keys_in, nullptr, values_out, offsets, num_in, stream, probing_length, status_out,
            launch ? *launch : default_launch());
    }

    // per-table status words (device array of num_tables)
    WC_H const Status* status_words() const noexcept { return statuses_; }
    WC_H slot_type* slots() const noexcept { return slots_; }
    WC_H int last_error() const noexcept { return last_error_; }

    WC_H static launch_type default_launch() noexcept { return launch_type{8, 64, 8}; }

private:
    template<class T>
    WC_H static bool malloc_(T** p, std::size_t bytes) noexcept {
        *p = nullptr;
        return hipMalloc(reinterpret_cast<void**>(p), bytes ? bytes : 1) == hipSuccess;
    }

    template<class StatusHandler, bool IsInsert>
    WC_H void run(const key_type* keys_in, const value_type* values_in, value_type* values_out,
                  const index_type* offsets, index_type num_in, hipStream_t stream, index_type probing_length,
                  typename StatusHandler::base_type* status_out, launch_type launch) noexcept {
        last_error_ = 0;
        if (!ok_ || num_in == 0) return;
        const index_type ik = launch.item_keys < 1 ? 1 : launch.item_keys;
        // items: at most SDIV(num_in, ik) + num_tables
        const index_type max_items = SDIV(num_in, ik) + num_tables_;
        if (max_items > item_capacity_) {
            if (item_table_) (void)hipFree(item_table_);
            item_table_ = nullptr;
            item_capacity_ = 0;
            if (!malloc_(&item_table_, sizeof(std::uint32_t) * max_items)) {
                last_error_ = int(hipErrorOutOfMemory);
                return;
            }
            item_capacity_ = max_items;
        }
        kernels::batch_count_items<index_type><<<detail::grid_for(num_tables_, 1, 8), block_size(), 0, stream>>>(
            offsets, num_tables_, ik, items_);
        detail::device_scan(items_, first_item_, num_tables_, parts_, false, stream);
        kernels::batch_fill_items<index_type><<<detail::grid_for(num_tables_, 1, 8), block_size(), 0, stream>>>(
            first_item_, offsets, num_tables_, ik, item_table_);
        storage::kernels::fill<index_type><<<1, 64, 0, stream>>>(counter_, 1, index_type(0));
        const unsigned grid = unsigned(detail::num_compute_units() * (launch.blocks_per_cu < 1 ? 1 : launch.blocks_per_cu));
        if (launch.keys_per_batch <= 4)
            launch_persistent<StatusHandler, (CGSize < 4 ? int(CGSize) : 4), IsInsert>(
                grid, keys_in, values_in, values_out, offsets, ik, stream, probing_length, status_out);
        else
            launch_persistent<StatusHandler, (CGSize < 8 ? int(CGSize) : 8), IsInsert>(
                grid, keys_in, values_in, values_out, offsets, ik, stream, probing_length, status_out);
        last_error_ = int(hipGetLastError());
    }

    template<class StatusHandler, int U, bool IsInsert>
    WC_H void launch_persistent(unsigned grid, const key_type* keys_in, const value_type* values_in,
                                value_type* values_out, const index_type* offsets, index_type ik, hipStream_t stream,
                                index_type probing_length, typename StatusHandler::base_type* status_out) noexcept {
        kernels::batch_persistent<view_type, StatusHandler, U, IsInsert><<<grid, block_size(), 0, stream>>>(
            keys_in, values_in, values_out, offsets, num_tables_, first_item_, item_table_, ik, counter_, tables_,
            slots_, statuses_, seed_, probing_length, status_out);
    }

    index_type num_tables_ = 0;
    key_type seed_;
    index_type total_slots_ = 0;
    std::vector<index_type> capacities_;
    bool ok_ = false;
    int last_error_ = 0;
    detail::BatchTableDesc* tables_ = nullptr;
    slot_type* slots_ = nullptr;
    Status* statuses_ = nullptr;
    index_type* parts_ = nullptr;
    index_type* counter_ = nullptr;
    index_type* items_ = nullptr;
    index_type* first_item_ = nullptr;
    std::uint32_t* item_table_ = nullptr;
    index_type item_capacity_ = 0;
};

}  // namespace warpcore
