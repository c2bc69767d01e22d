// TableBatch: many small single-value hash tables in one slot arena, all
// inserted into or looked up in one persistent launch, with the probe
// geometry of every table chosen by an autotuner.
//
// No reference counterpart: warpcore builds one table per object, and its
// capacity rule (first prime >= 131072 windows) makes every table at least
// 131101 * cg slots. A batch of thousands of small tables has its own rule:
// table t, probed in windows of W_t slots, gets p_t = the smallest prime
// >= max(2, SDIV(min_t, W_t)) windows, so a 64-slot request is a few
// windows, not a megabyte. Each table is otherwise a SingleValueHashTable
// <Key, Value> with double hashing over aligned windows of W_t slots: the
// same hashes, insert protocol, statuses and per-table status word, so
// per-key results equal those of a table of that geometry fed the same keys
// (only slot placement differs).
//
// Configurations. A table's configuration is (lanes, window): groups of
// `lanes` lanes read windows of `window` slots, one slot per lane or, for
// 8-byte slots, two adjacent slots per lane with one 16-byte load
// (owner_probing.hpp). 8-byte slots: (4,4) (4,8) (8,8) (8,16) (16,16);
// 16-byte slots: (4,4) (8,8) (16,16).
//
// A bulk call takes the keys of all tables concatenated, table t owning
// keys [offsets[t], offsets[t+1]). The call cuts every table's range into
// work items of up to `item_keys` keys and lists them on the device (count,
// scan, fill) in the batch's table order (capacity descending, so the
// smallest items come last); one persistent kernel lets each wavefront take
// items from an atomic counter (work stealing) and run the lane-owner code
// of the item's configuration on its table, 64 keys in flight per
// wavefront. Launch parameters (keys per item, blocks per CU, block size)
// come from a BatchLaunch. When every table has the same configuration the
// launch uses a kernel compiled for that configuration alone (registers of
// that configuration only).
//
// Autotuner (tune()). Tables fall into classes by requested capacity
// (bands [4^k, 4^(k+1)) slots) and fill (keys over requested capacity:
// below 0.6, 0.6-0.7, 0.7-0.8, from 0.8); the key width is the batch's.
// For every configuration the tuner lays all tables out with it and times
// each class alone: insert, then retrieve, of the caller's workload
// restricted to the class's tables. Each class keeps its fastest
// configuration among those whose slot total for the class is within 10%
// of the smallest (a wide window rounds a tiny table up to more slots).
// With the chosen configurations laid out, the tuner times the launch
// candidates on the whole workload and keeps the fastest. Tuning leaves the
// batch empty.
#pragma once

#include "base.hpp"
#include "capacity.hpp"
#include "defaults.hpp"
#include "engine.hpp"
#include "owner_probing.hpp"
#include "scan.hpp"
#include "single_value_hash_table.hpp"
#include "status.hpp"
#include "storage.hpp"

#include <algorithm>
#include <numeric>
#include <type_traits>
#include <vector>

namespace warpcore {

namespace detail {

// Per-table geometry in device memory.
struct BatchTableDesc {
    index_t slot_offset;   // a multiple of the table's window
    probing_schemes::Geometry geo;
    std::uint32_t config;
};

// (lanes, window) choices by slot size
template<std::size_t SlotBytes>
struct BatchConfigs;
template<>
struct BatchConfigs<8> {
    static constexpr int count = 5;
    static constexpr unsigned lanes[count] = {4, 4, 8, 8, 16};
    static constexpr unsigned window[count] = {4, 8, 8, 16, 16};
};
template<>
struct BatchConfigs<16> {
    static constexpr int count = 3;
    static constexpr unsigned lanes[count] = {4, 8, 16};
    static constexpr unsigned window[count] = {4, 8, 16};
};

// small-table capacity rule (see the file comment)
WC_H index_t batch_table_windows(index_t min_capacity, index_t window) noexcept {
    const index_t w = SDIV(min_capacity, window);
    return next_prime_at_least(w < 2 ? 2 : w);
}

// Tables of at most this many bytes of slots are inserted into and looked
// up in LDS (batch_lds_insert / batch_lds_retrieve), one block per CU beside
// the work-item kernel's blocks (64, 96, 128 and 160 KiB measured on the
// batched configuration: 36.2, 36.8, 38.8 and 37.5 Gops/s).
constexpr index_t kBatchLdsBytes = 128 * 1024;

struct BatchLaunch {
    index_t item_keys;   // keys per work item
    int blocks_per_cu;   // persistent grid: CUs * blocks_per_cu blocks
    int block_size;      // threads per block: 64..512, a multiple of 64
};

// Tuning class of a table: capacity band * 4 + fill band.
WC_H std::uint32_t batch_tune_class(index_t min_capacity, index_t keys) noexcept {
    const unsigned lg = min_capacity ? unsigned(63 - __builtin_clzll(min_capacity)) : 0u;
    const unsigned band = lg / 2 < 15 ? lg / 2 : 15u;
    const double fill = min_capacity ? double(keys) / double(min_capacity) : 1.0;
    const unsigned f = fill < 0.6 ? 0u : fill < 0.7 ? 1u : fill < 0.8 ? 2u : 3u;
    return band * 4 + f;
}

// One class of a tuning run.
struct BatchTuneClass {
    std::uint32_t id;   // capacity band * 4 + fill band
    index_t tables, keys, queries;
    int fixed_config, chosen_config;
    double fixed_ms, chosen_ms;   // the class alone: insert + retrieve
    index_t fixed_slots, chosen_slots;
};

struct BatchTuneReport {
    std::vector<BatchTuneClass> classes;
    double fixed_ms = 0.0;   // whole workload: every table at the fixed configuration, default launch
    double tuned_ms = 0.0;   // whole workload: chosen configurations and launch
    BatchLaunch launch{};
};

WC_D index_t wave_first(index_t v) noexcept {
    const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(v >> 32));
    return (index_t(hi) << 32) | lo;
}

template<int I, int N, class F>
WC_D void batch_dispatch(unsigned config, F& f) {
    if constexpr (I < N) {
        if (config == unsigned(I)) f(std::integral_constant<int, I>{});
        else batch_dispatch<I + 1, N>(config, f);
    }
}

}  // namespace detail

namespace kernels {

// items[i] = SDIV(ends[t] - begins[t], item_keys) for t = order[i], or 0
// from position `limit` on (tables handled elsewhere)
template<class T>
__global__ __launch_bounds__(256) void batch_count_items(const T* __restrict__ begins, const T* __restrict__ ends,
                                                         const std::uint32_t* __restrict__ order, index_t n,
                                                         T item_keys, T* __restrict__ items, index_t limit) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const std::uint32_t t = order[i];
        items[i] = i < limit ? SDIV(ends[t] - begins[t], item_keys) : T(0);
    }
}

// The items of position i (table t = order[i]) from first[i] on: table t
// and the first key of each. first = exclusive scan of items.
template<class T>
__global__ __launch_bounds__(256) void batch_fill_items(const T* __restrict__ first, const T* __restrict__ begins,
                                                        const T* __restrict__ ends,
                                                        const std::uint32_t* __restrict__ order, index_t n,
                                                        T item_keys, std::uint32_t* __restrict__ item_table,
                                                        T* __restrict__ item_begin, index_t limit) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < limit && i < n; i += stride) {
        const std::uint32_t t = order[i];
        const T b = begins[t];
        const T m = SDIV(ends[t] - b, item_keys);
        for (T j = 0; j < m; ++j) {
            item_table[first[i] + j] = t;
            item_begin[first[i] + j] = b + j * item_keys;
        }
    }
}

// Inserts into the batch's smallest tables (positions [first, n) of the
// table order, each table at most kBatchLdsBytes of slots), one block per
// table at a time: the block copies the table's slots into LDS, inserts the
// table's keys there with the per-key protocol of the global kernels
// (windows in probe order; the key found in a window is a duplicate whose
// value is overwritten, else the window's lowest EMPTY slot is claimed by a
// compare-and-swap, and a lost swap rescans the window; a window with
// neither sends the key on, up to the probing budget) and writes the slots
// back. Statuses and the per-table status words are those of the global
// kernels; only slot placement may differ.
template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class StatusHandler>
__global__ __launch_bounds__(1024) void batch_lds_insert(
    const Key* __restrict__ keys_in, const Value* __restrict__ values_in, const index_t* __restrict__ begins,
    const index_t* __restrict__ ends, const std::uint32_t* __restrict__ order, index_t first, index_t n,
    const detail::BatchTableDesc* __restrict__ tables, storage::key_value::pair_t<Key, Value>* slots,
    Status* __restrict__ statuses, Key seed, index_t probing_length,
    typename StatusHandler::base_type* __restrict__ status_out) {
    using Slot = storage::key_value::pair_t<Key, Value>;
    using Configs = detail::BatchConfigs<sizeof(Slot)>;
    constexpr bool packed = sizeof(Slot) == 8;
    static_assert(sizeof(Slot) == 8 || sizeof(Slot) == 16, "8- or 16-byte slots");
    __shared__ Slot lslots[detail::kBatchLdsBytes / sizeof(Slot)];
    for (index_t i = first + blockIdx.x; i < n; i += gridDim.x) {
        const std::uint32_t t = order[i];
        const detail::BatchTableDesc d = tables[t];
        const index_t cap = d.geo.capacity;
        const index_t b = begins[t], e = ends[t];
        if (b >= e) continue;   // uniform over the block
        Slot* g = slots + d.slot_offset;
        for (index_t s = threadIdx.x; s < cap; s += blockDim.x) lslots[s] = g[s];
        __syncthreads();
        detail::DeferredJoin join;
        auto run = [&](auto ic) {
            constexpr unsigned WIN = Configs::window[decltype(ic)::value];
            using Cursor = detail::WindowCursor<defaults::probing_scheme_t<Key, WIN>>;
            const index_t budget = probing_schemes::detail::window_budget(probing_length, WIN, d.geo.windows);
            for (index_t k = b + threadIdx.x; k < e; k += blockDim.x) {
                const Key key = keys_in[k];
                const Value val = values_in[k];
                Status st = Status::probing_length_exceeded();
                if (key == EmptyKey || key == TombstoneKey) {
                    st = Status::invalid_key();
                } else {
                    Cursor c;
                    c.begin(key, seed, d.geo);
                    index_t w = c.w;
                    for (index_t r = 0; r < budget; ++r) {
                        Slot* win = lslots + w * WIN;
                        bool done = false, next = false;
                        while (!done && !next) {
                            // the window's keys, all read before any is compared (other
                            // threads' swaps change them: read again on every pass)
                            __asm__ __volatile__("" ::: "memory");
                            Key wk[WIN];
                            std::uint64_t wword[packed ? WIN : 1];
#pragma unroll
                            for (unsigned j = 0; j < WIN; ++j) {
                                if constexpr (packed) {
                                    wword[j] = reinterpret_cast<const std::uint64_t*>(win)[j];
                                    __builtin_memcpy(&wk[j], &wword[j], sizeof(Key));
                                } else {
                                    wk[j] = win[j].key;
                                }
                            }
                            int empty_at = -1, found_at = -1;
#pragma unroll
                            for (unsigned j = 0; j < WIN; ++j) {
                                if (found_at < 0 && wk[j] == key) found_at = int(j);
                                if (empty_at < 0 && wk[j] == EmptyKey) empty_at = int(j);
                            }
                            if (found_at >= 0) {   // the key is stored: its value is overwritten
                                win[found_at].value = val;
                                st = Status::duplicate_key();
                                done = true;
                            } else if (empty_at < 0) {
                                next = true;
                            } else if constexpr (packed) {
                                Slot want;
                                want.key = key;
                                want.value = val;
                                std::uint64_t desired, expected = 0;
                                __builtin_memcpy(&desired, &want, 8);
#pragma unroll
                                for (unsigned j = 0; j < WIN; ++j)
                                    if (int(j) == empty_at) expected = wword[j];
                                if (__hip_atomic_compare_exchange_strong(
                                        reinterpret_cast<std::uint64_t*>(win + empty_at), &expected, desired,
                                        __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                                    st = Status::none();
                                    done = true;
                                }
                            } else {
                                Key expected = EmptyKey;
                                if (__hip_atomic_compare_exchange_strong(&win[empty_at].key, &expected, key,
                                                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                         __HIP_MEMORY_SCOPE_WORKGROUP)) {
                                    win[empty_at].value = val;
                                    st = Status::none();
                                    done = true;
                                }
                            }
                        }
                        if (done) break;
                        w += c.step;
                        if (w >= d.geo.windows) w -= d.geo.windows;
                    }
                }
                if (st.has_any()) join(st);
                StatusHandler::handle(st, status_out, k);
            }
        };
        detail::batch_dispatch<0, Configs::count>(d.config, run);
        join.flush(statuses + t);
        __syncthreads();
        for (index_t s = threadIdx.x; s < cap; s += blockDim.x) g[s] = lslots[s];
        __syncthreads();
    }
}

// Lookups in the same small tables, one block per table: slots copied into
// LDS, each key's windows read there in probe order (a hit writes the value;
// a window with an EMPTY slot ends the key as missing, its output left
// alone; the probing budget as in the global kernels). Statuses and status
// words are those of the global lookups.
template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class StatusHandler>
__global__ __launch_bounds__(1024) void batch_lds_retrieve(
    const Key* __restrict__ keys_in, Value* __restrict__ values_out, const index_t* __restrict__ begins,
    const index_t* __restrict__ ends, const std::uint32_t* __restrict__ order, index_t first, index_t n,
    const detail::BatchTableDesc* __restrict__ tables, const storage::key_value::pair_t<Key, Value>* slots,
    Status* __restrict__ statuses, Key seed, index_t probing_length,
    typename StatusHandler::base_type* __restrict__ status_out) {
    using Slot = storage::key_value::pair_t<Key, Value>;
    using Configs = detail::BatchConfigs<sizeof(Slot)>;
    __shared__ Slot lslots[detail::kBatchLdsBytes / sizeof(Slot)];
    for (index_t i = first + blockIdx.x; i < n; i += gridDim.x) {
        const std::uint32_t t = order[i];
        const detail::BatchTableDesc d = tables[t];
        const index_t cap = d.geo.capacity;
        const index_t b = begins[t], e = ends[t];
        if (b >= e) continue;   // uniform over the block
        const Slot* g = slots + d.slot_offset;
        for (index_t s = threadIdx.x; s < cap; s += blockDim.x) lslots[s] = g[s];
        __syncthreads();
        detail::DeferredJoin join;
        auto run = [&](auto ic) {
            constexpr unsigned WIN = Configs::window[decltype(ic)::value];
            using Cursor = detail::WindowCursor<defaults::probing_scheme_t<Key, WIN>>;
            const index_t budget = probing_schemes::detail::window_budget(probing_length, WIN, d.geo.windows);
            for (index_t k = b + threadIdx.x; k < e; k += blockDim.x) {
                const Key key = keys_in[k];
                Status st = Status::probing_length_exceeded();
                if (key == EmptyKey || key == TombstoneKey) {
                    st = Status::invalid_key();
                } else {
                    Cursor c;
                    c.begin(key, seed, d.geo);
                    index_t w = c.w;
                    for (index_t r = 0; r < budget; ++r) {
                        const Slot* win = lslots + w * WIN;
                        Key wk[WIN];
#pragma unroll
                        for (unsigned j = 0; j < WIN; ++j) wk[j] = win[j].key;
                        int found_at = -1;
                        bool empty = false;
#pragma unroll
                        for (unsigned j = 0; j < WIN; ++j) {
                            if (found_at < 0 && wk[j] == key) found_at = int(j);
                            empty = empty || wk[j] == EmptyKey;
                        }
                        if (found_at >= 0) {
                            values_out[k] = win[found_at].value;
                            st = Status::none();
                            break;
                        }
                        if (empty) {
                            st = Status::key_not_found();
                            break;
                        }
                        w += c.step;
                        if (w >= d.geo.windows) w -= d.geo.windows;
                    }
                }
                if (st.has_any()) join(st);
                StatusHandler::handle(st, status_out, k);
            }
        };
        detail::batch_dispatch<0, Configs::count>(d.config, run);
        join.flush(statuses + t);
        __syncthreads();
    }
}

// Persistent lane-owner kernel over the work items: each wavefront takes
// items until none is left. Only >= 0: every table has that configuration.
template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class StatusHandler, bool IsInsert, int Only>
__global__ __launch_bounds__(512) void batch_owner(
    const Key* __restrict__ keys_in, const Value* __restrict__ values_in, Value* __restrict__ values_out,
    const index_t* __restrict__ ends, const index_t* __restrict__ items, const index_t* __restrict__ first_item,
    index_t num_tables, const std::uint32_t* __restrict__ item_table, const index_t* __restrict__ item_begin,
    index_t item_keys, index_t* __restrict__ counter, const detail::BatchTableDesc* __restrict__ tables,
    storage::key_value::pair_t<Key, Value>* slots, Status* __restrict__ statuses, Key seed, index_t probing_length,
    typename StatusHandler::base_type* __restrict__ status_out) {
    using Slot = storage::key_value::pair_t<Key, Value>;
    using Configs = detail::BatchConfigs<sizeof(Slot)>;
    __shared__ detail::OwnerEntry<Key, std::uint32_t> lds[512];
    auto* my = lds + (threadIdx.x & ~63u);
    const index_t total = num_tables ? first_item[num_tables - 1] + items[num_tables - 1] : 0;
    while (true) {
        index_t item = 0;
        if (lane_id() == 0) item = detail::atomic_add(counter, index_t(1));
        item = detail::wave_first(item);
        if (item >= total) break;
        const std::uint32_t t = item_table[item];
        const index_t begin = item_begin[item];
        const index_t tend = ends[t];
        const index_t end = tend - begin > item_keys ? begin + item_keys : tend;
        const detail::BatchTableDesc d = tables[t];
        auto run = [&](auto ic) {
            constexpr int I = decltype(ic)::value;
            constexpr unsigned WIN = Configs::window[I];
            constexpr unsigned SPL = WIN / Configs::lanes[I];
            using Scheme = defaults::probing_scheme_t<Key, WIN>;
            using View = detail::SingleValueView<Key, Value, EmptyKey, TombstoneKey, Scheme, Slot>;
            const View view{slots + d.slot_offset, statuses + t, d.geo, seed, true, nullptr};
            const index_t budget = probing_schemes::detail::window_budget(probing_length, WIN, d.geo.windows);
            detail::DeferredJoin join;
            if constexpr (IsInsert)
                detail::insert_owner_range<View, StatusHandler, std::uint32_t, true, SPL>(
                    keys_in, values_in, begin, end, view, budget, status_out, my, join);
            else
                detail::retrieve_owner_range<View, StatusHandler, std::uint32_t, SPL>(
                    keys_in, begin, end, values_out, view, budget, status_out, my, join);
            join.flush(statuses + t);
        };
        if constexpr (Only >= 0) run(std::integral_constant<int, Only>{});
        else detail::batch_dispatch<0, Configs::count>(d.config, run);
    }
}

}  // namespace kernels

template<class Key, class Value, Key EmptyKey = defaults::empty_key<Key>(),
         Key TombstoneKey = defaults::tombstone_key<Key>()>
class TableBatch {
    static_assert(checks::is_valid_key_type<Key>(), "invalid key type");

public:
    using key_type = Key;
    using value_type = Value;
    using index_type = index_t;
    using status_type = Status;
    using slot_type = storage::key_value::pair_t<Key, Value>;
    using configs = detail::BatchConfigs<sizeof(slot_type)>;
    using launch_type = detail::BatchLaunch;
    using tune_report_type = detail::BatchTuneReport;

    WC_H static constexpr int num_configs() noexcept { return configs::count; }
    WC_H static unsigned config_lanes(int c) noexcept { return configs::lanes[c]; }
    WC_H static unsigned config_window(int c) noexcept { return configs::window[c]; }
    // index of configuration (lanes, window), or -1
    WC_H static int config_index(unsigned lanes, unsigned window) noexcept {
        for (int c = 0; c < configs::count; ++c)
            if (configs::lanes[c] == lanes && configs::window[c] == window) return c;
        return -1;
    }
    WC_H static launch_type default_launch() noexcept { return launch_type{2048, 8, 256}; }

    // min_capacities: host array of num_tables requested capacities (slots);
    // every table starts with configuration `config`
    WC_H TableBatch(const index_type* min_capacities, index_type num_tables,
                    key_type seed = defaults::seed<key_type>(), int config = config_index(16, 16)) noexcept
        : num_tables_(num_tables), seed_(seed), min_caps_(min_capacities, min_capacities + num_tables),
          fixed_config_(config) {
        ok_ = num_tables > 0 && num_tables < (index_type(1) << 32) && config >= 0 && config < configs::count &&
              malloc_(&tables_, sizeof(detail::BatchTableDesc) * num_tables) &&
              malloc_(&statuses_, sizeof(Status) * num_tables) &&
              malloc_(&order_, sizeof(std::uint32_t) * num_tables) &&
              malloc_(&parts_, sizeof(index_type) * detail::kScanParts) && malloc_(&counter_, sizeof(index_type)) &&
              malloc_(&items_, sizeof(index_type) * num_tables) &&
              malloc_(&first_item_, sizeof(index_type) * num_tables);
        if (ok_) ok_ = set_config(config);
        // the LDS insert runs on a stream of its own beside the work items
        if (ok_)
            ok_ = hipStreamCreateWithFlags(&side_, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&fork_, hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&join_, hipEventDisableTiming) == hipSuccess;
    }
    TableBatch(const TableBatch&) = delete;
    TableBatch& operator=(const TableBatch&) = delete;
    WC_H ~TableBatch() noexcept {
        if (side_) (void)hipStreamSynchronize(side_);   // its LDS kernels read and write the arena
        for (void* p : {(void*)tables_, (void*)slots_, (void*)statuses_, (void*)order_, (void*)parts_,
                        (void*)counter_, (void*)items_, (void*)first_item_, (void*)item_table_, (void*)item_begin_})
            if (p) (void)hipFree(p);
        if (join_) (void)hipEventDestroy(join_);
        if (fork_) (void)hipEventDestroy(fork_);
        if (side_) (void)hipStreamDestroy(side_);
    }

    WC_H bool ok() const noexcept { return ok_; }
    WC_H index_type num_tables() const noexcept { return num_tables_; }
    WC_H index_type capacity(index_type t) const noexcept { return capacities_[t]; }
    WC_H int table_config(index_type t) const noexcept { return int(config_of_[t]); }
    WC_H int fixed_config() const noexcept { return fixed_config_; }
    WC_H index_type total_slots() const noexcept { return total_slots_; }
    WC_H index_type bytes_total() const noexcept {
        return total_slots_ * sizeof(slot_type) + num_tables_ * (sizeof(detail::BatchTableDesc) + sizeof(Status));
    }

    // Lays every table out with configuration c and empties the batch.
    WC_H bool set_config(int c) noexcept {
        if (c < 0 || c >= configs::count) return false;
        std::vector<std::uint8_t> all(num_tables_, std::uint8_t(c));
        return set_configs(all.data());
    }
    // Lays table t out with configuration per_table[t] (host array) and
    // empties the batch. False (batch unchanged) on an unknown configuration
    // or when the arena cannot grow.
    WC_H bool set_configs(const std::uint8_t* per_table) noexcept {
        for (index_type t = 0; t < num_tables_; ++t)
            if (per_table[t] >= configs::count) return false;
        std::vector<detail::BatchTableDesc> h(num_tables_);
        std::vector<index_type> caps(num_tables_);
        index_type slots = 0;
        for (index_type t = 0; t < num_tables_; ++t) {
            const index_type win = configs::window[per_table[t]];
            const index_type p = detail::batch_table_windows(min_caps_[t], win);
            if (p >= (index_type(1) << 31)) return false;   // window indices are 32-bit in the kernel
            slots = SDIV(slots, win) * win;
            h[t].slot_offset = slots;
            h[t].geo = probing_schemes::Geometry::make(p * win, win);
            h[t].config = per_table[t];
            caps[t] = p * win;
            slots += p * win;
        }
        if (side_) (void)hipStreamSynchronize(side_);   // LDS kernels of earlier calls use the arena
        if (slots > arena_slots_) {
            slot_type* grown = nullptr;
            if (!malloc_(&grown, sizeof(slot_type) * slots)) return false;
            if (slots_) (void)hipFree(slots_);
            slots_ = grown;
            arena_slots_ = slots;
        }
        std::vector<std::uint32_t> order(num_tables_);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(),
                         [&](std::uint32_t a, std::uint32_t b) { return caps[a] > caps[b]; });
        if (hipMemcpy(tables_, h.data(), sizeof(detail::BatchTableDesc) * num_tables_, hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(order_, order.data(), sizeof(std::uint32_t) * num_tables_, hipMemcpyHostToDevice) != hipSuccess)
            return false;
        // the order's suffix of tables small enough for the LDS insert
        lds_first_ = num_tables_;
        while (lds_first_ > 0 && caps[order[lds_first_ - 1]] * sizeof(slot_type) <= detail::kBatchLdsBytes)
            --lds_first_;
        capacities_ = std::move(caps);
        config_of_.assign(per_table, per_table + num_tables_);
        total_slots_ = slots;
        single_config_ = int(per_table[0]);
        for (index_type t = 1; t < num_tables_; ++t)
            if (per_table[t] != per_table[0]) single_config_ = -1;
        init();
        return hipStreamSynchronize(nullptr) == hipSuccess;
    }

    // every slot EMPTY, every status word none
    WC_H void init(hipStream_t stream = 0) noexcept {
        if (!slots_) return;
        storage::kernels::fill_pairs<<<detail::grid_for(total_slots_, 1, 8), block_size(), 0, stream>>>(
            slots_, total_slots_, EmptyKey, Value{});
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
        run<StatusHandler, true>(keys_in, values_in, nullptr, offsets, offsets + 1, num_in, stream, probing_length,
                                 status_out, launch ? *launch : launch_);
    }
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void retrieve(const key_type* keys_in, const index_type* offsets, index_type num_in,
                       value_type* values_out, hipStream_t stream = 0,
                       index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr,
                       const launch_type* launch = nullptr) noexcept {
        run<StatusHandler, false>(keys_in, nullptr, values_out, offsets, offsets + 1, num_in, stream, probing_length,
                                  status_out, launch ? *launch : launch_);
    }

    // Autotuner (see the file comment). keys/values/offsets: an insert
    // workload, queries/query_offsets (num_queries may be 0): a retrieve
    // workload, values_out: num_queries outputs. Returns 0 or a HIP error;
    // the batch is left empty with the chosen configurations and launch.
    WC_H int tune(const key_type* keys_in, const value_type* values_in, const index_type* offsets,
                  index_type num_in, const key_type* queries, const index_type* query_offsets,
                  index_type num_queries, value_type* values_out, hipStream_t stream,
                  tune_report_type* report = nullptr) noexcept;

    WC_H const launch_type& launch() const noexcept { return launch_; }
    // Inserts and lookups in tables of at most kBatchLdsBytes of slots run
    // in LDS (on by default); off: every table through the work items.
    WC_H void set_lds_insert(bool on) noexcept { lds_insert_ = on; }
    WC_H bool lds_insert() const noexcept { return lds_insert_; }
    // tables the LDS insert takes (the smallest ones)
    WC_H index_type lds_tables() const noexcept { return num_tables_ - lds_first_; }
    WC_H bool set_launch(const launch_type& l) noexcept {
        if (l.item_keys < 1 || l.blocks_per_cu < 1 || l.block_size < 64 || l.block_size > 512 ||
            l.block_size % 64 != 0)
            return false;
        launch_ = l;
        return true;
    }

    // per-table status words (device array of num_tables)
    WC_H const Status* status_words() const noexcept { return statuses_; }
    WC_H slot_type* slots() const noexcept { return slots_; }
    WC_H int last_error() const noexcept { return last_error_; }

private:
    template<class T>
    WC_H static bool malloc_(T** p, std::size_t bytes) noexcept {
        *p = nullptr;
        return hipMalloc(reinterpret_cast<void**>(p), bytes ? bytes : 1) == hipSuccess;
    }

    template<class StatusHandler, bool IsInsert>
    WC_H void run(const key_type* keys_in, const value_type* values_in, value_type* values_out,
                  const index_type* begins, const index_type* ends, index_type num_in, hipStream_t stream,
                  index_type probing_length, typename StatusHandler::base_type* status_out,
                  launch_type launch) noexcept {
        last_error_ = 0;
        if (!ok_ || num_in == 0) return;
        if (launch.item_keys < 1 || launch.blocks_per_cu < 1 || launch.block_size < 64 || launch.block_size > 512 ||
            launch.block_size % 64 != 0) {
            last_error_ = int(hipErrorInvalidValue);
            return;
        }
        const index_type ik = launch.item_keys;
        // items: at most SDIV(num_in, ik) + num_tables
        const index_type max_items = SDIV(num_in, ik) + num_tables_;
        if (max_items > item_capacity_) {
            for (void* p : {(void*)item_table_, (void*)item_begin_})
                if (p) (void)hipFree(p);
            item_table_ = nullptr;
            item_begin_ = nullptr;
            item_capacity_ = 0;
            if (!malloc_(&item_table_, sizeof(std::uint32_t) * max_items) ||
                !malloc_(&item_begin_, sizeof(index_type) * max_items)) {
                last_error_ = int(hipErrorOutOfMemory);
                return;
            }
            item_capacity_ = max_items;
        }
        // the smallest tables run in LDS; the work items cover the others
        const index_type limit = lds_insert_ ? lds_first_ : num_tables_;
        if (limit < num_tables_) {   // on the side stream, after what the caller's stream holds
            const index_type n = num_tables_ - limit;
            const index_type cap = index_type(detail::num_compute_units());
            const unsigned grid = unsigned(n < cap ? n : cap);
            (void)hipEventRecord(fork_, stream);
            (void)hipStreamWaitEvent(side_, fork_, 0);
            if constexpr (IsInsert)
                kernels::batch_lds_insert<key_type, value_type, EmptyKey, TombstoneKey, StatusHandler>
                    <<<grid, 1024, 0, side_>>>(keys_in, values_in, begins, ends, order_, limit, num_tables_, tables_,
                                               slots_, statuses_, seed_, probing_length, status_out);
            else
                kernels::batch_lds_retrieve<key_type, value_type, EmptyKey, TombstoneKey, StatusHandler>
                    <<<grid, 1024, 0, side_>>>(keys_in, values_out, begins, ends, order_, limit, num_tables_,
                                               tables_, slots_, statuses_, seed_, probing_length, status_out);
            (void)hipEventRecord(join_, side_);
        }
        kernels::batch_count_items<index_type><<<detail::grid_for(num_tables_, 1, 8), block_size(), 0, stream>>>(
            begins, ends, order_, num_tables_, ik, items_, limit);
        detail::device_scan(items_, first_item_, num_tables_, parts_, false, stream);
        kernels::batch_fill_items<index_type><<<detail::grid_for(num_tables_, 1, 8), block_size(), 0, stream>>>(
            first_item_, begins, ends, order_, num_tables_, ik, item_table_, item_begin_, limit);
        storage::kernels::fill<index_type><<<1, 64, 0, stream>>>(counter_, 1, index_type(0));
        const unsigned grid = unsigned(detail::num_compute_units() * launch.blocks_per_cu);
        launch_owner<StatusHandler, IsInsert, -1>(single_config_, grid, unsigned(launch.block_size), keys_in,
                                                  values_in, values_out, ends, ik, stream, probing_length,
                                                  status_out);
        if (limit < num_tables_) (void)hipStreamWaitEvent(stream, join_, 0);   // the caller's stream holds both
        last_error_ = int(hipGetLastError());
    }

    // kernel for a single configuration `only` (or all, -1)
    template<class StatusHandler, bool IsInsert, int I>
    WC_H void launch_owner(int only, unsigned grid, unsigned block, const key_type* keys_in,
                           const value_type* values_in, value_type* values_out, const index_type* ends,
                           index_type ik, hipStream_t stream, index_type probing_length,
                           typename StatusHandler::base_type* status_out) noexcept {
        if constexpr (I < configs::count) {
            if (only != I) {
                launch_owner<StatusHandler, IsInsert, I + 1>(only, grid, block, keys_in, values_in, values_out, ends,
                                                             ik, stream, probing_length, status_out);
                return;
            }
        }
        constexpr int Only = I < configs::count ? I : -1;
        kernels::batch_owner<key_type, value_type, EmptyKey, TombstoneKey, StatusHandler, IsInsert, Only>
            <<<grid, block, 0, stream>>>(keys_in, values_in, values_out, ends, items_, first_item_, num_tables_,
                                         item_table_, item_begin_, ik, counter_, tables_, slots_, statuses_, seed_,
                                         probing_length, status_out);
    }

    // insert + retrieve of the given ranges between two events; ms or < 0
    WC_H double time_pass(const key_type* keys_in, const value_type* values_in, const index_type* begins,
                          const index_type* ends, index_type num_in, const key_type* queries,
                          const index_type* qbegins, const index_type* qends, index_type num_queries,
                          value_type* values_out, hipStream_t stream, const launch_type& l, hipEvent_t e0,
                          hipEvent_t e1) noexcept {
        using SH = status_handlers::ReturnNothing;
        init(stream);
        (void)hipEventRecord(e0, stream);
        run<SH, true>(keys_in, values_in, nullptr, begins, ends, num_in, stream, defaults::probing_length(),
                      nullptr, l);
        if (last_error_ != 0) return -1.0;
        if (num_queries) {
            run<SH, false>(queries, nullptr, values_out, qbegins, qends, num_queries, stream,
                           defaults::probing_length(), nullptr, l);
            if (last_error_ != 0) return -1.0;
        }
        (void)hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return double(ms);
    }

    index_type num_tables_ = 0;
    key_type seed_;
    std::vector<index_type> min_caps_;
    std::vector<index_type> capacities_;
    std::vector<std::uint8_t> config_of_;
    int fixed_config_;   // the configuration the batch was created with
    int single_config_ = -1;
    index_type total_slots_ = 0;
    index_type arena_slots_ = 0;
    launch_type launch_ = default_launch();
    index_type lds_first_ = 0;   // order positions from here on: tables small enough for the LDS insert
    bool lds_insert_ = true;
    hipStream_t side_ = nullptr;
    hipEvent_t fork_ = nullptr, join_ = nullptr;
    bool ok_ = false;
    int last_error_ = 0;
    detail::BatchTableDesc* tables_ = nullptr;
    slot_type* slots_ = nullptr;
    Status* statuses_ = nullptr;
    std::uint32_t* order_ = nullptr;
    index_type* parts_ = nullptr;
    index_type* counter_ = nullptr;
    index_type* items_ = nullptr;
    index_type* first_item_ = nullptr;
    std::uint32_t* item_table_ = nullptr;
    index_type* item_begin_ = nullptr;
    index_type item_capacity_ = 0;
};

template<class Key, class Value, Key EmptyKey, Key TombstoneKey>
WC_H int TableBatch<Key, Value, EmptyKey, TombstoneKey>::tune(const key_type* keys_in, const value_type* values_in,
                                                             const index_type* offsets, index_type num_in,
                                                             const key_type* queries,
                                                             const index_type* query_offsets,
                                                             index_type num_queries, value_type* values_out,
                                                             hipStream_t stream, tune_report_type* report) noexcept {
    if (!ok_) return int(hipErrorInvalidValue);
    const index_type n = num_tables_;
    std::vector<index_type> off(n + 1), qoff(n + 1, 0);
    if (hipMemcpy(off.data(), offsets, sizeof(index_type) * (n + 1), hipMemcpyDeviceToHost) != hipSuccess)
        return int(hipGetLastError());
    if (num_queries &&
        hipMemcpy(qoff.data(), query_offsets, sizeof(index_type) * (n + 1), hipMemcpyDeviceToHost) != hipSuccess)
        return int(hipGetLastError());

    // classes, and per class the end of every table's range (tables of
    // other classes get empty ranges)
    std::vector<std::uint32_t> cls(n);
    std::vector<std::uint32_t> ids;
    for (index_type t = 0; t < n; ++t) {
        cls[t] = detail::batch_tune_class(min_caps_[t], off[t + 1] - off[t]);
        ids.push_back(cls[t]);
    }
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    const std::size_t nc = ids.size();
    std::vector<index_type> ends_h(2 * nc * n);
    for (std::size_t c = 0; c < nc; ++c)
        for (index_type t = 0; t < n; ++t) {
            const bool in = cls[t] == ids[c];
            ends_h[(2 * c) * n + t] = in ? off[t + 1] : off[t];
            ends_h[(2 * c + 1) * n + t] = in ? qoff[t + 1] : qoff[t];
        }
    index_type* ends_d = nullptr;
    if (!malloc_(&ends_d, sizeof(index_type) * ends_h.size())) return int(hipErrorOutOfMemory);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = int(hipMemcpy(ends_d, ends_h.data(), sizeof(index_type) * ends_h.size(), hipMemcpyHostToDevice));
    if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = int(hipErrorOutOfMemory);

    const int fixed = fixed_config_;
    const launch_type base = default_launch();
    auto best_of = [&](const index_type* ends, const index_type* qends, index_type nk, index_type nq,
                       const launch_type& l) {
        double best = -1.0;
        for (int rep = 0; rep < 3 && rc == 0; ++rep) {   // the first repetition warms up
            const double ms = time_pass(keys_in, values_in, offsets, ends, nk, queries, query_offsets, qends, nq,
                                        values_out, stream, l, e0, e1);
            if (ms < 0.0) rc = last_error_ ? last_error_ : int(hipErrorUnknown);
            else if (rep > 0 && (best < 0.0 || ms < best)) best = ms;
        }
        return best;
    };

    // per configuration: every class alone
    std::vector<double> ms(configs::count * nc, -1.0);
    std::vector<index_type> slots_of(configs::count * nc, 0);
    std::vector<index_type> keys_of(nc, 0), queries_of(nc, 0), tables_of(nc, 0);
    for (index_type t = 0; t < n; ++t) {
        const std::size_t c = std::size_t(std::lower_bound(ids.begin(), ids.end(), cls[t]) - ids.begin());
        keys_of[c] += off[t + 1] - off[t];
        queries_of[c] += qoff[t + 1] - qoff[t];
        tables_of[c] += 1;
    }
    double fixed_total = -1.0;
    for (int k = 0; k < configs::count && rc == 0; ++k) {
        if (!set_config(k)) {
            rc = int(hipErrorOutOfMemory);
            break;
        }
        for (index_type t = 0; t < n; ++t) {
            const std::size_t c = std::size_t(std::lower_bound(ids.begin(), ids.end(), cls[t]) - ids.begin());
            slots_of[k * nc + c] += capacities_[t];
        }
        for (std::size_t c = 0; c < nc && rc == 0; ++c)
            ms[k * nc + c] = best_of(ends_d + (2 * c) * n, ends_d + (2 * c + 1) * n, keys_of[c],
                                     num_queries ? queries_of[c] : 0, base);
        if (k == fixed && rc == 0)
            fixed_total = best_of(offsets + 1, query_offsets ? query_offsets + 1 : nullptr, num_in, num_queries, base);
    }

    // per class: the fastest configuration within 10% of the fewest slots
    std::vector<int> chosen(nc, fixed);
    for (std::size_t c = 0; c < nc && rc == 0; ++c) {
        index_type fewest = ~index_type(0);
        for (int k = 0; k < configs::count; ++k) fewest = std::min(fewest, slots_of[k * nc + c]);
        double best = -1.0;
        for (int k = 0; k < configs::count; ++k) {
            if (double(slots_of[k * nc + c]) > 1.1 * double(fewest)) continue;
            if (best < 0.0 || ms[k * nc + c] < best) {
                best = ms[k * nc + c];
                chosen[c] = k;
            }
        }
    }

    // chosen layout, then the launch candidates on the whole workload
    double tuned_total = -1.0;
    launch_type best_launch = base;
    if (rc == 0) {
        std::vector<std::uint8_t> per_table(n);
        for (index_type t = 0; t < n; ++t)
            per_table[t] = std::uint8_t(chosen[std::size_t(std::lower_bound(ids.begin(), ids.end(), cls[t]) -
                                                           ids.begin())]);
        if (!set_configs(per_table.data())) rc = int(hipErrorOutOfMemory);
    }
    const index_type item_choices[] = {512, 2048, 8192};
    const int block_choices[] = {128, 256, 512};
    const int waves_per_cu[] = {16, 32};
    for (index_type ik : item_choices)
        for (int bs : block_choices)
            for (int wpc : waves_per_cu) {
                if (rc != 0) break;
                const launch_type l{ik, wpc * 64 / bs, bs};
                const double t = best_of(offsets + 1, query_offsets ? query_offsets + 1 : nullptr, num_in,
                                         num_queries, l);
                if (rc == 0 && (tuned_total < 0.0 || t < tuned_total)) {
                    tuned_total = t;
                    best_launch = l;
                }
            }
    if (rc == 0) {
        launch_ = best_launch;
        init(stream);
        rc = int(hipStreamSynchronize(stream));
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(ends_d);
    if (rc == 0 && report) {
        report->classes.clear();
        for (std::size_t c = 0; c < nc; ++c) {
            detail::BatchTuneClass r;
            r.id = ids[c];
            r.tables = tables_of[c];
            r.keys = keys_of[c];
            r.queries = num_queries ? queries_of[c] : 0;
            r.fixed_config = fixed;
            r.chosen_config = chosen[c];
            r.fixed_ms = ms[fixed * nc + c];
            r.chosen_ms = ms[chosen[c] * nc + c];
            r.fixed_slots = slots_of[fixed * nc + c];
            r.chosen_slots = slots_of[chosen[c] * nc + c];
            report->classes.push_back(r);
        }
        report->fixed_ms = fixed_total;
        report->tuned_ms = tuned_total;
        report->launch = launch_;
    }
    return rc;
}

}  // namespace warpcore
