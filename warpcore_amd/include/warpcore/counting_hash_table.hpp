// CountingHashTable: key -> number of occurrences (drop-in for
// include/warpcore/counting_hash_table.cuh). The key is found or claimed
// exactly as in SingleValueHashTable, then one lane adds 1 to the slot's
// counter with an atomicAdd while it is below max_count
// (counting_hash_table.cuh:141-172); lookups report min(count, max_count).
// A counter that wraps is detected from the atomic's own return value (the
// old count was the type's maximum), flagged index_overflow and clamped.
#pragma once

#include "scan.hpp"
#include "single_value_hash_table.hpp"

namespace warpcore {

namespace detail {

// AddValue: the value handed to finish_insert is the increment (a count of
// occurrences combined before the insert) instead of 1.
template<class Base, bool AddValue = false>
struct CountingView {
    using key_type = typename Base::key_type;
    using value_type = typename Base::value_type;
    using index_type = index_t;
    using slot_type = typename Base::slot_type;
    using probing_scheme_type = typename Base::probing_scheme_type;
    using out_type = value_type;
    using probe_slot_type = slot_type;
    static constexpr bool publishes_value = false;
    Base base;
    value_type max_count;

    WC_HD static constexpr index_type cg_size() noexcept { return Base::cg_size(); }
    WC_HD index_type capacity() const noexcept { return base.capacity(); }
    WC_HD Status* status_word() const noexcept { return base.status_word(); }
    WC_HD static constexpr bool is_valid_key(key_type k) noexcept { return Base::is_valid_key(k); }
    WC_D key_type key_at(index_type i) const noexcept { return base.key_at(i); }
    WC_D auto slot_at(index_type i) const noexcept { return base.slot_at(i); }

    // ---- hooks of the batched kernels (batched_probing.hpp)
    WC_HD slot_type* slot_ptr() const noexcept { return base.slots; }
    WC_HD static key_type key_of(const slot_type& s) noexcept { return s.key; }
    WC_HD static value_type value_of(const slot_type& s) noexcept { return s.value; }
    WC_HD static constexpr bool is_empty_key(key_type k) noexcept { return Base::is_empty_key(k); }
    WC_D value_type found(const value_type& c) const noexcept { return c > max_count ? max_count : c; }
    WC_D Status missing(value_type&) const noexcept { return Status::key_not_found(); }
    using cas_word = key_type;
    using insert_probe_type = key_type;
    WC_D static key_type load_insert_probe(const slot_type* p) noexcept { return p->key; }
    WC_HD static key_type probe_key(key_type k) noexcept { return k; }
    WC_HD static key_type empty_probe() noexcept { return Base::empty_key_value(); }
    WC_D cas_word cas_issue(slot_type* p, key_type seen, key_type key_in, const value_type&) const noexcept {
        return atomic_cas(&(p->key), seen, key_in);
    }
    WC_D int cas_outcome(cas_word old, key_type& seen, key_type key_in) const noexcept {
        return old == seen ? 1 : (old == key_in ? 2 : 0);
    }
    // counting_hash_table.cuh:154-170
    // With 64-bit counts and the default max_count (the largest count) the
    // check before the add and the wrap guard after it can only act after
    // 2^64 increments of one key, so the add is issued without reading the
    // count and without waiting for the old value.
    //
    // A combined increment (AddValue: `inc` occurrences of the key at once)
    // ends where `inc` single increments would: each of those adds one only
    // while the count is below max_count, so the count becomes
    // min(count + inc, max_count) and never wraps (no index_overflow, which
    // a single add of `inc` past the type's maximum would raise). All but the
    // first of those occurrences find the key stored, so the table's status
    // joins duplicate_key, as the reference does for each duplicate insert.
    template<class Join>
    WC_D void finish_insert(index_type slot, const value_type& inc, bool, Status& status, Join& join) const noexcept {
        if constexpr (AddValue)
            if (inc > value_type(1)) join(Status::duplicate_key());
        const value_type add = AddValue ? inc : value_type(1);
        value_type* v = &(base.slots[slot].value);
        if constexpr (sizeof(value_type) == 8) {
            if (max_count == std::numeric_limits<value_type>::max()) {
                __hip_atomic_fetch_add(v, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
        }
        if constexpr (AddValue) {
            value_type cur = atomic_load_relaxed(v);
            while (cur < max_count) {
                const value_type want = max_count - cur < add ? max_count : value_type(cur + add);
                const value_type prev = atomic_cas(v, cur, want);
                if (prev == cur) break;
                cur = prev;
            }
            return;
        }
        if (atomic_load_relaxed(v) < max_count) {
            const value_type old = atomic_add(v, add);
            if (old > std::numeric_limits<value_type>::max() - add) {   // the counter wrapped
                status += Status::index_overflow();
                join(Status::index_overflow());
                atomic_exch(v, max_count);
            }
        }
    }
    WC_HD bool ready() const noexcept { return base.initialized; }
    WC_HD key_type hash_seed() const noexcept { return base.seed; }
    WC_HD const probing_schemes::Geometry& geometry() const noexcept { return base.geo; }

    template<class Join>
    WC_D Status insert(key_type key_in, const tile<cg_size()>& group, index_type probing_length,
                       Join& join) const noexcept {
        Status status = Status::unknown_error();
        bool published = false;
        const index_type s = base.insert_impl(key_in, status, group, probing_length, join, nullptr, published);
        if (group.thread_rank() == 0 && s != ~index_type(0)) {
            value_type* v = &(base.slots[s].value);
            if (atomic_load_relaxed(v) < max_count) {
                const value_type old = atomic_add(v, value_type(1));
                if (old == std::numeric_limits<value_type>::max()) {
                    status += Status::index_overflow();
                    join(Status::index_overflow());
                    atomic_exch(v, max_count);
                }
            }
        }
        return status;
    }

    template<class Join>
    WC_D Status retrieve(key_type key_in, value_type& value_out, const tile<cg_size()>& group,
                         index_type probing_length, Join& join) const noexcept {
        value_type out{};
        const Status status = base.retrieve(key_in, out, group, probing_length, join);
        value_out = status.has_any() ? value_type(0) : (out > max_count ? max_count : out);
        return status;
    }

    template<class Join>
    WC_D Status erase(key_type key_in, const tile<cg_size()>& group, index_type probing_length,
                      Join& join) const noexcept {
        return base.erase(key_in, group, probing_length, join);
    }
};

// Combining counts before the insert (see CountingHashTable::insert): a
// whole-table process setting, 1 = on (default), 0 = every occurrence is
// inserted on its own.
inline std::atomic<int>& counting_combine_setting() noexcept {
    static std::atomic<int> m{1};
    return m;
}
WC_H int counting_combine() noexcept { return counting_combine_setting().load(std::memory_order_relaxed); }
WC_H void set_counting_combine(int m) noexcept { counting_combine_setting().store(m, std::memory_order_relaxed); }

constexpr unsigned kCombineTile = 1024;             // occurrences per block and tile
constexpr unsigned kCombineSlots = 2 * kCombineTile;  // LDS hash slots (load <= 1/2)
constexpr unsigned kCombineHot = 512;                 // block-persistent map of keys frequent in a tile
constexpr unsigned kCombineHotMin = 2;                // occurrences in a tile that make a key frequent
constexpr unsigned kCombineHotProbes = 8;             // slots a key tries in it before it is emitted per tile
constexpr unsigned kCombineBlocks = 1024;             // largest grid of count_combine
constexpr std::uint32_t kHotEntry = 0x80000000u;      // pos[i] flag: entry via hot_entry[pos & ~flag]

}  // namespace detail

namespace kernels {

// One tile of 1024 occurrences per block: identical keys are combined in an
// LDS hash map (CAS on the key, atomicAdd on the count, atomicMin on the
// first occurrence). A key that occurs kCombineHotMin times or more in the
// tile moves to a second LDS map that lives for all the block's tiles
// (kCombineHot slots, kCombineHotProbes tries), so a key that is frequent in the whole input
// becomes one entry per block instead of one per tile (the entries' count
// adds to the table are memory-side atomics, and those on one slot
// serialise). The other distinct keys of the tile are appended to keys_out
// / counts_out with rep_out = their first occurrence, and every occurrence
// gets pos[i] = its entry, or kHotEntry | its hot slot, whose entry the
// block writes to hot_entry when it flushes the second map at its end.
// Invalid keys (EMPTY, TOMBSTONE) get an entry each, so the insert reports
// them. Grid: at most kCombineBlocks blocks.
template<class K, class V, K EmptyKey, K TombstoneKey>
__global__ __launch_bounds__(256) void count_combine(const K* __restrict__ keys_in, index_t num_in,
                                                     K* __restrict__ keys_out, V* __restrict__ counts_out,
                                                     std::uint32_t* __restrict__ rep_out,
                                                     std::uint32_t* __restrict__ pos, index_t* __restrict__ num_out,
                                                     std::uint32_t* __restrict__ hot_entry) {
    constexpr unsigned T = detail::kCombineTile, S = detail::kCombineSlots, P = T / 256, Q = S / 256;
    constexpr unsigned H = detail::kCombineHot, QH = H / 256;
    static_assert(QH >= 1 && (H & (H - 1)) == 0, "hot map: a power of two of at least 256 slots");
    __shared__ K lkey[S];
    __shared__ std::uint32_t lcnt[S], lrep[S], lent[S];
    __shared__ K hkey[H];
    __shared__ std::uint32_t hcnt[H], hrep[H];
    __shared__ index_t base_sh;
    for (unsigned q = 0; q < QH; ++q) {
        const unsigned h = q * 256 + threadIdx.x;
        hkey[h] = EmptyKey;
        hcnt[h] = 0;
        hrep[h] = ~0u;
    }
    for (index_t tile = index_t(blockIdx.x) * T; tile < num_in; tile += index_t(gridDim.x) * T) {
        for (unsigned q = 0; q < Q; ++q) {
            const unsigned s = q * 256 + threadIdx.x;
            lkey[s] = EmptyKey;
            lcnt[s] = 0;
            lrep[s] = ~0u;
        }
        __syncthreads();
        K key[P];
        unsigned slot[P];   // LDS slot of the key, ~0u = invalid key, ~1u = past the end
        unsigned invalid = 0;
#pragma unroll
        for (unsigned j = 0; j < P; ++j) {
            const index_t i = tile + j * 256 + threadIdx.x;
            slot[j] = ~1u;
            if (i >= num_in) continue;
            key[j] = keys_in[i];
            if (key[j] == EmptyKey || key[j] == TombstoneKey) {
                slot[j] = ~0u;
                ++invalid;
                continue;
            }
            unsigned h = unsigned(hashers::MurmurHash<K>::hash(key[j])) & (S - 1);
            while (true) {
                const K prev = detail::atomic_cas(lkey + h, EmptyKey, key[j]);
                if (prev == EmptyKey || prev == key[j]) break;
                h = (h + 1) & (S - 1);
            }
            atomicAdd(lcnt + h, 1u);
            atomicMin(lrep + h, unsigned(j * 256 + threadIdx.x));
            slot[j] = h;
        }
        __syncthreads();
        // keys repeated in this tile join the block's hot map (their tile
        // count goes there; lcnt 0 marks the slot as moved)
        unsigned mine = invalid;   // entries of this thread: its un-moved LDS slots, then its invalid keys
        for (unsigned q = 0; q < Q; ++q) {
            const unsigned s = threadIdx.x * Q + q;
            const K k = lkey[s];
            if (k == EmptyKey) continue;
            bool moved = false;
            if (lcnt[s] >= detail::kCombineHotMin) {
                unsigned h = unsigned(hashers::MurmurHash<K>::hash(k)) & (H - 1);
                for (unsigned t = 0; t < detail::kCombineHotProbes; ++t) {
                    const K prev = detail::atomic_cas(hkey + h, EmptyKey, k);
                    if (prev == EmptyKey || prev == k) {
                        atomicAdd(hcnt + h, lcnt[s]);
                        atomicMin(hrep + h, std::uint32_t(tile + lrep[s]));
                        lent[s] = detail::kHotEntry | (std::uint32_t(blockIdx.x) * H + h);
                        lcnt[s] = 0;
                        moved = true;
                        break;
                    }
                    h = (h + 1) & (H - 1);
                }
            }
            mine += moved ? 0u : 1u;
        }
        index_t total;
        const index_t incl = detail::block_inclusive_scan(index_t(mine), total);
        if (threadIdx.x == 0) base_sh = detail::atomic_add(num_out, total);
        __syncthreads();
        index_t e = base_sh + incl - mine;
        for (unsigned q = 0; q < Q; ++q) {
            const unsigned s = threadIdx.x * Q + q;
            if (lkey[s] != EmptyKey && lcnt[s] != 0) {
                keys_out[e] = lkey[s];
                counts_out[e] = V(lcnt[s]);
                rep_out[e] = std::uint32_t(tile + lrep[s]);
                lent[s] = std::uint32_t(e);
                ++e;
            }
        }
#pragma unroll
        for (unsigned j = 0; j < P; ++j) {
            if (slot[j] != ~0u) continue;
            const index_t i = tile + j * 256 + threadIdx.x;
            keys_out[e] = key[j];
            counts_out[e] = V(1);
            rep_out[e] = std::uint32_t(i);
            pos[i] = std::uint32_t(e);
            ++e;
        }
        __syncthreads();
#pragma unroll
        for (unsigned j = 0; j < P; ++j)
            if (slot[j] < S) pos[tile + j * 256 + threadIdx.x] = lent[slot[j]];
        __syncthreads();
    }
    // flush the hot map: one entry per key, its slot's entry to hot_entry
    unsigned mine = 0;
    for (unsigned q = 0; q < QH; ++q) mine += hkey[threadIdx.x * QH + q] != EmptyKey;
    index_t total;
    const index_t incl = detail::block_inclusive_scan(index_t(mine), total);
    if (threadIdx.x == 0 && total != 0) base_sh = detail::atomic_add(num_out, total);
    __syncthreads();
    if (total == 0) return;
    index_t e = base_sh + incl - mine;
    for (unsigned q = 0; q < QH; ++q) {
        const unsigned h = threadIdx.x * QH + q;
        if (hkey[h] != EmptyKey) {
            keys_out[e] = hkey[h];
            counts_out[e] = V(hcnt[h]);
            rep_out[e] = hrep[h];
            hot_entry[std::uint32_t(blockIdx.x) * H + h] = std::uint32_t(e);
            ++e;
        }
    }
}

// Per occurrence: its entry's status for the first occurrence of a key;
// for the others duplicate_key where the entry was stored (none or
// duplicate), else the entry's status (the table was full, ...).
template<class StatusHandler>
__global__ __launch_bounds__(256) void count_expand(const std::uint32_t* __restrict__ pos,
                                                    const std::uint32_t* __restrict__ hot_entry,
                                                    const std::uint32_t* __restrict__ rep,
                                                    const Status* __restrict__ entry_status, index_t first,
                                                    index_t num_in,
                                                    typename StatusHandler::base_type* __restrict__ status_out) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num_in; i += stride) {
        const std::uint32_t p = pos[i];
        const std::uint32_t e = (p & detail::kHotEntry) ? hot_entry[p & ~detail::kHotEntry] : p;
        Status s = entry_status[e];
        if (rep[e] != std::uint32_t(i)) {
            const Status stored = s - Status::index_overflow();
            if (!stored.has_any(Status::flags_except(Status::duplicate_key()))) s = Status::duplicate_key() + (s - stored);
        }
        StatusHandler::handle(s, status_out, first + i);
    }
}

}  // namespace kernels

namespace detail {

}  // namespace detail

template<class Key, class Value = index_t, Key EmptyKey = defaults::empty_key<Key>(),
         Key TombstoneKey = defaults::tombstone_key<Key>(),
         class ProbingScheme = defaults::probing_scheme_t<Key, 4>,
         class TableStorage = defaults::table_storage_t<Key, Value>,
         index_t TempMemoryBytes = defaults::temp_memory_bytes()>
class CountingHashTable {
    using base_type = SingleValueHashTable<Key, Value, EmptyKey, TombstoneKey, ProbingScheme, TableStorage,
                                           TempMemoryBytes>;
    static_assert(checks::is_valid_counter_type<Value>(), "counter type must be either int, std::uint32_t or std::uint64_t");

public:
    using key_type = Key;
    using value_type = Value;
    using index_type = index_t;
    using status_type = Status;
    using view_type = detail::CountingView<typename base_type::view_type>;
    using add_view_type = detail::CountingView<typename base_type::view_type, true>;

    WC_HD static constexpr key_type empty_key() noexcept { return EmptyKey; }
    WC_HD static constexpr key_type tombstone_key() noexcept { return TombstoneKey; }
    WC_HD static constexpr index_type cg_size() noexcept { return base_type::cg_size(); }

    // counting_hash_table.cuh:74-85
    WC_H explicit CountingHashTable(index_type min_capacity, key_type seed = defaults::seed<key_type>(),
                                    value_type max_count = std::numeric_limits<value_type>::max(),
                                    bool no_init = false) noexcept
        : base_table_(min_capacity, seed, true), max_count_(max_count), is_copy_(false) {
        if (!no_init) init(seed);
    }
    WC_HD CountingHashTable(const CountingHashTable& o) noexcept
        : base_table_(o.base_table_), max_count_(o.max_count_), is_copy_(true), scratch_(o.scratch_) {}
    WC_H CountingHashTable(CountingHashTable&& o) noexcept
        : base_table_(std::move(o.base_table_)), max_count_(o.max_count_), is_copy_(o.is_copy_),
          scratch_(o.scratch_) {
        o.is_copy_ = true;
    }
    WC_HD ~CountingHashTable() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_) scratch_.release();
#endif
    }

    // counting_hash_table.cuh:113-124; the EMPTY fill also zeroes the counters
    WC_H void init(key_type seed, hipStream_t stream = 0) noexcept { base_table_.init(seed, stream); }
    WC_H void init(hipStream_t stream = 0) noexcept { init(base_table_.seed(), stream); }

    WC_D status_type insert(key_type key_in, const tile<cg_size()>& group,
                            index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{base_table_.status_};
        return view().insert(key_in, group, probing_length, join);
    }
    WC_D status_type retrieve(key_type key_in, value_type& value_out, const tile<cg_size()>& group,
                              index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{base_table_.status_};
        return view().retrieve(key_in, value_out, group, probing_length, join);
    }
    WC_D status_type erase(key_type key_in, const tile<cg_size()>& group,
                           index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{base_table_.status_};
        return view().erase(key_in, group, probing_length, join);
    }

    // counting_hash_table.cuh:182-194. Batches of at least 4096 occurrences
    // are first combined (count_combine: one insert of n per distinct key and
    // tile of 1024 occurrences), so a hot key costs one atomicAdd per tile
    // instead of one per occurrence; counts, statuses (one `none` per new
    // key, duplicate_key for the other occurrences) and the table status are
    // those of per-occurrence inserts.
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert(const key_type* keys_in, index_type num_in, hipStream_t stream = 0,
                     index_type probing_length = defaults::probing_length(),
                     typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!base_table_.is_initialized() || num_in == 0) return;
        if (detail::counting_combine() == 1 && num_in >= 4096 && num_in < (index_type(1) << 32)) {
            insert_combined<StatusHandler>(keys_in, num_in, stream, probing_length, status_out);
            return;
        }
        insert_each<view_type, StatusHandler, false>(keys_in, nullptr, num_in, view(), stream, probing_length,
                                                     status_out);
    }

    // every occurrence inserted on its own (the reference's kernel shape)
    template<class View, class StatusHandler, bool HasValues>
    WC_H void insert_each(const key_type* keys_in, const value_type* values_in, index_type num_in, const View& v,
                          hipStream_t stream, index_type probing_length,
                          typename StatusHandler::base_type* status_out) noexcept {
        if constexpr (detail::LaneWindow<typename view_type::slot_type, cg_size()>::usable) {
            if (capacity() * sizeof(typename view_type::slot_type) <= detail::lane_insert_table_bytes()) {
                kernels::insert_lanes<View, StatusHandler, detail::keys_per_lane(), HasValues>
                    <<<detail::grid_for_lanes(num_in), block_size(), 0, stream>>>(keys_in, values_in, num_in, v,
                                                                                 probing_length, status_out);
                return;
            }
        }
        if constexpr (detail::fixed_step_scheme<typename View::probing_scheme_type>::value) {
            if (detail::refill_inserts()) {
                detail::launch_insert_refill<View, StatusHandler, detail::refill_insert_keys<unsigned(cg_size())>(), HasValues>(
                    keys_in, values_in, num_in, v, probing_length, status_out, stream);
                return;
            }
            if (detail::lockstep_inserts()) {
                detail::launch_insert_lockstep<View, StatusHandler,
                                               detail::lockstep_insert_keys<unsigned(cg_size())>(), HasValues>(
                    keys_in, values_in, num_in, v, probing_length, status_out, stream);
                return;
            }
        }
        kernels::insert_batched<View, StatusHandler, detail::keys_per_group(), HasValues>
            <<<detail::grid_for_batched(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, values_in, num_in, v, probing_length, status_out);
    }

    // Buffers of one combining pass over up to `capacity` occurrences, carved
    // from one device buffer: distinct keys, their counts, first occurrences,
    // each occurrence's entry, the entries' statuses and the entry count.
    struct CombineBuffers {
        key_type* keys = nullptr;
        value_type* counts = nullptr;
        std::uint32_t* rep = nullptr;
        std::uint32_t* pos = nullptr;
        Status* status = nullptr;
        std::uint32_t* hot = nullptr;   // entry of each block's hot-map slot
        index_type* num = nullptr;
        index_type capacity = 0;

        WC_HD static constexpr index_type align(index_type b) noexcept { return (b + 255) / 256 * 256; }
        WC_HD static constexpr index_type hot_bytes() noexcept {
            return align(sizeof(std::uint32_t) * detail::kCombineBlocks * detail::kCombineHot);
        }
        // bytes of a buffer for `n` occurrences
        WC_HD static constexpr index_type bytes_for(index_type n) noexcept {
            return align(sizeof(key_type) * n) + align(sizeof(value_type) * n) + 3 * align(sizeof(std::uint32_t) * n) +
                   hot_bytes() + align(sizeof(index_type));
        }
        // the buffers inside `base` (256-byte aligned, bytes_for(n) bytes)
        WC_HD static CombineBuffers carve(void* base, index_type n) noexcept {
            auto* p = static_cast<unsigned char*>(base);
            CombineBuffers b;
            b.keys = reinterpret_cast<key_type*>(p);
            p += align(sizeof(key_type) * n);
            b.counts = reinterpret_cast<value_type*>(p);
            p += align(sizeof(value_type) * n);
            b.rep = reinterpret_cast<std::uint32_t*>(p);
            p += align(sizeof(std::uint32_t) * n);
            b.pos = reinterpret_cast<std::uint32_t*>(p);
            p += align(sizeof(std::uint32_t) * n);
            b.status = reinterpret_cast<Status*>(p);
            p += align(sizeof(std::uint32_t) * n);
            b.hot = reinterpret_cast<std::uint32_t*>(p);
            p += hot_bytes();
            b.num = reinterpret_cast<index_type*>(p);
            b.capacity = n;
            return b;
        }
    };

    // Combining passes in chunks of the table's own buffers (grown on
    // demand); each pass reads its entry count back to size the insert.
    template<class StatusHandler>
    WC_H void insert_combined(const key_type* keys_in, index_type num_in, hipStream_t stream,
                              index_type probing_length, typename StatusHandler::base_type* status_out) noexcept {
        constexpr index_type chunk = index_type(1) << 26;
        const index_type cap = num_in < chunk ? num_in : chunk;
        if (!scratch_.reserve(cap)) {   // no memory for the combining buffers: insert each occurrence
            insert_each<view_type, StatusHandler, false>(keys_in, nullptr, num_in, view(), stream, probing_length,
                                                         status_out);
            return;
        }
        const CombineBuffers buf = scratch_.buffers();
        for (index_type first = 0; first < num_in; first += cap) {
            const index_type n = num_in - first < cap ? num_in - first : cap;
            combine(keys_in + first, n, buf, stream);
            if constexpr (detail::fixed_step_scheme<ProbingScheme>::value) {
                // the insert reads the entry count on the device: no host round trip
                insert_entries<StatusHandler>(buf, n, stream, probing_length, status_out);
            } else {
                index_type entries = 0;
                (void)hipMemcpyAsync(&entries, buf.num, sizeof(index_type), hipMemcpyDeviceToHost, stream);
                (void)hipStreamSynchronize(stream);
                if (entry_statuses<StatusHandler>(status_out))
                    insert_each<add_view_type, status_handlers::ReturnStatus, true>(
                        buf.keys, buf.counts, entries, add_view(), stream, probing_length, buf.status);
                else
                    insert_each<add_view_type, status_handlers::ReturnNothing, true>(
                        buf.keys, buf.counts, entries, add_view(), stream, probing_length, nullptr);
            }
            finish_combined<StatusHandler>(buf, first, n, stream, status_out);
        }
    }

    // One combining pass over num_in <= buf.capacity occurrences in
    // caller-provided buffers, with no host round trip: the insert reads the
    // entry count on the device (refill kernel, sized by num_in). Results,
    // counts and statuses are those of insert(). Requires a fixed-step
    // probing scheme.
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert_combined_async(const key_type* keys_in, index_type num_in, const CombineBuffers& buf,
                                    hipStream_t stream = 0, index_type probing_length = defaults::probing_length(),
                                    typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(detail::fixed_step_scheme<ProbingScheme>::value, "refill kernels need a fixed-step scheme");
        // entries index below 2^31 (the top bit of an occurrence's entry marks a hot-map entry)
        if (!base_table_.is_initialized() || num_in == 0 || num_in > buf.capacity ||
            num_in >= (index_type(1) << 31))
            return;
        combine(keys_in, num_in, buf, stream);
        insert_entries<StatusHandler>(buf, num_in, stream, probing_length, status_out);
        finish_combined<StatusHandler>(buf, 0, num_in, stream, status_out);
    }

private:
    // the distinct keys and counts of keys_in[0 .. n) into buf (entry count
    // in *buf.num, cleared first by a fill kernel)
    WC_H void combine(const key_type* keys_in, index_type n, const CombineBuffers& buf, hipStream_t stream) noexcept {
        storage::kernels::fill<index_type><<<1, 64, 0, stream>>>(buf.num, 1, index_type(0));
        constexpr auto k = kernels::count_combine<key_type, value_type, EmptyKey, TombstoneKey>;
        // one round of resident blocks (a second round would run as a tail),
        // each combining many tiles in its hot map
        static const index_type per_cu = [] {
            int b = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, 256, 0) != hipSuccess || b < 1) b = 1;
            return index_type(b);
        }();
        const index_type tiles = SDIV(n, index_type(detail::kCombineTile));
        index_type limit = index_type(detail::num_compute_units()) * per_cu;
        if (limit > index_type(detail::kCombineBlocks)) limit = detail::kCombineBlocks;
        k<<<unsigned(tiles < limit ? tiles : limit), 256, 0, stream>>>(keys_in, n, buf.keys, buf.counts, buf.rep,
                                                                      buf.pos, buf.num, buf.hot);
    }

    // whether the entries' statuses are needed (per-occurrence statuses asked for)
    template<class StatusHandler>
    WC_HD static bool entry_statuses(const typename StatusHandler::base_type* status_out) noexcept {
        return !std::is_same<StatusHandler, status_handlers::ReturnNothing>::value && status_out != nullptr;
    }

    // insert of the combined entries of one pass (at most n; the kernel
    // reads their number on the device), writing the entries' statuses only
    // when per-occurrence statuses follow
    template<class StatusHandler>
    WC_H void insert_entries(const CombineBuffers& buf, index_type n, hipStream_t stream, index_type probing_length,
                             typename StatusHandler::base_type* status_out) noexcept {
        constexpr int U = detail::refill_insert_keys<unsigned(cg_size())>();
        if (entry_statuses<StatusHandler>(status_out))
            detail::launch_insert_refill<add_view_type, status_handlers::ReturnStatus, U, true>(
                buf.keys, buf.counts, n, add_view(), probing_length, buf.status, stream, buf.num);
        else
            detail::launch_insert_refill<add_view_type, status_handlers::ReturnNothing, U, true>(
                buf.keys, buf.counts, n, add_view(), probing_length, nullptr, stream, buf.num);
    }

    // per-occurrence statuses of occurrences [first, first + n) from the
    // entries' statuses (the table's status joined duplicate_key in the
    // insert, for every entry of more than one occurrence)
    template<class StatusHandler>
    WC_H void finish_combined(const CombineBuffers& buf, index_type first, index_type n, hipStream_t stream,
                              typename StatusHandler::base_type* status_out) noexcept {
        if constexpr (!std::is_same<StatusHandler, status_handlers::ReturnNothing>::value)
            if (status_out != nullptr)
                kernels::count_expand<StatusHandler><<<detail::grid_for(n, 1, 8), block_size(), 0, stream>>>(
                    buf.pos, buf.hot, buf.rep, buf.status, first, n, status_out);
    }

public:
    // counting_hash_table.cuh:236-249
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void retrieve(const key_type* keys_in, index_type num_in, value_type* values_out, hipStream_t stream = 0,
                       index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!base_table_.is_initialized() || num_in == 0) return;
        if constexpr (detail::LaneWindow<typename view_type::slot_type, cg_size()>::usable) {
            if (capacity() * sizeof(typename view_type::slot_type) <= detail::lane_table_bytes()) {
                kernels::retrieve_lanes<view_type, StatusHandler, detail::keys_per_lane()>
                    <<<detail::grid_for_lanes(num_in), block_size(), 0, stream>>>(keys_in, num_in, values_out, view(),
                                                                                 probing_length, status_out);
                return;
            }
        }
        if constexpr (detail::fixed_step_scheme<typename view_type::probing_scheme_type>::value) {
            if (detail::refill_lookups()) {
                detail::launch_retrieve_refill<view_type, StatusHandler, detail::refill_keys<unsigned(cg_size())>()>(
                    keys_in, num_in, values_out, view(), probing_length, status_out, stream);
                return;
            }
            if (detail::lockstep_lookups()) {
                detail::launch_retrieve_lockstep<view_type, StatusHandler, detail::lockstep_keys<unsigned(cg_size())>()>(
                    keys_in, num_in, values_out, view(), probing_length, status_out, stream);
                return;
            }
        }
        kernels::retrieve_batched<view_type, StatusHandler, detail::keys_per_group()>
            <<<detail::grid_for_batched(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, num_in, values_out, view(), probing_length, status_out);
    }

    WC_H void retrieve_all(key_type* keys_out, value_type* values_out, index_type& num_out,
                           hipStream_t stream = 0) const noexcept {
        base_table_.retrieve_all(keys_out, values_out, num_out, stream);
    }

    // counting_hash_table.cuh:273-301 (stream and probing length in their places)
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void erase(const key_type* keys_in, index_type num_in, hipStream_t stream = 0,
                    index_type probing_length = defaults::probing_length(),
                    typename StatusHandler::base_type* status_out = nullptr) noexcept {
        base_table_.template erase<StatusHandler>(keys_in, num_in, stream, probing_length, status_out);
    }

    template<class Func>
    WC_H void for_each(Func f, hipStream_t stream = 0, index_type smem_bytes = 0) const noexcept {
        base_table_.for_each(f, stream, smem_bytes);
    }

    WC_H index_type size(hipStream_t stream = 0) const noexcept { return base_table_.size(stream); }
    WC_H float load_factor(hipStream_t stream = 0) const noexcept { return base_table_.load_factor(stream); }
    WC_H float storage_density(hipStream_t stream = 0) const noexcept { return base_table_.storage_density(stream); }
    WC_HD index_type capacity() const noexcept { return base_table_.capacity(); }
    WC_H index_type bytes_total() const noexcept { return base_table_.bytes_total(); }
    WC_H status_type peek_status(hipStream_t stream = 0) const noexcept { return base_table_.peek_status(stream); }
    WC_H status_type pop_status(hipStream_t stream = 0) noexcept { return base_table_.pop_status(stream); }
    WC_HD bool is_initialized() const noexcept { return base_table_.is_initialized(); }
    WC_HD static constexpr bool is_empty_key(key_type key) noexcept { return base_type::is_empty_key(key); }
    WC_HD static constexpr bool is_tombstone_key(key_type key) noexcept { return base_type::is_tombstone_key(key); }
    WC_HD static constexpr bool is_valid_key(key_type key) noexcept { return base_type::is_valid_key(key); }
    WC_HD key_type seed() const noexcept { return base_table_.seed(); }
    WC_HD value_type max_count() const noexcept { return max_count_; }
    WC_HD bool is_copy() const noexcept { return is_copy_; }

    WC_HD view_type view() const noexcept { return view_type{base_table_.view(), max_count_}; }
    WC_HD add_view_type add_view() const noexcept { return add_view_type{base_table_.view(), max_count_}; }

private:
    // device buffer of insert_combined, grown on demand
    struct Scratch {
        void* base = nullptr;
        index_type capacity = 0;
        WC_H bool reserve(index_type n) noexcept {
            if (n <= capacity) return true;
            release();
            if (hipMalloc(&base, CombineBuffers::bytes_for(n)) != hipSuccess) {
                base = nullptr;
                return false;
            }
            capacity = n;
            return true;
        }
        WC_H CombineBuffers buffers() const noexcept { return CombineBuffers::carve(base, capacity); }
        WC_H void release() noexcept {
            if (base) (void)hipFree(base);
            base = nullptr;
            capacity = 0;
        }
    };

    base_type base_table_;
    value_type max_count_;
    bool is_copy_;
    Scratch scratch_;
};

}  // namespace warpcore
