// MultiValueHashTable: key -> many values, every (key, value) pair in a slot
// of its own along the key's probe sequence. Drop-in for
// include/warpcore/multi_value_hash_table.cuh: same template parameters and
// host/device calls.
//
// Insert (multi_value_hash_table.cuh:200-289): walk the key's windows,
// counting the key's slots; at max_values_per_key the pair is refused
// (duplicate_key + max_values_for_key_reached); otherwise the lowest EMPTY
// lane of the window claims its slot with a CAS; a CAS that loses to the
// same key counts one more value. The first value of a key returns none
// and bumps the distinct-key counter, later ones return duplicate_key (not
// joined into the table status, as in the reference).
//
// Per-key traversal (for_each, :605-661): values come in probe order; it
// stops at the first window holding an EMPTY slot or once max_values_per_key
// values were seen, with probing capped at the capacity.
//
// Retrieve is two-pass (:370-510): count the values of every key into
// end_offsets, prefix-sum the counts, set begin_offsets, then every key
// writes its values at begin_offsets[i]. A key whose count changed between
// the passes reports invalid_phase_overlap; its writes stay inside its own
// [begin, end) segment (the reference writes past it).
#pragma once

#include "base.hpp"
#include "capacity.hpp"
#include "defaults.hpp"
#include "engine.hpp"
#include "probing_schemes.hpp"
#include "refill_probing.hpp"
#include "scan.hpp"
#include "single_value_hash_table.hpp"
#include "status.hpp"
#include "storage.hpp"
#include "tile.hpp"

namespace warpcore {

namespace detail {

// Distinct-key counter of a batch kernel: one atomicAdd per wavefront at the
// end instead of one per new key.
struct DeferredCount {
    index_t acc = 0;
    WC_D void operator()(index_t n) noexcept { acc += n; }
    WC_D void flush(index_t* word) noexcept {
        const index_t all = tile<64>::reduce_add(acc);
        if (lane_id() == 0 && all != 0 && word != nullptr) atomic_add(word, all);
    }
};
struct ImmediateCount {
    index_t* word;
    WC_D void operator()(index_t n) const noexcept { if (word != nullptr && n) atomic_add(word, n); }
};

// Core of the bulk insert with several pairs in flight per group
// (refill_probing.hpp): the slot hooks of the single-value table, the
// per-key maximum, and the distinct-key counter, added to once per
// wavefront at the end.
template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class ProbingScheme, class Slot>
struct MultiValueInsertCore : SingleValueView<Key, Value, EmptyKey, TombstoneKey, ProbingScheme, Slot> {
    static constexpr bool multi_value = true;
    index_t max_values;
    index_t* num_keys;
    WC_D void add_keys(index_t n) const noexcept {
        const index_t all = tile<64>::reduce_add(n);
        if (lane_id() == 0 && all != 0 && num_keys != nullptr) atomic_add(num_keys, all);
    }
};

// Cores of the two lookup passes with several keys in flight per group
// (refill_probing.hpp): the count pass (Gather = false) writes each key's
// value count and adds them up into *total; the gather pass writes the
// j-th value of key i to values_out[begin_in[i] + j] while j is below the
// key's room end_in[i] - begin_in[i], and flags invalid_phase_overlap when
// the key's count differs from its room (gpu_engine.cuh:305-350, 466-504).
template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class ProbingScheme, class Slot, bool Gather>
struct MultiValueLookupCore : SingleValueView<Key, Value, EmptyKey, TombstoneKey, ProbingScheme, Slot> {
    static constexpr bool multi_value = true;
    index_t max_values;
    index_t* per_key_out;       // count pass, nullable
    index_t* total;             // count pass
    const index_t* begin_in;    // gather pass
    const index_t* end_in;
    Value* values_out;
    WC_D void range_of(index_t i, index_t& base, index_t& room) const noexcept {
        if constexpr (Gather) {
            base = begin_in[i];
            room = end_in[i] - base;
        } else {
            base = 0;
            room = 0;
        }
    }
    WC_D void emit(index_t base, index_t room, index_t j, const Value& v) const noexcept {
        if constexpr (Gather)
            if (j < room) values_out[base + j] = v;
    }
    template<class Join>
    WC_D void finish(index_t i, index_t count, index_t room, Status& st, Join& join, index_t* tally) const noexcept {
        if constexpr (Gather) {
            if (count != room) {
                join(Status::invalid_phase_overlap());
                st += Status::invalid_phase_overlap();
            }
        } else {
            if (per_key_out != nullptr) per_key_out[i] = count;
            *tally += count;
        }
    }
    WC_D void add_tally(index_t n) const noexcept {
        if constexpr (!Gather) {
            const index_t all = tile<64>::reduce_add(n);
            if (lane_id() == 0 && all != 0 && total != nullptr) atomic_add(total, all);
        }
    }
};

template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class ProbingScheme, class Slot>
struct MultiValueView {
    using key_type = Key;
    using value_type = Value;
    using slot_type = Slot;
    using index_type = index_t;
    static constexpr bool packed = is_packed_pair<Slot>();

    Slot* slots;
    Status* status;
    index_t* num_keys;
    probing_schemes::Geometry geo;
    Key seed;
    index_t max_values;
    bool initialized;

    WC_HD static constexpr index_type cg_size() noexcept { return ProbingScheme::cg_size(); }
    WC_HD static constexpr bool is_empty_key(Key k) noexcept { return k == EmptyKey; }
    WC_HD static constexpr bool is_valid_key(Key k) noexcept { return k != EmptyKey && k != TombstoneKey; }
    WC_HD index_type capacity() const noexcept { return geo.capacity; }
    WC_HD Status* status_word() const noexcept { return status; }
    WC_D Key key_at(index_type i) const noexcept { return slots[i].key; }
    WC_D Slot slot_at(index_type i) const noexcept { return load_slot(slots + i); }

    template<class Join, class Count>
    WC_D Status insert(Key key_in, const Value& value_in, const tile<cg_size()>& group, index_type probing_length,
                       Join& join, Count& count_key) const noexcept {
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) {
            join(Status::invalid_key());
            return Status::invalid_key();
        }
        const Status full = Status::duplicate_key() + Status::max_values_for_key_reached();
        ProbingScheme iter(geo, probing_length, group);
        index_type num_values = 0;
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Slot s = load_slot(slots + i);
            const Key table_key = s.key;
            auto empty_mask = group.ballot(is_empty_key(table_key));
            num_values += __builtin_popcountll(group.ballot(table_key == key_in));
            if (num_values >= max_values) {
                join(full);
                return full;
            }
            while (empty_mask) {
                const unsigned leader = unsigned(ffs64(empty_mask) - 1);
                bool success = false, collision = false;
                if (group.thread_rank() == leader) {
                    if constexpr (packed) {
                        Slot want;
                        want.key = key_in;
                        want.value = value_in;
                        std::uint64_t expected = slot_bits(s);
                        while (true) {
                            const std::uint64_t old = atomic_cas(reinterpret_cast<std::uint64_t*>(slots + i),
                                                                 expected, slot_bits(want));
                            if (old == expected) { success = true; break; }
                            const Key old_key = slot_from_bits<Slot>(old).key;
                            if (old_key != EmptyKey) { collision = old_key == key_in; break; }
                            expected = old;   // an EMPTY slot with other value bits: retry
                        }
                    } else {
                        const Key old = atomic_cas(&(slots[i].key), table_key, key_in);
                        success = old == table_key;
                        collision = old == key_in;
                        if (success) slots[i].value = value_in;
                    }
                    if (success && num_values == 0) count_key(1);
                }
                if (group.any(success)) return num_values > 0 ? Status::duplicate_key() : Status::none();
                num_values += group.any(collision);
                if (num_values >= max_values) {
                    join(full);
                    return full;
                }
                empty_mask &= empty_mask - 1;
            }
        }
        const Status st = num_values > 0 ? Status::probing_length_exceeded() + Status::duplicate_key()
                                         : Status::probing_length_exceeded();
        join(st);
        return st;
    }

    // f(key, value, j) for the j-th value of key_in in probe order
    template<class Func, class Join>
    WC_D Status for_each(Func&& f, Key key_in, index_type& num_out, const tile<cg_size()>& group,
                         index_type probing_length, Join& join) const noexcept {
        num_out = 0;
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) {
            join(Status::invalid_key());
            return Status::invalid_key();
        }
        ProbingScheme iter(geo, probing_length < capacity() ? probing_length : capacity(), group);
        index_type num = 0;
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Slot s = load_slot(slots + i);
            const bool hit = s.key == key_in;
            const auto hit_mask = group.ballot(hit);
            if (hit) f(key_in, s.value, num + __builtin_popcountll(hit_mask & group.lanemask_lt()));
            num += __builtin_popcountll(hit_mask);
            if (group.any(is_empty_key(s.key)) || num >= max_values) {
                num_out = num;
                if (num == 0) {
                    join(Status::key_not_found());
                    return Status::key_not_found();
                }
                return Status::none();
            }
        }
        num_out = num;
        join(Status::probing_length_exceeded());
        return Status::probing_length_exceeded();
    }

    // slot `at` holds the first value of its key along the key's probe sequence
    WC_D bool is_first_of_key(Key key, index_type at, const tile<cg_size()>& group) const noexcept {
        ProbingScheme iter(geo, capacity(), group);
        for (index_type i = iter.begin(key, seed); i != iter.end(); i = iter.next()) {
            const auto hit_mask = group.ballot(slots[i].key == key);
            if (hit_mask) return group.shfl(i, unsigned(ffs64(hit_mask) - 1)) == at;
        }
        return false;
    }
};

}  // namespace detail

namespace kernels {

template<class View, class StatusHandler>
__global__ __launch_bounds__(256) void mv_insert(const typename View::key_type* __restrict__ keys_in,
                                                 const typename View::value_type* __restrict__ values_in,
                                                 index_t num_in, View view, index_t probing_length,
                                                 typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    detail::DeferredCount count;
    const detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        const Status st = view.insert(keys_in[gid], values_in[gid], group, probing_length, join, count);
        if (group.thread_rank() == 0) StatusHandler::handle(st, status_out, gid);
    }
    join.flush(view.status_word());
    count.flush(view.num_keys);
}

// per-key value counts and their total (gpu_engine.cuh:466-504)
template<class View, class StatusHandler>
__global__ __launch_bounds__(256) void mv_num_values(const typename View::key_type* __restrict__ keys_in,
                                                     index_t num_in, index_t* __restrict__ total_out,
                                                     index_t* __restrict__ num_per_key_out, View view,
                                                     index_t probing_length,
                                                     typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    detail::DeferredCount total;
    const detail::GroupRange<CG> r;
    auto nothing = [](typename View::key_type, const typename View::value_type&, index_t) {};
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        index_t num = 0;
        const Status st = view.for_each(nothing, keys_in[gid], num, group, probing_length, join);
        if (group.thread_rank() == 0) {
            if (num_per_key_out != nullptr) num_per_key_out[gid] = num;
            total(num);
            StatusHandler::handle(st, status_out, gid);
        }
    }
    join.flush(view.status_word());
    total.flush(total_out);
}

// gather pass of the two-pass retrieve (gpu_engine.cuh:305-350)
template<class View, class StatusHandler>
__global__ __launch_bounds__(256) void mv_retrieve(const typename View::key_type* __restrict__ keys_in,
                                                   index_t num_in, const index_t* __restrict__ begin_in,
                                                   const index_t* __restrict__ end_in,
                                                   typename View::value_type* __restrict__ values_out, View view,
                                                   index_t probing_length,
                                                   typename StatusHandler::base_type* __restrict__ status_out) {
    using V = typename View::value_type;
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    const detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        const index_t begin = begin_in[gid];
        const index_t room = end_in[gid] - begin;
        V* out = values_out + begin;
        index_t num = 0;
        Status st = view.for_each(
            [=](typename View::key_type, const V& v, index_t j) {
                if (j < room) out[j] = v;
            },
            keys_in[gid], num, group, probing_length, join);
        if (group.thread_rank() == 0) {
            if (num != room) {
                join(Status::invalid_phase_overlap());
                st += Status::invalid_phase_overlap();
            }
            StatusHandler::handle(st, status_out, gid);
        }
    }
    join.flush(view.status_word());
}

// f(key, value, j) over the values of each key (gpu_engine.cuh:95-125)
template<class Func, class View, class StatusHandler>
__global__ __launch_bounds__(256) void mv_for_each(Func f, const typename View::key_type* __restrict__ keys_in,
                                                   index_t num_in, View view, index_t probing_length,
                                                   typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    const detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        index_t num = 0;
        const Status st = view.for_each(f, keys_in[gid], num, group, probing_length, join);
        if (group.thread_rank() == 0) StatusHandler::handle(st, status_out, gid);
    }
    join.flush(view.status_word());
}

// Each distinct key once: slot i is emitted when it holds the first value
// of its key along the key's probe sequence (gpu_engine.cuh:48-93). A
// wavefront loads 64 consecutive slots; each group of CG lanes then walks
// the probe sequences of its lanes' keys together.
template<class View>
__global__ __launch_bounds__(256) void mv_unique_keys(typename View::key_type* __restrict__ keys_out,
                                                      index_t* __restrict__ num_out, View view) {
    using K = typename View::key_type;
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    const index_t cap = view.capacity();
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    const index_t wave_first = index_t(blockIdx.x) * blockDim.x + threadIdx.x - lane_id();
    for (index_t base = wave_first; base < cap; base += stride) {
        const index_t i = base + lane_id();
        const K key = i < cap ? view.key_at(i) : K{};
        const bool valid = i < cap && view.is_valid_key(key);
        bool first = false;
        auto todo = group.ballot(valid);
        while (todo) {
            const unsigned r = unsigned(ffs64(todo) - 1);
            const K k = group.shfl(key, r);
            const index_t at = group.shfl(i, r);
            const bool is_first = view.is_first_of_key(k, at, group);
            if (group.thread_rank() == r) first = is_first;
            todo &= todo - 1;
        }
        const index_t out = atomic_agg_inc(num_out, first);
        if (first) keys_out[out] = key;
    }
}

// the same status for every key of a batch (dry runs)
template<class StatusHandler>
__global__ __launch_bounds__(256) void handle_all(Status st, typename StatusHandler::base_type* __restrict__ out,
                                                  index_t num) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num; i += stride)
        StatusHandler::handle(st, out, i);
}

// begin[i] = end[i - 1], begin[0] = 0 (the offsets shift of
// multi_value_hash_table.cuh:476-488)
template<class T>
__global__ __launch_bounds__(256) void shift_offsets(T* begin_out, const T* end_in, index_t num) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num; i += stride)
        begin_out[i] = i == 0 ? 0 : end_in[i - 1];
}

}  // namespace kernels

template<class Key, class Value, Key EmptyKey = defaults::empty_key<Key>(),
         Key TombstoneKey = defaults::tombstone_key<Key>(),
         class ProbingScheme = defaults::probing_scheme_t<Key, 8>,
         class TableStorage = defaults::table_storage_t<Key, Value>,
         index_t TempMemoryBytes = defaults::temp_memory_bytes()>
class MultiValueHashTable {
    static_assert(checks::is_valid_key_type<Key>(), "invalid key type");
    static_assert(EmptyKey != TombstoneKey, "empty key and tombstone key must not be identical");
    static_assert(checks::is_cycle_free_probing_scheme<ProbingScheme>(), "not a valid probing scheme type");
    static_assert(std::is_same<typename ProbingScheme::key_type, Key>::value,
                  "probing key type differs from table's key type");
    static_assert(checks::is_key_value_storage<TableStorage>(), "not a valid storage type");
    static_assert(std::is_same<typename TableStorage::key_type, Key>::value,
                  "storage's key type differs from table's key type");
    static_assert(std::is_same<typename TableStorage::value_type, Value>::value,
                  "storage's value type differs from table's value type");
    static_assert(TempMemoryBytes >= sizeof(index_t), "temporary storage must at least be of size index_type");

    using temp_type = storage::CyclicStore<index_t>;

public:
    using key_type = Key;
    using value_type = Value;
    using index_type = index_t;
    using status_type = Status;
    using probing_scheme_type = ProbingScheme;
    using slot_type = typename TableStorage::slot_type;
    using view_type = detail::MultiValueView<Key, Value, EmptyKey, TombstoneKey, ProbingScheme, slot_type>;
    using insert_core_type = detail::MultiValueInsertCore<Key, Value, EmptyKey, TombstoneKey, ProbingScheme, slot_type>;
    template<bool Gather>
    using lookup_core_type =
        detail::MultiValueLookupCore<Key, Value, EmptyKey, TombstoneKey, ProbingScheme, slot_type, Gather>;

    WC_HD static constexpr key_type empty_key() noexcept { return EmptyKey; }
    WC_HD static constexpr key_type tombstone_key() noexcept { return TombstoneKey; }
    WC_HD static constexpr index_type cg_size() noexcept { return ProbingScheme::cg_size(); }

    // multi_value_hash_table.cuh:93-123
    WC_H explicit MultiValueHashTable(index_type min_capacity, key_type seed = defaults::seed<key_type>(),
                                      index_type max_values_per_key = std::numeric_limits<index_type>::max(),
                                      bool no_init = false) noexcept
        : status_(nullptr), table_(detail::get_valid_capacity(min_capacity, cg_size())),
          temp_(TempMemoryBytes / sizeof(index_type)), seed_(seed), max_values_per_key_(max_values_per_key),
          num_keys_(nullptr), scan_parts_(nullptr), is_copy_(false), is_initialized_(false),
          geo_(probing_schemes::Geometry::make(table_.capacity(), cg_size())) {
        if (hipMalloc(&status_, sizeof(status_type)) != hipSuccess) status_ = nullptr;
        if (hipMalloc(&num_keys_, sizeof(index_type)) != hipSuccess) num_keys_ = nullptr;
        if (hipMalloc(&scan_parts_, sizeof(index_type) * detail::kScanParts) != hipSuccess) scan_parts_ = nullptr;
        status_type st = table_.status() + temp_.status();
        if (num_keys_ == nullptr || scan_parts_ == nullptr) st += Status::out_of_memory() + Status::not_initialized();
        assign_status(st);
        if (!no_init) init();
    }
    WC_HD MultiValueHashTable(const MultiValueHashTable& o) noexcept
        : status_(o.status_), table_(o.table_), temp_(o.temp_), seed_(o.seed_),
          max_values_per_key_(o.max_values_per_key_), num_keys_(o.num_keys_), scan_parts_(o.scan_parts_),
          is_copy_(true), is_initialized_(o.is_initialized_), geo_(o.geo_) {}
    WC_H MultiValueHashTable(MultiValueHashTable&& o) noexcept
        : status_(o.status_), table_(std::move(o.table_)), temp_(std::move(o.temp_)), seed_(o.seed_),
          max_values_per_key_(o.max_values_per_key_), num_keys_(o.num_keys_), scan_parts_(o.scan_parts_),
          is_copy_(o.is_copy_), is_initialized_(o.is_initialized_), geo_(o.geo_) {
        o.is_copy_ = true;
    }
    WC_HD ~MultiValueHashTable() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_) {
            if (status_ != nullptr) (void)hipFree(status_);
            if (num_keys_ != nullptr) (void)hipFree(num_keys_);
            if (scan_parts_ != nullptr) (void)hipFree(scan_parts_);
        }
#endif
    }

    // :160-177; empties every slot (key and value bits)
    WC_H void init(hipStream_t stream = 0) noexcept {
        is_initialized_ = false;
        if (!table_.status().has_not_initialized() && !temp_.status().has_not_initialized() &&
            num_keys_ != nullptr && scan_parts_ != nullptr) {
            table_.init_pairs(empty_key(), value_type{}, stream);
            assign_status(table_.status() + temp_.status(), stream);
            (void)hipMemsetAsync(num_keys_, 0, sizeof(index_type), stream);
            is_initialized_ = true;
        }
    }

    // ---- device calls (tile of cg_size() converged lanes)
    WC_D status_type insert(key_type key_in, const value_type& value_in, const tile<cg_size()>& group,
                            index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        detail::ImmediateCount count{num_keys_};
        return view().insert(key_in, value_in, group, probing_length, join, count);
    }
    // :325-355: values_out == nullptr counts only (dry run)
    WC_D status_type retrieve(key_type key_in, value_type* values_out, index_type& num_out,
                              const tile<cg_size()>& group,
                              index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        if (values_out == nullptr) {
            const status_type st = num_values(key_in, num_out, group, probing_length);
            join(Status::dry_run());
            return Status::dry_run() + st;
        }
        return view().for_each([=](key_type, const value_type& v, index_type j) { values_out[j] = v; }, key_in,
                               num_out, group, probing_length, join);
    }
    template<class Func>
    WC_D status_type for_each(Func f, key_type key_in, index_type& num_values_out, const tile<cg_size()>& group,
                              index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().for_each(f, key_in, num_values_out, group, probing_length, join);
    }
    WC_D status_type num_values(key_type key_in, index_type& num_out, const tile<cg_size()>& group,
                                index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().for_each([](key_type, const value_type&, index_type) {}, key_in, num_out, group,
                               probing_length, join);
    }

    // ---- host calls
    // :297-319
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert(const key_type* keys_in, const value_type* values_in, index_type num_in,
                     hipStream_t stream = 0, index_type probing_length = defaults::probing_length(),
                     typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_ || num_in == 0) return;
        if constexpr (detail::fixed_step_scheme<ProbingScheme>::value) {
            if (detail::refill_inserts()) {   // several pairs in flight per group
                const insert_core_type core{{table_.data(), status_, geo_, seed_, is_initialized_},
                                            max_values_per_key_, num_keys_};
                detail::launch_insert_refill<insert_core_type, StatusHandler,
                                             detail::refill_insert_keys<unsigned(cg_size())>(), true>(
                    keys_in, values_in, num_in, core, probing_length, status_out, stream);
                return;
            }
        }
        kernels::mv_insert<view_type, StatusHandler>
            <<<detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(keys_in, values_in, num_in, view(),
                                                                                probing_length, status_out);
    }

    // :370-510. With values_out == nullptr only the counts are computed:
    // end_offsets_out[i] = number of values of keys_in[i] (not scanned),
    // num_out = their sum, and status dry_run.
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void retrieve(const key_type* keys_in, index_type num_in, index_type* begin_offsets_out,
                       index_type* end_offsets_out, value_type* values_out, index_type& num_out,
                       hipStream_t stream = 0, index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_) return;
        if (num_in > index_type(std::numeric_limits<int>::max())) {   // :395-400
            join_status(Status::index_overflow(), stream);
            return;
        }
        if (num_in == 0) {
            num_out = 0;
            return;
        }
        num_values(keys_in, num_in, num_out, end_offsets_out, stream, probing_length);
        if (values_out != nullptr) {
            detail::device_scan(end_offsets_out, end_offsets_out, num_in, scan_parts_, true, stream);
            kernels::shift_offsets<index_type><<<detail::grid_for(num_in, 1, 8), block_size(), 0, stream>>>(
                begin_offsets_out, end_offsets_out, num_in);
            bool gathered = false;
            if constexpr (detail::fixed_step_scheme<ProbingScheme>::value) {
                if (detail::refill_lookups()) {   // several keys in flight per group
                    const lookup_core_type<true> core{{table_.data(), status_, geo_, seed_, is_initialized_},
                                                      max_values_per_key_, nullptr, nullptr,
                                                      begin_offsets_out, end_offsets_out, values_out};
                    detail::launch_retrieve_refill<lookup_core_type<true>, StatusHandler, 4>(
                        keys_in, num_in, nullptr, core, probing_length, status_out, stream);
                    gathered = true;
                }
            }
            if (!gathered)
                kernels::mv_retrieve<view_type, StatusHandler>
                    <<<detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(
                        keys_in, num_in, begin_offsets_out, end_offsets_out, values_out, view(), probing_length,
                        status_out);
        } else {
            if (status_out != nullptr)
                kernels::handle_all<StatusHandler><<<detail::grid_for(num_in, 1, 8), block_size(), 0, stream>>>(
                    Status::dry_run(), status_out, num_in);
            join_status(Status::dry_run(), stream);
        }
        if (stream == 0) (void)hipStreamSynchronize(stream);
    }

    // :520-548. The reference types num_values_out as value_type&, which only
    // compiles where value_type is index_type; this takes index_type&.
    WC_H void retrieve_all(key_type* keys_out, index_type& num_keys_out, index_type* begin_offsets_out,
                           index_type* end_offsets_out, value_type* values_out, index_type& num_values_out,
                           hipStream_t stream = 0) const noexcept {
        if (!is_initialized_) return;
        retrieve_all_keys(keys_out, num_keys_out, stream);
        if (keys_out != nullptr)
            retrieve(keys_out, num_keys_out, begin_offsets_out, end_offsets_out, values_out, num_values_out,
                     stream);
        if (stream == 0) (void)hipStreamSynchronize(stream);
    }

    // :556-591; order unspecified
    WC_H void retrieve_all_keys(key_type* keys_out, index_type& num_out, hipStream_t stream = 0) const noexcept {
        if (!is_initialized_) return;
        if (keys_out != nullptr) {
            index_type* tmp = temp_.get();
            (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
            kernels::mv_unique_keys<view_type>
                <<<detail::grid_for(capacity(), 1, 8), block_size(), 0, stream>>>(keys_out, tmp, view());
            (void)hipMemcpyAsync(&num_out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
            (void)hipStreamSynchronize(stream);
        } else {
            num_out = num_keys(stream);
            join_status(Status::dry_run(), stream);
        }
    }

    // :670-681: f(key, value) over every stored pair
    template<class Func>
    WC_H void for_each(Func f, hipStream_t stream = 0, index_type smem_bytes = 0) const noexcept {
        if (!is_initialized_) return;
        kernels::for_each<Func, view_type>
            <<<detail::grid_for(capacity(), 1, 8), block_size(), smem_bytes, stream>>>(f, view());
    }

    // :695-714: f(key, value, j) over the values of each key
    template<class Func, class StatusHandler = defaults::status_handler_t>
    WC_H void for_each(Func f, const key_type* keys_in, index_type num_in, hipStream_t stream = 0,
                       index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr,
                       index_type smem_bytes = 0) const noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_ || num_in == 0) return;
        kernels::mv_for_each<Func, view_type, StatusHandler>
            <<<detail::grid_for(num_in, cg_size()), block_size(), smem_bytes, stream>>>(f, keys_in, num_in, view(),
                                                                                         probing_length, status_out);
    }

    // :721-731
    WC_H index_type num_keys(hipStream_t stream = 0) const noexcept {
        index_type num = 0;
        if (num_keys_ == nullptr) return 0;
        (void)hipMemcpyAsync(&num, num_keys_, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
        return num;
    }

    // :766-793
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void num_values(const key_type* keys_in, index_type num_in, index_type& num_out,
                         index_type* num_per_key_out = nullptr, hipStream_t stream = 0,
                         index_type probing_length = defaults::probing_length(),
                         typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        if (!is_initialized_) return;
        index_type* tmp = temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        bool counted = num_in == 0;
        if constexpr (detail::fixed_step_scheme<ProbingScheme>::value) {
            if (!counted && detail::refill_lookups()) {   // several keys in flight per group
                const lookup_core_type<false> core{{table_.data(), status_, geo_, seed_, is_initialized_},
                                                   max_values_per_key_, num_per_key_out, tmp,
                                                   nullptr, nullptr, nullptr};
                detail::launch_retrieve_refill<lookup_core_type<false>, StatusHandler, 4>(
                    keys_in, num_in, nullptr, core, probing_length, status_out, stream);
                counted = true;
            }
        }
        if (!counted)
            kernels::mv_num_values<view_type, StatusHandler>
                <<<detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(
                    keys_in, num_in, tmp, num_per_key_out, view(), probing_length, status_out);
        (void)hipMemcpyAsync(&num_out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
    }

    WC_H index_type num_values(hipStream_t stream = 0) const noexcept { return size(stream); }

    // :809-834: number of stored values
    WC_H index_type size(hipStream_t stream = 0) const noexcept {
        if (!is_initialized_) return 0;
        index_type out = 0;
        index_type* tmp = temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        kernels::size<view_type><<<detail::grid_for(capacity(), 1, 8), block_size(), 0, stream>>>(tmp, view());
        (void)hipMemcpyAsync(&out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
        return out;
    }
    WC_H float load_factor(hipStream_t stream = 0) const noexcept { return float(size(stream)) / float(capacity()); }
    // :851-862
    WC_H float storage_density(hipStream_t stream = 0) const noexcept {
        const index_type key_bytes = num_keys(stream) * sizeof(key_type);
        const index_type value_bytes = num_values(stream) * sizeof(value_type);
        return float(key_bytes + value_bytes) / float(table_.bytes_total());
    }
    WC_HD index_type capacity() const noexcept { return table_.capacity(); }
    WC_H index_type bytes_total() const noexcept {
        return table_.bytes_total() + temp_.bytes_total() + sizeof(status_type) +
               sizeof(index_type) * (detail::kScanParts + 1);
    }
    WC_HD bool is_initialized() const noexcept { return is_initialized_; }
    WC_HD index_type max_values_per_key() const noexcept { return max_values_per_key_; }
    WC_HD key_type seed() const noexcept { return seed_; }

    WC_H status_type peek_status(hipStream_t stream = 0) const noexcept {
        status_type status = status_type::not_initialized();
        if (status_ != nullptr) {
            (void)hipMemcpyAsync(&status, status_, sizeof(status_type), hipMemcpyDeviceToHost, stream);
            (void)hipStreamSynchronize(stream);
        }
        return status;
    }
    WC_H status_type pop_status(hipStream_t stream = 0) noexcept {
        const status_type status = peek_status(stream);
        if (status_ != nullptr) assign_status(table_.status(), stream);
        return status;
    }

    WC_HD static constexpr bool is_empty_key(key_type key) noexcept { return key == empty_key(); }
    WC_HD static constexpr bool is_tombstone_key(key_type key) noexcept { return key == tombstone_key(); }
    WC_HD static constexpr bool is_valid_key(key_type key) noexcept {
        return key != empty_key() && key != tombstone_key();
    }
    WC_HD bool is_copy() const noexcept { return is_copy_; }

    WC_HD view_type view() const noexcept {
        return view_type{table_.data(), status_, num_keys_, geo_, seed_, max_values_per_key_, is_initialized_};
    }

private:
    WC_H void assign_status(status_type status, hipStream_t stream = 0) const noexcept {
        if (status_ == nullptr) return;
        (void)hipMemcpyAsync(status_, &status, sizeof(status_type), hipMemcpyHostToDevice, stream);
        (void)hipStreamSynchronize(stream);
    }
    WC_H void join_status(status_type status, hipStream_t stream = 0) const noexcept {
        if (status_ == nullptr) return;
        const status_type peeked = peek_status(stream);
        const status_type joined = peeked + status;
        if (joined != peeked) assign_status(joined, stream);
    }

    status_type* status_;
    TableStorage table_;
    temp_type temp_;
    key_type seed_;
    index_type max_values_per_key_;
    index_type* num_keys_;
    index_type* scan_parts_;
    bool is_copy_;
    bool is_initialized_;
    probing_schemes::Geometry geo_;
};

}  // namespace warpcore
