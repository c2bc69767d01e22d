// CountingHashTable: key -> number of occurrences (drop-in for
// include/warpcore/counting_hash_table.cuh). The key is found or claimed
// exactly as in SingleValueHashTable, then one lane adds 1 to the slot's
// counter with an atomicAdd while it is below max_count
// (counting_hash_table.cuh:141-172); lookups report min(count, max_count).
// A counter that wraps is detected from the atomic's own return value (the
// old count was the type's maximum), flagged index_overflow and clamped.
#pragma once

#include "single_value_hash_table.hpp"

namespace warpcore {

namespace detail {

template<class Base>
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
    template<class Join>
    WC_D void finish_insert(index_type slot, const value_type&, bool, Status& status, Join& join) const noexcept {
        value_type* v = &(base.slots[slot].value);
        if (atomic_load_relaxed(v) < max_count) {
            const value_type old = atomic_add(v, value_type(1));
            if (old == std::numeric_limits<value_type>::max()) {
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
        : base_table_(o.base_table_), max_count_(o.max_count_), is_copy_(true) {}
    WC_H CountingHashTable(CountingHashTable&& o) noexcept
        : base_table_(std::move(o.base_table_)), max_count_(o.max_count_), is_copy_(o.is_copy_) {
        o.is_copy_ = true;
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

    // counting_hash_table.cuh:182-194
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert(const key_type* keys_in, index_type num_in, hipStream_t stream = 0,
                     index_type probing_length = defaults::probing_length(),
                     typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!base_table_.is_initialized() || num_in == 0) return;
        if constexpr (detail::LaneWindow<typename view_type::slot_type, cg_size()>::usable) {
            if (capacity() * sizeof(typename view_type::slot_type) <= detail::lane_table_bytes()) {
                kernels::insert_lanes<view_type, StatusHandler, detail::keys_per_lane(), false>
                    <<<detail::grid_for_lanes(num_in), block_size(), 0, stream>>>(keys_in, nullptr, num_in, view(),
                                                                                 probing_length, status_out);
                return;
            }
        }
        if constexpr (detail::fixed_step_scheme<typename view_type::probing_scheme_type>::value) {
            if (detail::lockstep_inserts()) {
                detail::launch_insert_lockstep<view_type, StatusHandler,
                                               detail::lockstep_insert_keys<unsigned(cg_size())>(), false>(
                    keys_in, nullptr, num_in, view(), probing_length, status_out, stream);
                return;
            }
        }
        kernels::insert_batched<view_type, StatusHandler, detail::keys_per_group(), false>
            <<<detail::grid_for_batched(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, nullptr, num_in, view(), probing_length, status_out);
    }

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

private:
    base_type base_table_;
    value_type max_count_;
    bool is_copy_;
};

}  // namespace warpcore
