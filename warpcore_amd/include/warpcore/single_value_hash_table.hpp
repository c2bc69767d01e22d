// SingleValueHashTable: key -> one value, open addressing over line-aligned
// windows. Drop-in for include/warpcore/single_value_hash_table.cuh: same
// template parameters, host calls (insert / retrieve / erase / retrieve_all /
// for_each / size / load_factor / peek_status / pop_status ...) and device
// calls taking a tile<CG> (the reference's thread_block_tile<CG>).
//
// Insert protocol per probe window (single_value_hash_table.cuh:742-823):
// every lane reads one slot; a hit anywhere in the window means duplicate
// (the value is overwritten); otherwise the lowest EMPTY lane tries a
// compare-and-swap, the window's other EMPTY lanes are tried in order, and a
// CAS that returns the key itself is a duplicate. Tombstones are never
// reused. For 8-byte u32/u32 slots the CAS swaps key and value together, so
// a reader never sees the key without its value; wider slots publish the
// key first and store the value after it, as the reference does.
#pragma once

#include "base.hpp"
#include "capacity.hpp"
#include "defaults.hpp"
#include "batched_probing.hpp"
#include "lane_probing.hpp"
#include "lockstep_probing.hpp"
#include "refill_probing.hpp"
#include "engine.hpp"
#include "probing_schemes.hpp"
#include "status.hpp"
#include "storage.hpp"
#include "tile.hpp"

namespace warpcore {

template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class ProbingScheme, class TableStorage,
         index_t TempMemoryBytes>
class CountingHashTable;

template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class ValueStore, class ProbingScheme>
class BucketListHashTable;

namespace detail {

// Trivially copyable device state of a single-value table; batch kernels
// receive this by value.
template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class ProbingScheme, class Slot>
struct SingleValueView {
    using key_type = Key;
    using value_type = Value;
    using slot_type = Slot;
    using index_type = index_t;
    using probing_scheme_type = ProbingScheme;
    using out_type = Value;
    using probe_slot_type = Slot;
    static constexpr bool packed = is_packed_pair<Slot>();
    static constexpr bool publishes_value = packed;

    Slot* slots;
    Status* status;
    probing_schemes::Geometry geo;
    Key seed;
    bool initialized;
    // overflow filter (one byte per window, see mark_overflow), or null:
    // kept by every insert path of the table, read by the lane-owner lookups
    std::uint32_t* filter;

    WC_HD static constexpr index_type cg_size() noexcept { return ProbingScheme::cg_size(); }
    WC_HD static constexpr bool is_empty_key(Key k) noexcept { return k == EmptyKey; }
    WC_HD static constexpr Key empty_key_value() noexcept { return EmptyKey; }
    WC_HD static constexpr bool is_tombstone_key(Key k) noexcept { return k == TombstoneKey; }
    WC_HD static constexpr bool is_valid_key(Key k) noexcept { return k != EmptyKey && k != TombstoneKey; }
    WC_HD index_type capacity() const noexcept { return geo.capacity; }
    WC_HD Status* status_word() const noexcept { return status; }
    WC_D Key key_at(index_type i) const noexcept { return slots[i].key; }
    WC_D Slot slot_at(index_type i) const noexcept { return load_slot(slots + i); }

    // ---- hooks of the batched kernels (batched_probing.hpp)
    WC_HD bool ready() const noexcept { return initialized; }
    WC_HD Key hash_seed() const noexcept { return seed; }
    WC_HD const probing_schemes::Geometry& geometry() const noexcept { return geo; }
    WC_HD Slot* slot_ptr() const noexcept { return slots; }
    WC_HD static Key key_of(const Slot& s) noexcept { return s.key; }
    WC_HD static Value value_of(const Slot& s) noexcept { return s.value; }
    WC_D Value found(const Value& v) const noexcept { return v; }
    WC_D Status missing(Value&) const noexcept { return Status::key_not_found(); }
    // erase of the key found at slot i (single_value_hash_table.cuh:446-455)
    WC_D void erase_slot(index_type i) const noexcept { slots[i].key = TombstoneKey; }
    // Claiming CAS of the batched insert, split so that the U CASes of a
    // group are all in flight before any result is examined. cas_issue()
    // swaps the EMPTY slot this lane saw as `seen` for the key (packed slots:
    // key and value in one 64-bit word) and returns the old word;
    // cas_outcome() classifies it: 1 claimed, 2 the key was already there,
    // 0 another key won the slot, 3 the EMPTY slot's value bits moved (retry
    // with `seen` refreshed).
    using cas_word = typename std::conditional<packed, std::uint64_t, Key>::type;
    // what insert reads of a slot: the whole 8-byte packed pair, else the key
    using insert_probe_type = typename std::conditional<packed, Slot, Key>::type;
    WC_D static insert_probe_type load_insert_probe(const Slot* p) noexcept {
        if constexpr (packed) return load_slot(p);
        else return p->key;
    }
    // expected word of a claiming CAS on an EMPTY slot: EMPTY and the value
    // bits the EMPTY fill writes (cas_outcome() refreshes it if they differ)
    WC_HD static insert_probe_type empty_probe() noexcept {
        if constexpr (packed) {
            Slot s;
            s.key = EmptyKey;
            s.value = Value{};
            return s;
        } else {
            return EmptyKey;
        }
    }
    WC_HD static Key probe_key(const insert_probe_type& s) noexcept {
        if constexpr (packed) return s.key;
        else return s;
    }
    WC_D cas_word cas_issue(Slot* p, const insert_probe_type& seen, Key key_in, const Value& value_in) const noexcept {
        if constexpr (packed) {
            Slot want;
            want.key = key_in;
            want.value = value_in;
            return atomic_cas(reinterpret_cast<std::uint64_t*>(p), slot_bits(seen), slot_bits(want));
        } else {
            return atomic_cas(&(p->key), seen, key_in);
        }
    }
    WC_D int cas_outcome(cas_word old, insert_probe_type& seen, Key key_in) const noexcept {
        if constexpr (packed) {
            if (old == slot_bits(seen)) return 1;
            const Key old_key = slot_from_bits<Slot>(old).key;
            if (old_key == EmptyKey) { seen = slot_from_bits<Slot>(old); return 3; }
            return old_key == key_in ? 2 : 0;
        } else {
            return old == seen ? 1 : (old == key_in ? 2 : 0);
        }
    }
    // after insert_batched found the key's slot (single_value_hash_table.cuh:222-225)
    template<class Join>
    WC_D void finish_insert(index_type slot, const Value& value_in, bool published, Status&, Join&) const noexcept {
        if (!published) slots[slot].value = value_in;
    }

    // Finds or claims the slot of `key_in`. Returns the slot index, or ~0 if
    // none (invalid key, probing exhausted). With `publish` and a packed
    // slot the claiming CAS also writes *publish as the value, and
    // `published` is set.
    template<class Join>
    WC_D index_type insert_impl(Key key_in, Status& status_out, const tile<cg_size()>& group,
                                index_type probing_length, Join& join, const Value* publish,
                                bool& published) const noexcept {
        published = false;
        if (!initialized) { status_out = Status::not_initialized(); return ~index_type(0); }
        if (!is_valid_key(key_in)) {
            join(Status::invalid_key());
            status_out = Status::invalid_key();
            return ~index_type(0);
        }
        ProbingScheme iter(geo, probing_length, group);
        index_type step = 0;   // the key's window step, for the overflow filter
        if constexpr (fixed_step_scheme<ProbingScheme>::value) {
            if (filter != nullptr) {
                WindowCursor<ProbingScheme> c;
                c.begin(key_in, seed, geo);
                step = c.step;
            }
        }
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Slot s = load_slot(slots + i);
            const Key table_key = s.key;
            const auto hit_mask = group.ballot(table_key == key_in);
            if (hit_mask) {
                join(Status::duplicate_key());
                status_out = Status::duplicate_key();
                return group.shfl(i, unsigned(ffs64(hit_mask) - 1));
            }
            auto empty_mask = group.ballot(is_empty_key(table_key));
            while (empty_mask) {
                const unsigned leader = unsigned(ffs64(empty_mask) - 1);
                bool success = false;
                bool duplicate = false;
                if (group.thread_rank() == leader) {
                    if constexpr (packed) {
                        std::uint64_t expected = slot_bits(s);
                        if (publish != nullptr) {
                            Slot want;
                            want.key = key_in;
                            want.value = *publish;
                            const std::uint64_t desired = slot_bits(want);
                            while (true) {
                                const std::uint64_t old =
                                    atomic_cas(reinterpret_cast<std::uint64_t*>(slots + i), expected, desired);
                                if (old == expected) { success = true; break; }
                                const Key old_key = slot_from_bits<Slot>(old).key;
                                if (old_key != EmptyKey) { duplicate = old_key == key_in; break; }
                                expected = old;  // value bits of an EMPTY slot moved; retry
                            }
                            published = success;
                        } else {
                            const Key old = atomic_cas(&(slots[i].key), table_key, key_in);
                            success = old == table_key;
                            duplicate = old == key_in;
                        }
                    } else {
                        const Key old = atomic_cas(&(slots[i].key), table_key, key_in);
                        success = old == table_key;
                        duplicate = old == key_in;
                    }
                }
                if (group.any(duplicate)) {
                    join(Status::duplicate_key());
                    status_out = Status::duplicate_key();
                    return group.shfl(i, leader);
                }
                if (group.any(success)) {
                    status_out = Status::none();
                    published = group.any(published);
                    return group.shfl(i, leader);
                }
                empty_mask ^= std::uint64_t(1) << leader;
            }
            // the key moves on past a window without room
            if constexpr (fixed_step_scheme<ProbingScheme>::value)
                if (group.thread_rank() == 0) mark_overflow(*this, i / cg_size(), step);
        }
        join(Status::probing_length_exceeded());
        status_out = Status::probing_length_exceeded();
        return ~index_type(0);
    }

    // single_value_hash_table.cuh:210-228
    template<class Join>
    WC_D Status insert(Key key_in, const Value& value_in, const tile<cg_size()>& group, index_type probing_length,
                       Join& join) const noexcept {
        Status status = Status::unknown_error();
        bool published = false;
        const index_type s = insert_impl(key_in, status, group, probing_length, join, &value_in, published);
        if (group.thread_rank() == 0 && s != ~index_type(0) && !published) slots[s].value = value_in;
        return status;
    }

    // single_value_hash_table.cuh:267-307; the value comes from the same
    // slot read that found the key
    template<class Join>
    WC_D Status retrieve(Key key_in, Value& value_out, const tile<cg_size()>& group, index_type probing_length,
                         Join& join) const noexcept {
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) { join(Status::invalid_key()); return Status::invalid_key(); }
        ProbingScheme iter(geo, probing_length, group);
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Slot s = load_slot(slots + i);
            const auto hit_mask = group.ballot(s.key == key_in);
            if (hit_mask) {
                value_out = group.shfl(s.value, unsigned(ffs64(hit_mask) - 1));
                return Status::none();
            }
            if (group.any(is_empty_key(s.key))) {
                join(Status::key_not_found());
                return Status::key_not_found();
            }
        }
        join(Status::probing_length_exceeded());
        return Status::probing_length_exceeded();
    }

    // single_value_hash_table.cuh:425-468; misses are returned, not joined
    template<class Join>
    WC_D Status erase(Key key_in, const tile<cg_size()>& group, index_type probing_length, Join& join) const noexcept {
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) { join(Status::invalid_key()); return Status::invalid_key(); }
        ProbingScheme iter(geo, probing_length, group);
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Key table_key = slots[i].key;
            const auto hit_mask = group.ballot(table_key == key_in);
            if (hit_mask) {
                if (group.thread_rank() == unsigned(ffs64(hit_mask) - 1)) slots[i].key = TombstoneKey;
                return Status::none();
            }
            if (group.any(is_empty_key(table_key))) return Status::key_not_found();
        }
        return Status::probing_length_exceeded();
    }
};

// Device state of a single-value table over storage::key_value::SoAStore:
// keys and values in two arrays. The per-window protocol is the one of
// SingleValueView (single_value_hash_table.cuh:742-823, 267-307, 425-468);
// a lane reads only its slot's key, the CAS publishes the key and the value
// is stored after it, and a lookup reads the value of the hit slot alone.
// Bulk calls run the one-group-per-key kernels of engine.hpp.
template<class Key, class Value, Key EmptyKey, Key TombstoneKey, class ProbingScheme>
struct SingleValueSoAView {
    using key_type = Key;
    using value_type = Value;
    using slot_type = storage::key_value::pair_t<Key, Value>;
    using index_type = index_t;
    using probing_scheme_type = ProbingScheme;
    using out_type = Value;

    Key* keys;
    Value* values;
    Status* status;
    probing_schemes::Geometry geo;
    Key seed;
    bool initialized;

    WC_HD static constexpr index_type cg_size() noexcept { return ProbingScheme::cg_size(); }
    WC_HD static constexpr bool is_empty_key(Key k) noexcept { return k == EmptyKey; }
    WC_HD static constexpr bool is_tombstone_key(Key k) noexcept { return k == TombstoneKey; }
    WC_HD static constexpr bool is_valid_key(Key k) noexcept { return k != EmptyKey && k != TombstoneKey; }
    WC_HD index_type capacity() const noexcept { return geo.capacity; }
    WC_HD Status* status_word() const noexcept { return status; }
    WC_D Key key_at(index_type i) const noexcept { return keys[i]; }
    WC_D slot_type slot_at(index_type i) const noexcept {
        slot_type s;
        s.key = keys[i];
        s.value = values[i];
        return s;
    }

    // Slot of `key_in` (found or claimed), ~0 if none.
    template<class Join>
    WC_D index_type insert_impl(Key key_in, Status& status_out, const tile<cg_size()>& group,
                                index_type probing_length, Join& join) const noexcept {
        if (!initialized) { status_out = Status::not_initialized(); return ~index_type(0); }
        if (!is_valid_key(key_in)) {
            join(Status::invalid_key());
            status_out = Status::invalid_key();
            return ~index_type(0);
        }
        ProbingScheme iter(geo, probing_length, group);
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Key table_key = keys[i];
            const auto hit_mask = group.ballot(table_key == key_in);
            if (hit_mask) {
                join(Status::duplicate_key());
                status_out = Status::duplicate_key();
                return group.shfl(i, unsigned(ffs64(hit_mask) - 1));
            }
            auto empty_mask = group.ballot(is_empty_key(table_key));
            while (empty_mask) {
                const unsigned leader = unsigned(ffs64(empty_mask) - 1);
                bool success = false;
                bool duplicate = false;
                if (group.thread_rank() == leader) {
                    const Key old = atomic_cas(keys + i, table_key, key_in);
                    success = old == table_key;
                    duplicate = old == key_in;
                }
                if (group.any(duplicate)) {
                    join(Status::duplicate_key());
                    status_out = Status::duplicate_key();
                    return group.shfl(i, leader);
                }
                if (group.any(success)) {
                    status_out = Status::none();
                    return group.shfl(i, leader);
                }
                empty_mask ^= std::uint64_t(1) << leader;
            }
        }
        join(Status::probing_length_exceeded());
        status_out = Status::probing_length_exceeded();
        return ~index_type(0);
    }

    template<class Join>
    WC_D Status insert(Key key_in, const Value& value_in, const tile<cg_size()>& group, index_type probing_length,
                       Join& join) const noexcept {
        Status status = Status::unknown_error();
        const index_type s = insert_impl(key_in, status, group, probing_length, join);
        if (group.thread_rank() == 0 && s != ~index_type(0)) values[s] = value_in;
        return status;
    }

    template<class Join>
    WC_D Status retrieve(Key key_in, Value& value_out, const tile<cg_size()>& group, index_type probing_length,
                         Join& join) const noexcept {
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) { join(Status::invalid_key()); return Status::invalid_key(); }
        ProbingScheme iter(geo, probing_length, group);
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Key table_key = keys[i];
            const auto hit_mask = group.ballot(table_key == key_in);
            if (hit_mask) {
                const unsigned at = unsigned(ffs64(hit_mask) - 1);
                Value v{};
                if (group.thread_rank() == at) v = values[i];
                value_out = group.shfl(v, at);
                return Status::none();
            }
            if (group.any(is_empty_key(table_key))) {
                join(Status::key_not_found());
                return Status::key_not_found();
            }
        }
        join(Status::probing_length_exceeded());
        return Status::probing_length_exceeded();
    }

    template<class Join>
    WC_D Status erase(Key key_in, const tile<cg_size()>& group, index_type probing_length, Join& join) const noexcept {
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) { join(Status::invalid_key()); return Status::invalid_key(); }
        ProbingScheme iter(geo, probing_length, group);
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Key table_key = keys[i];
            const auto hit_mask = group.ballot(table_key == key_in);
            if (hit_mask) {
                if (group.thread_rank() == unsigned(ffs64(hit_mask) - 1)) keys[i] = TombstoneKey;
                return Status::none();
            }
            if (group.any(is_empty_key(table_key))) return Status::key_not_found();
        }
        return Status::probing_length_exceeded();
    }
};

template<class Storage> struct is_soa_store : std::false_type {};
template<class K, class V> struct is_soa_store<storage::key_value::SoAStore<K, V>> : std::true_type {};

}  // namespace detail

template<class Key, class Value, Key EmptyKey = defaults::empty_key<Key>(),
         Key TombstoneKey = defaults::tombstone_key<Key>(),
         class ProbingScheme = defaults::probing_scheme_t<Key, 8>,
         class TableStorage = defaults::table_storage_t<Key, Value>,
         index_t TempMemoryBytes = defaults::temp_memory_bytes()>
class SingleValueHashTable {
    static_assert(checks::is_valid_key_type<Key>(), "invalid key type");
    static_assert(EmptyKey != TombstoneKey, "empty key and tombstone key must not be identical");
    static_assert(checks::is_probing_scheme<ProbingScheme>(), "not a valid probing scheme type");
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
    static constexpr bool soa_storage = detail::is_soa_store<TableStorage>::value;
    using view_type = typename std::conditional<
        soa_storage, detail::SingleValueSoAView<Key, Value, EmptyKey, TombstoneKey, ProbingScheme>,
        detail::SingleValueView<Key, Value, EmptyKey, TombstoneKey, ProbingScheme, slot_type>>::type;

    WC_HD static constexpr key_type empty_key() noexcept { return EmptyKey; }
    WC_HD static constexpr key_type tombstone_key() noexcept { return TombstoneKey; }
    WC_HD static constexpr index_type cg_size() noexcept { return ProbingScheme::cg_size(); }

    // single_value_hash_table.cuh:111-128
    WC_H explicit SingleValueHashTable(index_type min_capacity, key_type seed = defaults::seed<key_type>(),
                                       bool no_init = false) noexcept
        : status_(nullptr),
          table_(detail::get_valid_capacity(min_capacity, cg_size())),
          temp_(TempMemoryBytes / sizeof(index_type)),
          seed_(seed),
          is_initialized_(false),
          is_copy_(false),
          geo_(probing_schemes::Geometry::make(table_.capacity(), cg_size())) {
        if (hipMalloc(&status_, sizeof(status_type)) != hipSuccess) status_ = nullptr;
        if constexpr (with_filter) {
            if (table_.capacity() != 0 && hipMalloc(&filter_, filter_bytes()) != hipSuccess) filter_ = nullptr;
        }
        assign_status(table_.status() + temp_.status());
        if (!no_init) init();
    }

    WC_HD SingleValueHashTable(const SingleValueHashTable& o) noexcept
        : status_(o.status_), table_(o.table_), temp_(o.temp_), seed_(o.seed_),
          is_initialized_(o.is_initialized_), is_copy_(true), geo_(o.geo_), filter_(o.filter_),
          filter_valid_(o.filter_valid_) {}

    WC_H SingleValueHashTable(SingleValueHashTable&& o) noexcept
        : status_(o.status_), table_(std::move(o.table_)), temp_(std::move(o.temp_)), seed_(o.seed_),
          is_initialized_(o.is_initialized_), is_copy_(o.is_copy_), geo_(o.geo_), filter_(o.filter_),
          filter_valid_(o.filter_valid_) {
        o.is_copy_ = true;
    }

    WC_HD ~SingleValueHashTable() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_ && status_ != nullptr) (void)hipFree(status_);
        if (!is_copy_ && filter_ != nullptr) (void)hipFree(filter_);
#endif
    }

    // single_value_hash_table.cuh:175-201
    WC_H void init(key_type seed, hipStream_t stream = 0) noexcept {
        is_initialized_ = false;
        seed_ = seed;
        if (!table_.status().has_not_initialized() && !temp_.status().has_not_initialized()) {
            table_.init_keys(empty_key(), stream);
            if (filter_ != nullptr)
                storage::kernels::fill<std::uint32_t>
                    <<<detail::grid_for(filter_bytes() / 4, 1, 8), block_size(), 0, stream>>>(
                        filter_, filter_bytes() / 4, 0u);
            filter_valid_ = filter_ != nullptr;
            assign_status(table_.status() + temp_.status(), stream);
            is_initialized_ = true;
        }
    }
    WC_H void init(hipStream_t stream = 0) noexcept { init(seed_, stream); }

    // ----------------------------------------------------------- device API
    WC_D status_type insert(key_type key_in, const value_type& value_in, const tile<cg_size()>& group,
                            index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().insert(key_in, value_in, group, probing_length, join);
    }
    WC_D status_type retrieve(key_type key_in, value_type& value_out, const tile<cg_size()>& group,
                              index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().retrieve(key_in, value_out, group, probing_length, join);
    }
    WC_D status_type erase(key_type key_in, const tile<cg_size()>& group,
                           index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().erase(key_in, group, probing_length, join);
    }

    // ------------------------------------------------------------- host API
    // single_value_hash_table.cuh:239-258
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert(const key_type* keys_in, const value_type* values_in, index_type num_in, hipStream_t stream = 0,
                     index_type probing_length = defaults::probing_length(),
                     typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_ || num_in == 0) return;
        if constexpr (soa_storage) {
            kernels::insert<view_type, StatusHandler><<<detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, values_in, num_in, view(), probing_length, status_out);
        } else {
            insert_aos<StatusHandler>(keys_in, values_in, num_in, stream, probing_length, status_out);
        }
    }

    // single_value_hash_table.cuh:318-337
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void retrieve(const key_type* keys_in, index_type num_in, value_type* values_out, hipStream_t stream = 0,
                       index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_ || num_in == 0) return;
        if constexpr (soa_storage) {
            kernels::retrieve<view_type, StatusHandler><<<detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, num_in, values_out, view(), probing_length, status_out);
        } else {
            retrieve_aos<StatusHandler>(keys_in, num_in, values_out, stream, probing_length, status_out);
        }
    }

    // single_value_hash_table.cuh:390-417 (order unspecified)
    WC_H void retrieve_all(key_type* keys_out, value_type* values_out, index_type& num_out,
                           hipStream_t stream = 0) const noexcept {
        if (!is_initialized_) return;
        index_type* tmp = temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        kernels::retrieve_all<view_type>
            <<<detail::grid_for(capacity(), 1, 8), block_size(), 0, stream>>>(keys_out, values_out, tmp, view());
        (void)hipMemcpyAsync(&num_out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
    }

    // single_value_hash_table.cuh:478-496
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void erase(const key_type* keys_in, index_type num_in, hipStream_t stream = 0,
                    index_type probing_length = defaults::probing_length(),
                    typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_ || num_in == 0) return;
        if constexpr (!soa_storage && detail::fixed_step_scheme<typename view_type::probing_scheme_type>::value) {
            if constexpr (detail::owner_capable<view_type>()) {
                if (detail::owner_lookups()) {
                    detail::launch_erase_owner<view_type, StatusHandler>(keys_in, num_in, view(), probing_length,
                                                                         status_out, stream);
                    return;
                }
            }
            if (detail::lockstep_lookups()) {
                detail::launch_erase_lockstep<view_type, StatusHandler, detail::lockstep_keys<unsigned(cg_size())>()>(
                    keys_in, num_in, view(), probing_length, status_out, stream);
                return;
            }
        }
        kernels::erase<view_type, StatusHandler>
            <<<detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(keys_in, num_in, view(),
                                                                                probing_length, status_out);
    }

    // f(key, value) on every stored pair (single_value_hash_table.cuh:504-528)
    template<class Func>
    WC_H void for_each(Func f, hipStream_t stream = 0, index_type smem_bytes = 0) const noexcept {
        if (!is_initialized_) return;
        kernels::for_each<Func, view_type>
            <<<detail::grid_for(capacity(), 1, 8), block_size(), smem_bytes, stream>>>(f, view());
    }

    // single_value_hash_table.cuh:534-592
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

    // slots that are not EMPTY (stored keys and tombstones; inserts never
    // reuse a tombstone, so these are the slots taken since init)
    WC_H index_type claimed_slots(hipStream_t stream = 0) const noexcept {
        if (!is_initialized_) return 0;
        index_type out = 0;
        index_type* tmp = temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        kernels::size<view_type><<<detail::grid_for(capacity(), 1, 8), block_size(), 0, stream>>>(tmp, view(), true);
        (void)hipMemcpyAsync(&out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
        return out;
    }

    WC_H float load_factor(hipStream_t stream = 0) const noexcept { return float(size(stream)) / float(capacity()); }
    WC_H float storage_density(hipStream_t stream = 0) const noexcept { return load_factor(stream); }
    WC_HD index_type capacity() const noexcept { return table_.capacity(); }
    WC_H index_type bytes_total() const noexcept {
        return table_.bytes_total() + temp_.bytes_total() + sizeof(status_type) + (filter_ ? filter_bytes() : 0);
    }
    WC_HD bool is_initialized() const noexcept { return is_initialized_; }

    // single_value_hash_table.cuh:645-687
    WC_H status_type peek_status(hipStream_t stream = 0) const noexcept {
        status_type status = status_type::not_initialized();
        if (status_ != nullptr) {
            (void)hipMemcpyAsync(&status, status_, sizeof(status_type), hipMemcpyDeviceToHost, stream);
            (void)hipStreamSynchronize(stream);
        }
        return status;
    }
    WC_H status_type pop_status(hipStream_t stream = 0) noexcept {
        status_type status = status_type::not_initialized();
        if (status_ != nullptr) {
            (void)hipMemcpyAsync(&status, status_, sizeof(status_type), hipMemcpyDeviceToHost, stream);
            assign_status(table_.status(), stream);
        }
        return status;
    }

    WC_HD static constexpr bool is_empty_key(key_type key) noexcept { return key == empty_key(); }
    WC_HD static constexpr bool is_tombstone_key(key_type key) noexcept { return key == tombstone_key(); }
    WC_HD static constexpr bool is_valid_key(key_type key) noexcept {
        return key != empty_key() && key != tombstone_key();
    }
    WC_HD key_type seed() const noexcept { return seed_; }
    WC_HD bool is_copy() const noexcept { return is_copy_; }

    // Device state handed to kernels.
    WC_HD view_type view() const noexcept {
        if constexpr (soa_storage) {
            return view_type{table_.keys(), table_.values(), status_, geo_, seed_, is_initialized_};
        } else {
            return view_type{table_.data(), status_, geo_, seed_, is_initialized_, filter_valid_ ? filter_ : nullptr};
        }
    }

private:
    // Bulk kernels of AoS tables: lane-per-key windows on small tables,
    // refill (default), lockstep or streaming cooperative kernels above.
    template<class StatusHandler>
    WC_H void insert_aos(const key_type* keys_in, const value_type* values_in, index_type num_in, hipStream_t stream,
                         index_type probing_length, typename StatusHandler::base_type* status_out) noexcept {
        if constexpr (detail::LaneWindow<slot_type, cg_size()>::usable) {
            if (capacity() * sizeof(slot_type) <= detail::lane_insert_table_bytes()) {
                filter_valid_ = false;   // these kernels do not keep the overflow filter
                kernels::insert_lanes<view_type, StatusHandler, detail::keys_per_lane(), true>
                    <<<detail::grid_for_lanes(num_in), block_size(), 0, stream>>>(keys_in, values_in, num_in, view(),
                                                                                 probing_length, status_out);
                return;
            }
        }
        if constexpr (detail::fixed_step_scheme<typename view_type::probing_scheme_type>::value) {
            if (detail::refill_inserts()) {
                detail::launch_insert_refill<view_type, StatusHandler, detail::refill_insert_keys<unsigned(cg_size())>(), true>(
                    keys_in, values_in, num_in, view(), probing_length, status_out, stream);
                return;
            }
            if (detail::lockstep_inserts()) {
                filter_valid_ = false;
                detail::launch_insert_lockstep<view_type, StatusHandler,
                                               detail::lockstep_insert_keys<unsigned(cg_size())>(), true>(
                    keys_in, values_in, num_in, view(), probing_length, status_out, stream);
                return;
            }
        }
        filter_valid_ = false;
        kernels::insert_batched<view_type, StatusHandler, detail::keys_per_group(), true>
            <<<detail::grid_for_batched(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, values_in, num_in, view(), probing_length, status_out);
    }

    template<class StatusHandler>
    WC_H void retrieve_aos(const key_type* keys_in, index_type num_in, value_type* values_out, hipStream_t stream,
                           index_type probing_length, typename StatusHandler::base_type* status_out) const noexcept {
        if constexpr (detail::LaneWindow<slot_type, cg_size()>::usable) {
            if (capacity() * sizeof(slot_type) <= detail::lane_table_bytes()) {
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

    WC_H void assign_status(status_type status, hipStream_t stream = 0) const noexcept {
        if (status_ != nullptr) {
            (void)hipMemcpyAsync(status_, &status, sizeof(status_type), hipMemcpyHostToDevice, stream);
            (void)hipStreamSynchronize(stream);
        }
    }
    WC_H void join_status(status_type status, hipStream_t stream = 0) const noexcept {
        if (status_ != nullptr) {
            const status_type joined = peek_status(stream) + status;
            assign_status(joined, stream);
        }
    }
    WC_D void device_join_status(status_type status) const noexcept {
        if (status_ != nullptr) status_->atomic_join(status);
    }

    // overflow filter: AoS tables probed with a fixed step keep one byte
    // per probe window (every insert path the table uses maintains it; a
    // lookup that meets a full window without its key and with its bit clear
    // ends there, see owner_probing.hpp)
    static constexpr bool with_filter =
        !soa_storage && detail::fixed_step_scheme<ProbingScheme>::value;
    WC_HD index_type filter_bytes() const noexcept { return SDIV(geo_.windows, index_type(4)) * 4; }

    status_type* status_;
    TableStorage table_;
    temp_type temp_;
    key_type seed_;
    bool is_initialized_;
    bool is_copy_;
    probing_schemes::Geometry geo_;
    std::uint32_t* filter_ = nullptr;
    // the filter holds every insert since init (cleared when an insert ran
    // on kernels that do not maintain it; lookups then ignore it)
    mutable bool filter_valid_ = false;

    template<class K_, class V_, K_ E_, K_ T_, class P_, class S_, index_t B_> friend class CountingHashTable;
    template<class K_, class V_, K_ E_, K_ T_, class VS_, class P_> friend class BucketListHashTable;
};

}  // namespace warpcore
