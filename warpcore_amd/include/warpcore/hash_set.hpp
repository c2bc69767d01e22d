// HashSet: a set of keys in one array of key slots (drop-in for
// include/warpcore/hash_set.cuh; default cg 16). Insert uses the same
// window protocol as SingleValueHashTable on bare keys. Membership follows
// the device retrieve of the reference (hash_set.cuh:291-329): a hit gives
// true, a window with an EMPTY slot gives false, both with status none. The
// reference's host retrieve cannot be instantiated (it needs a value_type
// the class does not define); here retrieve(keys, n, flags_out, ...) and its
// alias contains(...) launch that device semantics. Erase reports none
// whatever it finds (hash_set.cuh:395-438).
#pragma once

#include "single_value_hash_table.hpp"

namespace warpcore {

namespace detail {

template<class Key, Key EmptyKey, Key TombstoneKey, class ProbingScheme>
struct HashSetView {
    using key_type = Key;
    using value_type = bool;
    using index_type = index_t;
    struct slot_type { Key key; bool value; };

    Key* keys;
    Status* status;
    probing_schemes::Geometry geo;
    Key seed;
    bool initialized;

    WC_HD static constexpr index_type cg_size() noexcept { return ProbingScheme::cg_size(); }
    WC_HD static constexpr bool is_empty_key(Key k) noexcept { return k == EmptyKey; }
    WC_HD static constexpr bool is_valid_key(Key k) noexcept { return k != EmptyKey && k != TombstoneKey; }
    WC_HD index_type capacity() const noexcept { return geo.capacity; }
    WC_HD Status* status_word() const noexcept { return status; }
    WC_D Key key_at(index_type i) const noexcept { return keys[i]; }
    WC_D slot_type slot_at(index_type i) const noexcept { return slot_type{keys[i], true}; }

    // ---- hooks of the batched kernels (batched_probing.hpp); a slot is a bare key
    using probing_scheme_type = ProbingScheme;
    using out_type = bool;
    using probe_slot_type = Key;
    static constexpr bool publishes_value = false;
    WC_HD bool ready() const noexcept { return initialized; }
    WC_HD Key hash_seed() const noexcept { return seed; }
    WC_HD const probing_schemes::Geometry& geometry() const noexcept { return geo; }
    WC_HD Key* slot_ptr() const noexcept { return keys; }
    WC_HD static Key key_of(Key k) noexcept { return k; }
    WC_HD static bool value_of(Key) noexcept { return true; }
    WC_D bool found(bool) const noexcept { return true; }
    WC_D Status missing(bool& flag) const noexcept { flag = false; return Status::none(); }
    using cas_word = Key;
    using insert_probe_type = Key;
    WC_D static Key load_insert_probe(const Key* p) noexcept { return *p; }
    WC_HD static Key probe_key(Key k) noexcept { return k; }
    WC_HD static Key empty_probe() noexcept { return EmptyKey; }
    WC_D cas_word cas_issue(Key* p, Key seen, Key key_in, const bool&) const noexcept {
        return atomic_cas(p, seen, key_in);
    }
    WC_D int cas_outcome(cas_word old, Key& seen, Key key_in) const noexcept {
        return old == seen ? 1 : (old == key_in ? 2 : 0);
    }
    template<class Join>
    WC_D void finish_insert(index_type, const bool&, bool, Status&, Join&) const noexcept {}

    // hash_set.cuh:192-254
    template<class Join>
    WC_D Status insert(Key key_in, const tile<cg_size()>& group, index_type probing_length, Join& join) const noexcept {
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) { join(Status::invalid_key()); return Status::invalid_key(); }
        ProbingScheme iter(geo, probing_length, group);
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Key table_key = keys[i];
            if (group.any(table_key == key_in)) { join(Status::duplicate_key()); return Status::duplicate_key(); }
            auto empty_mask = group.ballot(is_empty_key(table_key));
            while (empty_mask) {
                const unsigned leader = unsigned(ffs64(empty_mask) - 1);
                bool success = false, duplicate = false;
                if (group.thread_rank() == leader) {
                    const Key old = atomic_cas(keys + i, table_key, key_in);
                    success = old == table_key;
                    duplicate = old == key_in;
                }
                if (group.any(duplicate)) { join(Status::duplicate_key()); return Status::duplicate_key(); }
                if (group.any(success)) return Status::none();
                empty_mask ^= std::uint64_t(1) << leader;
            }
        }
        join(Status::probing_length_exceeded());
        return Status::probing_length_exceeded();
    }

    // hash_set.cuh:291-329
    template<class Join>
    WC_D Status retrieve(Key key_in, bool& flag_out, const tile<cg_size()>& group, index_type probing_length,
                         Join& join) const noexcept {
        if (!initialized) return Status::not_initialized();
        if (!is_valid_key(key_in)) { join(Status::invalid_key()); return Status::invalid_key(); }
        ProbingScheme iter(geo, probing_length, group);
        for (index_type i = iter.begin(key_in, seed); i != iter.end(); i = iter.next()) {
            const Key table_key = keys[i];
            if (group.any(table_key == key_in)) { flag_out = true; return Status::none(); }
            if (group.any(is_empty_key(table_key))) { flag_out = false; return Status::none(); }
        }
        join(Status::probing_length_exceeded());
        return Status::probing_length_exceeded();
    }

    // hash_set.cuh:395-438
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
            if (group.any(is_empty_key(table_key))) return Status::none();
        }
        return Status::none();
    }
};

}  // namespace detail

namespace kernels {
template<class Func, class Core>
__global__ __launch_bounds__(256) void for_each_key(Func f, Core core) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < core.capacity(); i += stride) {
        const auto k = core.key_at(i);
        if (core.is_valid_key(k)) f(k);
    }
}
}  // namespace kernels

template<class Key, Key EmptyKey = defaults::empty_key<Key>(), Key TombstoneKey = defaults::tombstone_key<Key>(),
         class ProbingScheme = defaults::probing_scheme_t<Key, 16>,
         index_t TempMemoryBytes = defaults::temp_memory_bytes()>
class HashSet {
    static_assert(checks::is_valid_key_type<Key>(), "invalid key type");
    static_assert(EmptyKey != TombstoneKey, "empty key and tombstone key must not be identical");
    static_assert(checks::is_probing_scheme<ProbingScheme>(), "not a valid probing scheme type");
    static_assert(std::is_same<typename ProbingScheme::key_type, Key>::value,
                  "probing key type differs from table's key type");
    static_assert(TempMemoryBytes >= sizeof(index_t), "temporary storage must at least be of size index_type");

public:
    using key_type = Key;
    using index_type = index_t;
    using status_type = Status;
    using view_type = detail::HashSetView<Key, EmptyKey, TombstoneKey, ProbingScheme>;

    WC_HD static constexpr key_type empty_key() noexcept { return EmptyKey; }
    WC_HD static constexpr key_type tombstone_key() noexcept { return TombstoneKey; }
    WC_HD static constexpr index_type cg_size() noexcept { return ProbingScheme::cg_size(); }

    // hash_set.cuh:80-110
    WC_H explicit HashSet(index_type min_capacity, key_type seed = defaults::seed<key_type>()) noexcept
        : status_(nullptr), keys_(nullptr),
          capacity_(detail::get_valid_capacity(min_capacity, cg_size())),
          temp_(TempMemoryBytes / sizeof(index_type)), seed_(seed), is_initialized_(false), is_copy_(false),
          geo_(probing_schemes::Geometry::make(capacity_, cg_size())) {
        const std::size_t bytes = sizeof(key_type) * capacity_;
        if (hipMalloc(&status_, sizeof(status_type)) != hipSuccess) { status_ = nullptr; return; }
        if (capacity_ == 0) { assign_status(Status::invalid_configuration() + Status::not_initialized()); return; }
        if (detail::available_gpu_memory() < bytes || hipMalloc(&keys_, bytes) != hipSuccess) {
            keys_ = nullptr;
            assign_status(Status::out_of_memory() + Status::not_initialized());
            return;
        }
        assign_status(Status::none());
        is_initialized_ = true;
        init();
    }
    WC_HD HashSet(const HashSet& o) noexcept
        : status_(o.status_), keys_(o.keys_), capacity_(o.capacity_), temp_(o.temp_), seed_(o.seed_),
          is_initialized_(o.is_initialized_), is_copy_(true), geo_(o.geo_) {}
    WC_H HashSet(HashSet&& o) noexcept
        : status_(o.status_), keys_(o.keys_), capacity_(o.capacity_), temp_(std::move(o.temp_)), seed_(o.seed_),
          is_initialized_(o.is_initialized_), is_copy_(o.is_copy_), geo_(o.geo_) {
        o.is_copy_ = true;
    }
    WC_HD ~HashSet() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_) {
            if (keys_ != nullptr) (void)hipFree(keys_);
            if (status_ != nullptr) (void)hipFree(status_);
        }
#endif
    }

    // hash_set.cuh:142-161
    WC_H void init(key_type seed, hipStream_t stream = 0) noexcept {
        seed_ = seed;
        if (is_initialized_) {
            kernels::memset<key_type><<<detail::grid_for(capacity_, 1, 8), block_size(), 0, stream>>>(
                keys_, capacity_, empty_key());
            assign_status(Status::none(), stream);
        }
    }
    WC_H void init(hipStream_t stream = 0) noexcept { init(seed_, stream); }

    WC_D status_type insert(key_type key_in, const tile<cg_size()>& group,
                            index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().insert(key_in, group, probing_length, join);
    }
    WC_D status_type retrieve(key_type key_in, bool& flag_out, const tile<cg_size()>& group,
                              index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().retrieve(key_in, flag_out, group, probing_length, join);
    }
    WC_D status_type erase(key_type key_in, const tile<cg_size()>& group,
                           index_type probing_length = defaults::probing_length()) const noexcept {
        detail::ImmediateJoin join{status_};
        return view().erase(key_in, group, probing_length, join);
    }

    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert(const key_type* keys_in, index_type num_in, hipStream_t stream = 0,
                     index_type probing_length = defaults::probing_length(),
                     typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_ || num_in == 0) return;
        if constexpr (detail::LaneWindow<key_type, cg_size()>::usable) {
            if (capacity() * sizeof(key_type) <= detail::lane_insert_table_bytes()) {
                kernels::insert_lanes<view_type, StatusHandler, detail::keys_per_lane(), false>
                    <<<detail::grid_for_lanes(num_in), block_size(), 0, stream>>>(keys_in, nullptr, num_in, view(),
                                                                                 probing_length, status_out);
                return;
            }
        }
        if constexpr (detail::fixed_step_scheme<typename view_type::probing_scheme_type>::value) {
            if (detail::refill_inserts()) {
                detail::launch_insert_refill<view_type, StatusHandler, detail::refill_insert_keys<unsigned(cg_size())>(), false>(
                    keys_in, nullptr, num_in, view(), probing_length, status_out, stream);
                return;
            }
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

    // membership per key; flags_out[i] is written when status_out[i] is none
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void retrieve(const key_type* keys_in, index_type num_in, bool* flags_out, hipStream_t stream = 0,
                       index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!is_initialized_ || num_in == 0) return;
        if constexpr (detail::LaneWindow<key_type, cg_size()>::usable) {
            if (capacity() * sizeof(key_type) <= detail::lane_table_bytes()) {
                kernels::retrieve_lanes<view_type, StatusHandler, detail::keys_per_lane()>
                    <<<detail::grid_for_lanes(num_in), block_size(), 0, stream>>>(keys_in, num_in, flags_out, view(),
                                                                                 probing_length, status_out);
                return;
            }
        }
        if constexpr (detail::fixed_step_scheme<typename view_type::probing_scheme_type>::value) {
            if (detail::refill_lookups()) {
                detail::launch_retrieve_refill<view_type, StatusHandler, detail::refill_keys<unsigned(cg_size())>()>(
                    keys_in, num_in, flags_out, view(), probing_length, status_out, stream);
                return;
            }
            if (detail::lockstep_lookups()) {
                detail::launch_retrieve_lockstep<view_type, StatusHandler, detail::lockstep_keys<unsigned(cg_size())>()>(
                    keys_in, num_in, flags_out, view(), probing_length, status_out, stream);
                return;
            }
        }
        kernels::retrieve_batched<view_type, StatusHandler, detail::keys_per_group()>
            <<<detail::grid_for_batched(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, num_in, flags_out, view(), probing_length, status_out);
    }
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void contains(const key_type* keys_in, index_type num_in, bool* flags_out, hipStream_t stream = 0,
                       index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        retrieve<StatusHandler>(keys_in, num_in, flags_out, stream, probing_length, status_out);
    }

    // hash_set.cuh:366-387
    WC_H void retrieve_all(key_type* keys_out, index_type& num_out, hipStream_t stream = 0) const noexcept {
        if (!is_initialized_) return;
        index_type* tmp = temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        kernels::retrieve_all<view_type>
            <<<detail::grid_for(capacity(), 1, 8), block_size(), 0, stream>>>(keys_out, nullptr, tmp, view());
        (void)hipMemcpyAsync(&num_out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
    }

    template<class StatusHandler = defaults::status_handler_t>
    WC_H void erase(const key_type* keys_in, index_type num_in, hipStream_t stream = 0,
                    index_type probing_length = defaults::probing_length(),
                    typename StatusHandler::base_type* status_out = nullptr) noexcept {
        if (!is_initialized_ || num_in == 0) return;
        kernels::erase<view_type, StatusHandler>
            <<<detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(keys_in, num_in, view(),
                                                                                probing_length, status_out);
    }

    template<class Func>
    WC_H void for_each(Func f, hipStream_t stream = 0, index_type smem_bytes = 0) const noexcept {
        if (!is_initialized_) return;
        kernels::for_each_key<Func, view_type>
            <<<detail::grid_for(capacity(), 1, 8), block_size(), smem_bytes, stream>>>(f, view());
    }

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
    WC_H float storage_density(hipStream_t stream = 0) const noexcept { return load_factor(stream); }
    WC_HD index_type capacity() const noexcept { return capacity_; }
    WC_H index_type bytes_total() const noexcept {
        return capacity_ * sizeof(key_type) + temp_.bytes_total() + sizeof(status_type);
    }
    WC_HD bool is_initialized() const noexcept { return is_initialized_; }

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
        if (status_ != nullptr) assign_status(Status::none(), stream);  // hash_set.cuh pop_status
        return status;
    }

    WC_HD static constexpr bool is_empty_key(key_type key) noexcept { return key == empty_key(); }
    WC_HD static constexpr bool is_tombstone_key(key_type key) noexcept { return key == tombstone_key(); }
    WC_HD static constexpr bool is_valid_key(key_type key) noexcept {
        return key != empty_key() && key != tombstone_key();
    }
    WC_HD key_type seed() const noexcept { return seed_; }
    WC_HD bool is_copy() const noexcept { return is_copy_; }
    WC_HD view_type view() const noexcept { return view_type{keys_, status_, geo_, seed_, is_initialized_}; }

private:
    WC_H void assign_status(status_type status, hipStream_t stream = 0) const noexcept {
        if (status_ != nullptr) {
            (void)hipMemcpyAsync(status_, &status, sizeof(status_type), hipMemcpyHostToDevice, stream);
            (void)hipStreamSynchronize(stream);
        }
    }

    status_type* status_;
    key_type* keys_;
    index_type capacity_;
    storage::CyclicStore<index_type> temp_;
    key_type seed_;
    bool is_initialized_;
    bool is_copy_;
    probing_schemes::Geometry geo_;
};

}  // namespace warpcore
