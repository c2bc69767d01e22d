// BucketListHashTable: key -> growing list of value buckets. Drop-in for
// include/warpcore/bucket_list_hash_table.cuh with the value store
// storage::multi_value::BucketListStore (storage.cuh:656-1454).
//
// Keys live in a SingleValueHashTable<Key, handle>; a handle is one 64-bit
// word {state: 2 bits, bucket index, value count, values in the tail
// bucket} updated only by CAS/exchange. Values live in a pool of 8-byte
// cells handed out by atomicAdds on allocation counters. A key's first bucket holds
// min_bucket_size values and has no header; every later bucket is a header
// cell {previous bucket, bucket size} followed by its values, each bucket
// ceil(previous size * grow_factor) cells long, capped at max_bucket_size.
//
// Append (storage.cuh:1113-1284): an uninitialised handle or a full tail
// bucket is first switched to `blocking` by a CAS; the winner allocates the
// bucket, writes the value and publishes the new handle with an exchange;
// other appenders to the same key re-read the handle until it is no longer
// blocking. The winner's allocation path contains no wait, so a blocked
// handle is always released; the re-read loop is also bounded and returns
// unknown_error if that bound is ever reached. Otherwise the CAS bumps the
// count and the tail, and the winner stores the value in its cell.
//
// Allocation: the reference bumps one counter for the whole pool. Device
// atomics on one address are serialised at the memory side (12 ns each on
// MI355X, measured), which bounded inserts at 2.8 G values/s; so pools of
// at least 2^17 cells are cut into up to 256 arenas of at least 2^16 cells,
// each with its own counter on its own 128-byte line. A wavefront allocates
// from its home arena and moves on to the next ones when a request does not
// fit; an allocation fails (out_of_memory) only when it fits in none. A
// smaller pool is one arena and behaves exactly as the reference's.
//
// Values of a key come back tail bucket first (storage.cuh:1292-1345). A
// handle left `full` by an out-of-memory append still yields the values it
// holds; the reference yields none for it while reporting their count.
#pragma once

#include "base.hpp"
#include "defaults.hpp"
#include "engine.hpp"
#include "multi_value_hash_table.hpp"
#include "scan.hpp"
#include "single_value_hash_table.hpp"
#include "status.hpp"
#include "tile.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace warpcore {
namespace storage {
namespace multi_value {

namespace detail {

enum class LinkedListState : std::uint32_t { uninitialized = 0, initialized = 1, blocking = 2, full = 3 };

// One 64-bit word: state | index << 2 | count << (2 + IB) | tail << (2 + IB + CB).
template<index_t IB, index_t CB, index_t SB>
struct alignas(8) BucketListHandle {
    std::uint64_t pack;

    static constexpr std::uint64_t mask(index_t bits) noexcept {
        return bits >= 64 ? ~std::uint64_t(0) : (std::uint64_t(1) << bits) - 1;
    }
    WC_HD static constexpr index_t max_bucket_index() noexcept { return mask(IB); }
    WC_HD static constexpr index_t max_value_count() noexcept { return mask(CB); }
    WC_HD static constexpr index_t max_bucket_size() noexcept { return mask(SB); }

    WC_HD static constexpr BucketListHandle make(LinkedListState state, index_t index, index_t count,
                                                 index_t tail) noexcept {
        return BucketListHandle{std::uint64_t(state) | ((std::uint64_t(index) & mask(IB)) << 2) |
                                ((std::uint64_t(count) & mask(CB)) << (2 + IB)) |
                                ((std::uint64_t(tail) & mask(SB)) << (2 + IB + CB))};
    }
    WC_HD constexpr LinkedListState state() const noexcept { return LinkedListState(pack & 3); }
    WC_HD constexpr index_t bucket_index() const noexcept { return (pack >> 2) & mask(IB); }
    WC_HD constexpr index_t value_count() const noexcept { return (pack >> (2 + IB)) & mask(CB); }
    WC_HD constexpr index_t num_values_tail() const noexcept { return (pack >> (2 + IB + CB)) & mask(SB); }
    WC_HD constexpr bool is_uninitialized() const noexcept { return state() == LinkedListState::uninitialized; }
    WC_HD constexpr bool is_initialized() const noexcept { return state() == LinkedListState::initialized; }
    WC_HD constexpr bool is_blocking() const noexcept { return state() == LinkedListState::blocking; }
    WC_HD constexpr bool is_full() const noexcept { return state() == LinkedListState::full; }
};

// A pool cell: a value, or the header {previous bucket, bucket size} of a bucket.
template<class Value, index_t IB, index_t SB>
union alignas(alignof(Value) > 8 ? alignof(Value) : 8) Bucket {
    Value value;
    std::uint64_t info;

    WC_HD static constexpr std::uint64_t make_info(index_t previous, index_t size) noexcept {
        return (std::uint64_t(previous) & BucketListHandle<IB, 1, SB>::mask(IB)) |
               ((std::uint64_t(size) & BucketListHandle<IB, 1, SB>::mask(SB)) << IB);
    }
    WC_HD static constexpr index_t previous_of(std::uint64_t info) noexcept {
        return info & BucketListHandle<IB, 1, SB>::mask(IB);
    }
    WC_HD static constexpr index_t size_of(std::uint64_t info) noexcept {
        return (info >> IB) & BucketListHandle<IB, 1, SB>::mask(SB);
    }
};

// atomic_add(counter, n) of every active lane of the wavefront with one
// atomic: the lanes' amounts follow each other in lane order after the value
// the counter had. The counter ends where one atomic per lane would leave it
// and every lane's range lies below it exactly as one of those atomics would
// have placed some lane's; only which lane gets which range differs. (A bump
// allocator shared by the whole grid serialises its atomics at the memory
// side; this divides them by the lanes allocating together.)
WC_D index_t wave_atomic_add(index_t* counter, index_t n) noexcept {
    const std::uint64_t active = __builtin_amdgcn_read_exec();
    const unsigned lane = lane_id();
    index_t before = 0, total = 0;
    for (std::uint64_t m = active; m != 0; m &= m - 1) {
        const unsigned l = unsigned(__builtin_ctzll(m));
        const index_t v = (index_t(unsigned(__builtin_amdgcn_readlane(int(n >> 32), int(l)))) << 32) |
                          index_t(unsigned(__builtin_amdgcn_readlane(int(n), int(l))));
        if (l < lane) before += v;
        total += v;
    }
    const unsigned leader = unsigned(__builtin_ctzll(active));
    index_t base = 0;
    if (lane == leader) base = warpcore::detail::atomic_add(counter, total);
    base = (index_t(unsigned(__builtin_amdgcn_readlane(int(base >> 32), int(leader)))) << 32) |
           index_t(unsigned(__builtin_amdgcn_readlane(int(base), int(leader))));
    return base + before;
}

// Device half of the store, passed to kernels by value.
template<class Value, index_t IB, index_t CB, index_t SB>
struct BucketListStoreView {
    using handle_type = BucketListHandle<IB, CB, SB>;
    using bucket_type = Bucket<Value, IB, SB>;
    using value_type = Value;
    // re-reads of a blocked handle before an append gives up (unknown_error)
    static constexpr std::uint32_t kMaxBlockedReads = 1u << 24;

    bucket_type* buckets;
    index_t* counters;    // arena a's cells handed out: counters[a * kCounterStride]
    index_t capacity;
    float grow;
    index_t min_size;
    index_t max_size;
    index_t arenas;       // a power of two; every arena but the last holds arena_cells cells
    index_t arena_cells;

    static constexpr index_t kCounterStride = 16;   // 128 bytes between counters

    WC_HD index_t arena_begin(index_t a) const noexcept { return a * arena_cells; }
    WC_HD index_t arena_size(index_t a) const noexcept {
        return a + 1 < arenas ? arena_cells : capacity - a * arena_cells;
    }

    // `n` cells for every active lane, from the wavefront's home arena or
    // the ones after it; the first cell, or ~0 when no arena holds them.
    // Active lanes take part together (a wave-uniform walk over the arenas).
    WC_D index_t allocate(index_t n) const noexcept {
        const index_t home = arenas == 1 ? 0
            : index_t(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (arenas - 1);
        for (index_t t = 0; t < arenas; ++t) {
            const index_t a = (home + t) & (arenas - 1);
            const index_t at = wave_atomic_add(counters + a * kCounterStride, n);
            if (at + n <= arena_size(a)) return arena_begin(a) + at;
        }
        return ~index_t(0);
    }

    WC_HD static constexpr index_t head() noexcept { return handle_type::max_bucket_index(); }

    // A new bucket's link (info word) read by appends must be visible
    // before the handle that points to it: an append that cannot size the
    // tail from the count (size_by_count) reads the tail's link with a
    // device-scope atomic load once it sees the handle. (Links that the
    // count sizes are read by later launches only and are plain stores.)
    // Both are written
    // with device-scope atomics, which the memory side performs; the link's
    // exchange returns its old word, and waiting for that word before the
    // handle's exchange is issued orders the two at the memory side for
    // every XCD. Values are plain stores, read only by later launches. (A
    // device-scope fence here writes back and invalidates the L2 of every
    // XCD: 2.3x the whole insert's time.)
    WC_D static void wait_for(std::uint64_t returned) noexcept {
        asm volatile("" ::"v"(unsigned(returned)) : "memory");
    }

    // Size of the bucket that holds value number c (c >= 1) of a key:
    // buckets follow one sequence (min_size, then each the previous one
    // grown by `grow`, at most max_size, computed as the allocation below
    // does), so it is found from the count without reading the bucket's link
    // when the sequence settles within kSizeSteps steps; false otherwise.
    static constexpr int kSizeSteps = 64;
    WC_D bool size_by_count(index_t c, index_t& size) const noexcept {
        index_t cum = min_size;
        size = min_size;
        for (int k = 0; k < kSizeSteps && cum < c; ++k) {
            const index_t next = grown(size);
            if (next == size) return true;   // the same size from here on
            size = next;
            cum += size;
        }
        return cum >= c;
    }
    WC_D index_t grown(index_t size) const noexcept { return index_t(fminf(float(max_size), ceilf(float(size) * grow))); }

    WC_D Status append(handle_type* handle, const Value& value, index_t max_values_per_key) const noexcept {
        return append_from(handle, warpcore::detail::atomic_load_relaxed(&handle->pack), value, max_values_per_key);
    }

    // append with the handle as the caller read it (`seen`); a failed CAS
    // continues from the word it returned
    WC_D Status append_from(handle_type* handle, std::uint64_t seen, const Value& value,
                            index_t max_values_per_key) const noexcept {
        using warpcore::detail::atomic_cas;
        using warpcore::detail::atomic_exch;
        using LS = detail::LinkedListState;
        std::uint64_t* hp = &handle->pack;
        for (std::uint32_t reads = 0; reads < kMaxBlockedReads; ++reads) {
            const handle_type cur{seen};
            if (cur.is_blocking()) {   // another lane is linking a bucket: wait a little
                __builtin_amdgcn_s_sleep(1);
                seen = warpcore::detail::atomic_load_relaxed(hp);
                continue;
            }
            if (cur.is_uninitialized()) {
                const handle_type blocked = handle_type::make(LS::blocking, head(), 0, 0);
                seen = atomic_cas(hp, cur.pack, blocked.pack);
                if (seen != cur.pack) continue;
                const index_t alloc = allocate(min_size);
                if (alloc != ~index_t(0)) {
                    buckets[alloc].value = value;
                    atomic_exch(hp, handle_type::make(LS::initialized, alloc, 1, 1).pack);
                    return Status::none();
                }
                atomic_exch(hp, handle_type::make(LS::full, head(), 0, 0).pack);
                return Status::out_of_memory();
            }
            if (cur.is_full()) return Status::out_of_memory();
            if (cur.value_count() == max_values_per_key) return Status::max_values_for_key_reached();
            index_t tail_size = min_size;
            if (cur.value_count() > min_size && !size_by_count(cur.value_count(), tail_size))
                tail_size = bucket_type::size_of(warpcore::detail::atomic_load_relaxed(&buckets[cur.bucket_index()].info));
            if (cur.num_values_tail() == tail_size) {
                const handle_type blocked =
                    handle_type::make(LS::blocking, cur.bucket_index(), cur.value_count(), cur.num_values_tail());
                seen = atomic_cas(hp, cur.pack, blocked.pack);
                if (seen != cur.pack) continue;
                const index_t new_size = grown(tail_size);
                const index_t alloc = allocate(new_size + 1);
                if (alloc != ~index_t(0)) {
                    buckets[alloc + 1].value = value;
                    const std::uint64_t link = bucket_type::make_info(cur.bucket_index(), new_size);
                    index_t known = 0;
                    if (size_by_count(cur.value_count() + 1, known))
                        buckets[alloc].info = link;   // read only by later launches: a plain store
                    else
                        wait_for(atomic_exch(&buckets[alloc].info, link));   // appends will read this link
                    atomic_exch(hp, handle_type::make(LS::initialized, alloc, cur.value_count() + 1, 1).pack);
                    return Status::none();
                }
                atomic_exch(hp, handle_type::make(LS::full, cur.bucket_index(), cur.value_count(),
                                                  cur.num_values_tail())
                                    .pack);
                return Status::out_of_memory();
            }
            const handle_type next = handle_type::make(LS::initialized, cur.bucket_index(), cur.value_count() + 1,
                                                       cur.num_values_tail() + 1);
            seen = atomic_cas(hp, cur.pack, next.pack);
            if (seen == cur.pack) {
                const index_t j = cur.value_count() + 1 <= min_size ? cur.num_values_tail()
                                                                    : cur.num_values_tail() + 1;
                buckets[cur.bucket_index() + j].value = value;
                return Status::none();
            }
        }
        return Status::unknown_error();
    }

    // f(value, j) for j < count, lanes of `group` striding over the values
    template<class Func, unsigned G>
    WC_D void for_each(Func&& f, const handle_type& handle, const tile<G>& group) const noexcept {
        for_each_strided(f, handle, group.thread_rank(), G);
    }

    // One lane's walk of a handle's buckets, tail first, put(value, j) for
    // j < count. Bucket sizes come from the value counts (size_by_count), so
    // each bucket costs one round trip: its link and its values are loaded
    // together, B values at a time, before any of them is put.
    template<unsigned B, class Put>
    WC_D void walk_lane(const handle_type& handle, Put& put) const noexcept {
        if ((!handle.is_initialized() && !handle.is_full()) || handle.bucket_index() == head()) return;
        const index_t count = handle.value_count();
        index_t bucket = handle.bucket_index();
        index_t size = handle.num_values_tail();   // values to read in this bucket
        index_t base = 0;                          // values before this bucket, in walk order
        while (true) {
            const bool headed = count - base > min_size;   // not the first, headerless bucket
            const index_t off = headed ? 1 : 0;
            if (size == 0 || bucket + off + size > capacity) return;   // not on a well-formed chain
            const std::uint64_t info = headed ? buckets[bucket].info : 0;
            for (index_t p = 0; p < size; p += B) {
                Value v[B];
#pragma unroll
                for (unsigned b = 0; b < B; ++b)
                    if (p + b < size) v[b] = buckets[bucket + off + p + b].value;
#pragma unroll
                for (unsigned b = 0; b < B; ++b)
                    if (p + b < size) put(v[b], base + p + b);
            }
            base += size;
            if (base >= count || !headed) return;
            bucket = bucket_type::previous_of(info);
            if (bucket == head() || bucket >= capacity) return;
            if (count - base <= min_size) {
                size = min_size;
            } else if (!size_by_count(count - base, size)) {
                size = bucket_type::size_of(buckets[bucket].info);
            }
        }
    }

    // Lane `rank` of `group_size` lanes takes positions rank, rank + group_size,
    // ... of every bucket; all lanes walk the chain together, so each knows
    // which bucket it is in. (The reference advances lane-local and global
    // indices separately, storage.cuh:1310-1343, which for groups of more
    // than one lane mistakes an earlier bucket for the first, headerless one
    // and then follows a value as a link.)
    template<class Func>
    WC_D void for_each_strided(Func& f, const handle_type& handle, index_t rank, index_t group_size) const
        noexcept {
        if ((!handle.is_initialized() && !handle.is_full()) || handle.bucket_index() == head()) return;
        const index_t count = handle.value_count();
        const index_t tail = handle.num_values_tail();
        index_t bucket = handle.bucket_index();
        const index_t offset = count <= min_size ? 0 : 1;
        for (index_t p = rank; p < tail; p += group_size) f(buckets[bucket + offset + p].value, p);
        index_t base = tail;   // values before the current bucket, in walk order
        while (base < count) {
            bucket = bucket_type::previous_of(buckets[bucket].info);
            const bool last = base >= count - min_size;   // the first, headerless bucket
            const index_t size = last ? min_size : bucket_type::size_of(buckets[bucket].info);
            const index_t off = last ? 0 : 1;
            if (size == 0 || bucket == head() || bucket + off + size > capacity) return;   // not on a well-formed chain
            for (index_t p = rank; p < size; p += group_size) f(buckets[bucket + off + p].value, base + p);
            base += size;
        }
    }
};

}  // namespace detail

template<class Value, index_t BucketIndexBits = 32, index_t ValueCounterBits = 20, index_t BucketSizeBits = 10>
class BucketListStore {
    static_assert(checks::is_valid_value_type<Value>(), "Value type must be trivially copyable.");
    static_assert(BucketIndexBits + ValueCounterBits + BucketSizeBits + 2 <= 64,
                  "Too many bits for bucket index and value counter and bucket size.");
    static_assert(BucketIndexBits + BucketSizeBits <= 64, "bucket header does not fit 64 bits");

public:
    using value_type = Value;
    using handle_type = detail::BucketListHandle<BucketIndexBits, ValueCounterBits, BucketSizeBits>;
    using bucket_type = detail::Bucket<Value, BucketIndexBits, BucketSizeBits>;
    using view_type = detail::BucketListStoreView<Value, BucketIndexBits, ValueCounterBits, BucketSizeBits>;
    using index_type = index_t;
    using status_type = Status;
    using tag = tags::dynamic_value_storage;

    WC_HD static constexpr index_type bucket_index_bits() noexcept { return BucketIndexBits; }
    WC_HD static constexpr index_type value_counter_bits() noexcept { return ValueCounterBits; }
    WC_HD static constexpr index_type bucket_size_bits() noexcept { return BucketSizeBits; }
    WC_HD static constexpr index_type head() noexcept { return handle_type::max_bucket_index(); }
    WC_HD static constexpr handle_type uninitialized_handle() noexcept {
        return handle_type::make(detail::LinkedListState::uninitialized, head(), 0, 0);
    }
    WC_HD static constexpr index_type size(const handle_type& handle) noexcept { return handle.value_count(); }

    // storage.cuh:989-1029
    WC_H explicit BucketListStore(index_type max_capacity, float bucket_grow_factor = 1.1f,
                                  index_type min_bucket_size = 1,
                                  index_type max_bucket_size = handle_type::max_bucket_size()) noexcept
        : status_(Status::not_initialized()), capacity_(max_capacity), bucket_grow_factor_(bucket_grow_factor),
          min_bucket_size_(min_bucket_size), max_bucket_size_(max_bucket_size), buckets_(nullptr),
          counters_(nullptr), arenas_(arenas_for(max_capacity)), is_copy_(false) {
        if (capacity_ < handle_type::max_bucket_index() && bucket_grow_factor_ >= 1.0f && min_bucket_size_ >= 1 &&
            max_bucket_size_ >= min_bucket_size_ && max_bucket_size_ <= handle_type::max_bucket_size()) {
            const std::size_t bytes = sizeof(bucket_type) * capacity_ + counter_bytes();
            if (warpcore::detail::available_gpu_memory() >= bytes &&
                (capacity_ == 0 || hipMalloc(&buckets_, sizeof(bucket_type) * capacity_) == hipSuccess) &&
                hipMalloc(&counters_, counter_bytes()) == hipSuccess) {
                status_ = Status::none();
                init();
            } else {
                status_ += Status::out_of_memory();
            }
        } else {
            status_ += Status::invalid_configuration();
        }
    }
    WC_HD BucketListStore(const BucketListStore& o) noexcept
        : status_(o.status_), capacity_(o.capacity_), bucket_grow_factor_(o.bucket_grow_factor_),
          min_bucket_size_(o.min_bucket_size_), max_bucket_size_(o.max_bucket_size_), buckets_(o.buckets_),
          counters_(o.counters_), arenas_(o.arenas_), is_copy_(true) {}
    WC_H BucketListStore(BucketListStore&& o) noexcept
        : status_(o.status_), capacity_(o.capacity_), bucket_grow_factor_(o.bucket_grow_factor_),
          min_bucket_size_(o.min_bucket_size_), max_bucket_size_(o.max_bucket_size_), buckets_(o.buckets_),
          counters_(o.counters_), arenas_(o.arenas_), is_copy_(o.is_copy_) {
        o.is_copy_ = true;
    }
    WC_HD ~BucketListStore() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_) {
            if (buckets_ != nullptr) (void)hipFree(buckets_);
            if (counters_ != nullptr) (void)hipFree(counters_);
        }
#endif
    }

    // storage.cuh:1063-1090: every cell becomes a header {head, min size};
    // also empties the pool (every arena's counter = 0, by a fill kernel)
    WC_H void init(hipStream_t stream = 0) noexcept {
        if (status_.has_any() || counters_ == nullptr) return;
        if (capacity_ != 0)
            kernels::fill<std::uint64_t><<<warpcore::detail::grid_for(capacity_, 1, 8), block_size(), 0, stream>>>(
                reinterpret_cast<std::uint64_t*>(buckets_), capacity_ * (sizeof(bucket_type) / 8),
                bucket_type::make_info(head(), min_bucket_size_));
        kernels::fill<index_type><<<1, block_size(), 0, stream>>>(counters_, counter_bytes() / sizeof(index_type),
                                                                  index_type(0));
        status_ = Status::none();
    }

    static constexpr index_type kMaxArenas = 256;
    static constexpr index_type kMinArenaCells = index_type(1) << 16;
    // arenas of a pool: the largest power of two up to kMaxArenas leaving
    // every arena at least kMinArenaCells cells
    WC_HD static constexpr index_type arenas_for(index_type cap) noexcept {
        index_type a = 1;
        while (a < kMaxArenas && cap / (2 * a) >= kMinArenaCells) a *= 2;
        return a;
    }
    WC_HD index_type arenas() const noexcept { return arenas_; }
    WC_HD std::size_t counter_bytes() const noexcept {
        return arenas_ == 1 ? sizeof(index_type) : arenas_ * view_type::kCounterStride * sizeof(index_type);
    }

    WC_D status_type append(handle_type* handle, const value_type& value, index_type max_values_per_key) const
        noexcept {
        return view().append(handle, value, max_values_per_key);
    }

    // f(value, j) for j < size(handle), lanes of `group` striding over the
    // values (storage.cuh:1292-1345)
    template<class Func, unsigned G>
    WC_D void for_each(Func&& f, const handle_type& handle, const tile<G>& group) const noexcept {
        view().for_each(f, handle, group);
    }
    template<class Func>
    WC_D void for_each(Func&& f, const handle_type& handle) const noexcept {
        view().for_each(f, handle, tile<1>{});
    }

    WC_HD view_type view() const noexcept {
        return view_type{buckets_, counters_, capacity_, bucket_grow_factor_, min_bucket_size_, max_bucket_size_,
                         arenas_, capacity_ / arenas_};
    }

    WC_HD status_type status() const noexcept { return status_; }
    WC_HD index_type capacity() const noexcept { return capacity_; }
    WC_H index_type bytes_total() const noexcept { return capacity_ * sizeof(bucket_type) + counter_bytes(); }
    // cells handed out: with one arena its counter, as the reference's
    // next_free (it also counts the cells of failed allocations); with
    // several, each arena's counter up to the arena's size (an arena a
    // request did not fit counts as used up to its end)
    WC_H index_type next_free(hipStream_t stream = 0) const noexcept {
        if (counters_ == nullptr) return 0;
        std::vector<index_type> c(counter_bytes() / sizeof(index_type));
        (void)hipMemcpyAsync(c.data(), counters_, counter_bytes(), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
        if (arenas_ == 1) return c[0];
        const view_type v = view();
        index_type n = 0;
        for (index_type a = 0; a < arenas_; ++a) n += std::min(c[a * view_type::kCounterStride], v.arena_size(a));
        return n;
    }
    // storage.cuh:1370-1380
    WC_H float load_factor(hipStream_t stream = 0) const noexcept {
        return float(next_free(stream)) / float(capacity());
    }
    WC_H index_type bytes_occupied(hipStream_t stream = 0) const noexcept {
        return next_free(stream) * sizeof(bucket_type);
    }
    WC_HD float bucket_grow_factor() const noexcept { return bucket_grow_factor_; }
    WC_HD index_type min_bucket_size() const noexcept { return min_bucket_size_; }
    WC_HD index_type max_bucket_size() const noexcept { return max_bucket_size_; }
    WC_HD bool is_copy() const noexcept { return is_copy_; }

private:
    status_type status_;
    index_type capacity_;
    float bucket_grow_factor_;
    index_type min_bucket_size_;
    index_type max_bucket_size_;
    bucket_type* buckets_;
    index_type* counters_;
    index_type arenas_;
    bool is_copy_;
};

}  // namespace multi_value
}  // namespace storage

namespace defaults {
template<class Value> using value_storage_t = storage::multi_value::BucketListStore<Value>;
}  // namespace defaults

namespace detail {

// Device state of a bucket-list table: the key table's view plus the store.
template<class KeyView, class Store>
struct BucketListView {
    using key_type = typename KeyView::key_type;
    using value_type = typename Store::value_type;
    using handle_type = typename Store::handle_type;
    KeyView keys;
    typename Store::view_type store;
    index_t max_values;

    WC_HD static constexpr index_t cg_size() noexcept { return KeyView::cg_size(); }
    WC_HD Status* status_word() const noexcept { return keys.status_word(); }

    template<class Join>
    WC_D Status insert(key_type key_in, const value_type& value_in, const tile<cg_size()>& group,
                       index_t probing_length, Join& join) const noexcept {
        Status status = Status::unknown_error();
        bool published = false;
        const index_t slot = keys.insert_impl(key_in, status, group, probing_length, join,
                                              static_cast<const handle_type*>(nullptr), published);
        if (slot == ~index_t(0)) return status;
        handle_type* handle = &keys.slots[slot].value;
        index_t count = 0;
        if (group.thread_rank() == 0) count = handle_type{atomic_load_relaxed(&handle->pack)}.value_count();
        count = group.shfl(count, 0);
        if (count >= max_values) {
            join(Status::max_values_for_key_reached());
            return status + Status::max_values_for_key_reached();
        }
        status_base_t appended = 0;
        if (group.thread_rank() == 0) {
            const Status a = store.append(handle, value_in, max_values);
            if (a.has_any()) join(a);
            appended = a.base();
        }
        return status + Status::from_base(group.shfl(appended, 0));
    }

    // handle of key_in from the key table (status none when found)
    template<class Join>
    WC_D Status find(key_type key_in, handle_type& handle, const tile<cg_size()>& group, index_t probing_length,
                     Join& join) const noexcept {
        return keys.retrieve(key_in, handle, group, probing_length, join);
    }
};

// Key pass of a bucket-list insert on the lane-owner kernel: finds or
// claims each key's slot in the key table and records the slot of every
// input pair (slot_of; it stays ~0 where the key failed) for the append
// pass. The pair whose compare-and-swap claims a new key also makes that
// key's first append here: its first bucket (min_size cells, from the
// wavefront's arena) gets the pair's value and the handle becomes
// {initialized, bucket, 1, 1}, as the append's uninitialised path leaves
// it; the handle is a plain store, read by the append pass only (a later
// launch), and the pair is marked kSlotAppended. If no arena holds the
// bucket the handle becomes {full, head, 0, 0}, again as that path leaves
// it, and the pair is marked kSlotAppendFailed (the append pass reports its
// out_of_memory).
constexpr std::uint32_t kSlotAppended = ~1u;       // slot_of marks: appended by the key pass,
constexpr std::uint32_t kSlotAppendFailed = ~2u;   // or its allocation failed there

template<class KeyView, class StoreView>
struct BucketListKeyCore : KeyView {
    std::uint32_t* slot_of;
    const typename StoreView::value_type* values;   // the pairs' values, by position
    StoreView store;
    index_t max_values;
    template<class Join>
    WC_D void finish_insert(index_t, const typename KeyView::value_type&, bool, Status&, Join&) const noexcept {}
    WC_D void on_slot(index_t slot, index_t pos, bool claimed) const noexcept {
        if (claimed && max_values > 0) {
            using LS = storage::multi_value::detail::LinkedListState;
            using handle_type = typename KeyView::value_type;
            const index_t at = store.allocate(store.min_size);
            if (at != ~index_t(0)) {
                store.buckets[at].value = values[pos];
                this->slots[slot].value = handle_type::make(LS::initialized, at, 1, 1);
                slot_of[pos] = kSlotAppended;
            } else {
                this->slots[slot].value = handle_type::make(LS::full, StoreView::head(), 0, 0);
                slot_of[pos] = kSlotAppendFailed;
            }
            return;
        }
        slot_of[pos] = std::uint32_t(slot);
    }
};

}  // namespace detail

namespace kernels {

// Append pass of a bucket-list insert: one lane per pair, the pair's value
// appended to the handle of the slot the key pass recorded (the appends of
// a wavefront are independent); the pair's status is the key pass's
// (key_status, or none) + the append's (bucket_list_hash_table.cuh:234-339
// per pair).
template<class View, class StatusHandler>
__global__ __launch_bounds__(256) void bl_append(const typename View::value_type* __restrict__ values_in,
                                                 index_t num_in, const std::uint32_t* __restrict__ slot_of,
                                                 const Status* __restrict__ key_status, View view,
                                                 typename StatusHandler::base_type* __restrict__ status_out) {
    using handle_type = typename View::handle_type;
    warpcore::detail::DeferredJoin join;
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num_in; i += stride) {
        const std::uint32_t slot = slot_of[i];
        Status st = key_status != nullptr ? key_status[i] : Status::none();
        if (slot == detail::kSlotAppendFailed) {
            join(Status::out_of_memory());
            st = st + Status::out_of_memory();
        } else if (slot < detail::kSlotAppendFailed) {   // not failed, not appended by the key pass
            handle_type* h = &view.keys.slots[slot].value;
            const std::uint64_t seen = warpcore::detail::atomic_load_relaxed(&h->pack);
            Status a = Status::max_values_for_key_reached();
            if (handle_type{seen}.value_count() < view.max_values)
                a = view.store.append_from(h, seen, values_in[i], view.max_values);
            if (a.has_any()) join(a);
            st = st + a;
        }
        StatusHandler::handle(st, status_out, i);
    }
    join.flush(view.keys.status_word());
}

template<class View, class StatusHandler>
__global__ __launch_bounds__(256) void bl_insert(const typename View::key_type* __restrict__ keys_in,
                                                 const typename View::value_type* __restrict__ values_in,
                                                 index_t num_in, View view, index_t probing_length,
                                                 typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    warpcore::detail::DeferredJoin join;
    const warpcore::detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        const Status st = view.insert(keys_in[gid], values_in[gid], group, probing_length, join);
        if (group.thread_rank() == 0) StatusHandler::handle(st, status_out, gid);
    }
    join.flush(view.status_word());
}

// per-key value counts and their total (gpu_engine.cuh:466-504)
template<class View, class StatusHandler>
__global__ __launch_bounds__(256) void bl_num_values(const typename View::key_type* __restrict__ keys_in,
                                                     index_t num_in, index_t* __restrict__ total_out,
                                                     index_t* __restrict__ num_per_key_out, View view,
                                                     index_t probing_length,
                                                     typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    warpcore::detail::DeferredJoin join;
    warpcore::detail::DeferredCount total;
    const warpcore::detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        typename View::handle_type h{0};
        const Status st = view.find(keys_in[gid], h, group, probing_length, join);
        const index_t num = st.has_any() ? 0 : h.value_count();
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
__global__ __launch_bounds__(256) void bl_retrieve(const typename View::key_type* __restrict__ keys_in,
                                                   index_t num_in, const index_t* __restrict__ begin_in,
                                                   const index_t* __restrict__ end_in,
                                                   typename View::value_type* __restrict__ values_out, View view,
                                                   index_t probing_length,
                                                   typename StatusHandler::base_type* __restrict__ status_out) {
    using V = typename View::value_type;
    constexpr unsigned CG = unsigned(View::cg_size());
    const tile<CG> group;
    warpcore::detail::DeferredJoin join;
    const warpcore::detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        const index_t begin = begin_in[gid];
        const index_t room = end_in[gid] - begin;
        V* out = values_out + begin;
        typename View::handle_type h{0};
        Status st = view.find(keys_in[gid], h, group, probing_length, join);
        index_t num = 0;
        if (!st.has_any()) {
            view.store.for_each([=](const V& v, index_t j) { if (j < room) out[j] = v; }, h, group);
            num = h.value_count();
        }
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

// Snapshot retrieve, second step: counts[i] = value count of the handle
// the key lookup wrote to handles[i] (0 for a key it did not find)
template<class Handle>
__global__ __launch_bounds__(256) void bl_counts(const std::uint64_t* __restrict__ handles, index_t num_in,
                                                 index_t* __restrict__ counts) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num_in; i += stride)
        counts[i] = Handle{handles[i]}.value_count();
}

// Snapshot retrieve, last step: one lane per key walks the buckets of the
// handle in begin_io[i] (kept there by the lookup), replaces it by the key's
// first output position (end - count) and writes the values tail bucket
// first. A lane per key keeps 64 chains in flight per wavefront.
template<class View>
__global__ __launch_bounds__(256) void bl_walk(index_t num_in, index_t* __restrict__ begin_io,
                                               const index_t* __restrict__ end_in,
                                               typename View::value_type* __restrict__ values_out, View view) {
    using V = typename View::value_type;
    using handle_type = typename View::handle_type;
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num_in; i += stride) {
        const handle_type h{begin_io[i]};
        const index_t count = h.value_count();
        const index_t begin = end_in[i] - count;
        begin_io[i] = begin;
        V* out = values_out + begin;
        auto put = [=](const V& v, index_t j) { if (j < count) out[j] = v; };
        view.store.template walk_lane<4>(h, put);
    }
}

// sum of the value counts of all keys (bucket_list_hash_table.cuh:805-822)
template<class View>
__global__ __launch_bounds__(256) void bl_total_values(index_t* __restrict__ num_out, View view) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    index_t count = 0;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < view.keys.capacity(); i += stride) {
        const auto s = view.keys.slot_at(i);
        if (view.keys.is_valid_key(s.key)) count += s.value.value_count();
    }
    count = tile<64>::reduce_add(count);
    if (lane_id() == 0 && count) warpcore::detail::atomic_add(num_out, count);
}

}  // namespace kernels

template<class Key, class Value, Key EmptyKey = defaults::empty_key<Key>(),
         Key TombstoneKey = defaults::tombstone_key<Key>(), class ValueStore = defaults::value_storage_t<Value>,
         class ProbingScheme = defaults::probing_scheme_t<Key, 8>>
class BucketListHashTable {
    static_assert(checks::is_value_storage<ValueStore>(), "not a valid storage type");

public:
    using handle_type = typename ValueStore::handle_type;

private:
    using hash_table_type = SingleValueHashTable<Key, handle_type, EmptyKey, TombstoneKey, ProbingScheme>;
    using value_store_type = ValueStore;

public:
    using key_type = Key;
    using value_type = Value;
    using index_type = index_t;
    using status_type = Status;
    using view_type = detail::BucketListView<typename hash_table_type::view_type, ValueStore>;

    WC_HD static constexpr key_type empty_key() noexcept { return EmptyKey; }
    WC_HD static constexpr key_type tombstone_key() noexcept { return TombstoneKey; }
    WC_HD static constexpr bool is_valid_key(key_type key) noexcept {
        return key != empty_key() && key != tombstone_key();
    }
    WC_HD static constexpr index_type cg_size() noexcept { return hash_table_type::cg_size(); }
    WC_HD static constexpr index_type max_bucket_size() noexcept { return handle_type::max_bucket_size(); }

    // bucket_list_hash_table.cuh:106-125
    WC_H explicit BucketListHashTable(index_type key_capacity, index_type value_capacity,
                                      key_type seed = defaults::seed<key_type>(), float grow_factor = 1.1f,
                                      index_type min_bucket_size = 1, index_type max_bucket_size_ = max_bucket_size(),
                                      index_type max_values_per_key = handle_type::max_value_count(),
                                      bool no_init = false) noexcept
        : hash_table_(key_capacity, seed, true),
          value_store_(value_capacity, grow_factor, min_bucket_size, max_bucket_size_),
          max_values_per_key_(std::min(max_values_per_key, handle_type::max_value_count())), is_copy_(false),
          scan_parts_(nullptr) {
        if (hipMalloc(&scan_parts_, sizeof(index_type) * warpcore::detail::kScanParts) != hipSuccess) {
            scan_parts_ = nullptr;
            join_status(Status::out_of_memory());
        }
        join_status(value_store_.status());
        if (!no_init) init(seed);
    }
    WC_HD BucketListHashTable(const BucketListHashTable& o) noexcept
        : hash_table_(o.hash_table_), value_store_(o.value_store_), max_values_per_key_(o.max_values_per_key_),
          is_copy_(true), scan_parts_(o.scan_parts_) {}
    WC_H BucketListHashTable(BucketListHashTable&& o) noexcept
        : hash_table_(std::move(o.hash_table_)), value_store_(std::move(o.value_store_)),
          max_values_per_key_(o.max_values_per_key_), is_copy_(o.is_copy_), scan_parts_(o.scan_parts_) {
        o.is_copy_ = true;
    }
    WC_HD ~BucketListHashTable() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_ && scan_parts_ != nullptr) (void)hipFree(scan_parts_);
        if (!is_copy_ && scratch_ != nullptr) (void)hipFree(scratch_);
#endif
    }

    // :155-171
    WC_H void init(key_type seed, hipStream_t stream = 0) noexcept {
        const status_type status = hash_table_.peek_status(stream);
        if (!status.has_not_initialized()) {
            hash_table_.init(seed, stream);
            value_store_.init(stream);
            hash_table_.table_.init_pairs(empty_key(), ValueStore::uninitialized_handle(), stream);
        }
    }
    WC_H void init(hipStream_t stream = 0) noexcept { init(hash_table_.seed(), stream); }

    // ---- device calls
    WC_D status_type insert(key_type key_in, const value_type& value_in, const tile<cg_size()>& group,
                            index_type probing_length = defaults::probing_length()) const noexcept {
        warpcore::detail::ImmediateJoin join{hash_table_.status_};
        return view().insert(key_in, value_in, group, probing_length, join);
    }
    WC_D status_type retrieve(key_type key_in, value_type* values_out, index_type& num_out,
                              const tile<cg_size()>& group,
                              index_type probing_length = defaults::probing_length()) const noexcept {
        warpcore::detail::ImmediateJoin join{hash_table_.status_};
        handle_type h{0};
        const status_type st = view().find(key_in, h, group, probing_length, join);
        if (!st.has_any()) {
            value_store_.for_each([=](const value_type& v, index_type j) { values_out[j] = v; }, h, group);
            num_out = h.value_count();
        } else {
            num_out = 0;
        }
        return st;
    }
    template<class Func>
    WC_D status_type for_each(Func f, key_type key_in, const tile<cg_size()>& group,
                              index_type probing_length = defaults::probing_length()) const noexcept {
        warpcore::detail::ImmediateJoin join{hash_table_.status_};
        handle_type h{0};
        const status_type st = view().find(key_in, h, group, probing_length, join);
        if (!st.has_any()) value_store_.for_each(f, h, group);
        return st;
    }
    WC_D status_type num_values(key_type key_in, index_type& num_out, const tile<cg_size()>& group,
                                index_type probing_length = defaults::probing_length()) const noexcept {
        warpcore::detail::ImmediateJoin join{hash_table_.status_};
        handle_type h{0};
        const status_type st = view().find(key_in, h, group, probing_length, join);
        num_out = st.has_any() ? 0 : h.value_count();
        return st;
    }

    // ---- host calls
    // :188-292 (status per pair = key-table status + append status)
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void insert(const key_type* keys_in, const value_type* values_in, index_type num_in,
                     hipStream_t stream = 0, index_type probing_length = defaults::probing_length(),
                     typename StatusHandler::base_type* status_out = nullptr) noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!hash_table_.is_initialized_ || num_in == 0) return;
        using key_view = typename hash_table_type::view_type;
        using key_core = detail::BucketListKeyCore<key_view, typename value_store_type::view_type>;
        if constexpr (warpcore::detail::fixed_step_scheme<typename key_view::probing_scheme_type>::value &&
                      warpcore::detail::owner_insert_capable<key_core>()) {
            // key pass on the lane-owner kernel, then one lane per pair appends
            // (in chunks of the table's scratch: each pair's slot and key status)
            constexpr index_type chunk = index_type(1) << 24;
            const index_type cap = num_in < chunk ? num_in : chunk;
            // (slot indices below detail::kSlotAppendFailed fit slot_of)
            if (warpcore::detail::owner_inserts() && hash_table_.capacity() < index_type(detail::kSlotAppendFailed) &&
                reserve_scratch(cap)) {
                constexpr bool statuses = !std::is_same<StatusHandler, status_handlers::ReturnNothing>::value;
                std::uint32_t* slot_of = scratch_;
                Status* key_status = reinterpret_cast<Status*>(scratch_ + scratch_cap_);
                for (index_type first = 0; first < num_in; first += cap) {
                    const index_type n = num_in - first < cap ? num_in - first : cap;
                    storage::kernels::fill<std::uint32_t>
                        <<<warpcore::detail::grid_for(n, 1, 8), block_size(), 0, stream>>>(slot_of, n, ~0u);
                    const key_core core{hash_table_.view(), slot_of, values_in + first, value_store_.view(),
                                        max_values_per_key_};
                    Status* ks = statuses && status_out != nullptr ? key_status : nullptr;
                    if (ks != nullptr)
                        warpcore::detail::launch_insert_owner<key_core, status_handlers::ReturnStatus, false>(
                            keys_in + first, nullptr, n, core, probing_length, ks, stream, nullptr);
                    else
                        warpcore::detail::launch_insert_owner<key_core, status_handlers::ReturnNothing, false>(
                            keys_in + first, nullptr, n, core, probing_length, nullptr, stream, nullptr);
                    typename StatusHandler::base_type* out = nullptr;
                    if constexpr (statuses)
                        if (status_out != nullptr) out = status_out + first;
                    kernels::bl_append<view_type, StatusHandler>
                        <<<warpcore::detail::grid_for(n, 1, 32), block_size(), 0, stream>>>(
                            values_in + first, n, slot_of, ks, view(), out);
                }
                if (stream == 0) (void)hipStreamSynchronize(stream);
                return;
            }
        }
        kernels::bl_insert<view_type, StatusHandler>
            <<<warpcore::detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(
                keys_in, values_in, num_in, view(), probing_length, status_out);
        if (stream == 0) (void)hipStreamSynchronize(stream);
    }

    // :333-424, the same two passes as MultiValueHashTable::retrieve
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void retrieve(const key_type* keys_in, index_type num_in, index_type* begin_offsets_out,
                       index_type* end_offsets_out, value_type* values_out, index_type& num_out,
                       hipStream_t stream = 0, index_type probing_length = defaults::probing_length(),
                       typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!hash_table_.is_initialized_) return;
        if (num_in > index_type(std::numeric_limits<int>::max())) {
            join_status(Status::index_overflow(), stream);
            return;
        }
        if (num_in == 0) {
            num_out = 0;
            return;
        }
        if constexpr (warpcore::detail::owner_capable<typename hash_table_type::view_type>()) {
            if (values_out != nullptr && scan_parts_ != nullptr && warpcore::detail::owner_lookups()) {
                retrieve_snapshot<StatusHandler>(keys_in, num_in, begin_offsets_out, end_offsets_out, values_out,
                                                 num_out, stream, probing_length, status_out);
                return;
            }
        }
        num_values(keys_in, num_in, num_out, end_offsets_out, stream, probing_length);
        if (values_out != nullptr) {
            if (scan_parts_ == nullptr) {
                join_status(Status::out_of_memory(), stream);
                num_out = 0;
                return;
            }
            warpcore::detail::device_scan(end_offsets_out, end_offsets_out, num_in, scan_parts_, true, stream);
            kernels::shift_offsets<index_type><<<warpcore::detail::grid_for(num_in, 1, 8), block_size(), 0, stream>>>(
                begin_offsets_out, end_offsets_out, num_in);
            kernels::bl_retrieve<view_type, StatusHandler>
                <<<warpcore::detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(
                    keys_in, num_in, begin_offsets_out, end_offsets_out, values_out, view(), probing_length,
                    status_out);
        } else {
            if (status_out != nullptr)
                kernels::handle_all<StatusHandler>
                    <<<warpcore::detail::grid_for(num_in, 1, 8), block_size(), 0, stream>>>(Status::dry_run(),
                                                                                             status_out, num_in);
            join_status(Status::dry_run(), stream);
        }
        if (stream == 0) (void)hipStreamSynchronize(stream);
    }

    // retrieve with the lane-owner key lookup: each key's handle is read once
    // (into begin_offsets_out), its count becomes end_offsets_out, the scan
    // makes the counts end offsets, and one lane per key walks its buckets.
    // Counts and values come from the one handle read, so the two never
    // disagree (no invalid_phase_overlap); statuses are the lookup's.
    template<class StatusHandler>
    WC_H void retrieve_snapshot(const key_type* keys_in, index_type num_in, index_type* begin_offsets_out,
                                index_type* end_offsets_out, value_type* values_out, index_type& num_out,
                                hipStream_t stream, index_type probing_length,
                                typename StatusHandler::base_type* status_out) const noexcept {
        static_assert(sizeof(handle_type) == sizeof(index_type), "handles are kept in the begin offsets");
        // a key the lookup does not find keeps the word 0: an uninitialised handle of no values
        storage::kernels::fill<index_type><<<warpcore::detail::grid_for(num_in, 1, 8), block_size(), 0, stream>>>(
            begin_offsets_out, num_in, index_type(0));
        warpcore::detail::launch_retrieve_owner<typename hash_table_type::view_type, StatusHandler>(
            keys_in, num_in, reinterpret_cast<handle_type*>(begin_offsets_out), hash_table_.view(), probing_length,
            status_out, stream);
        kernels::bl_counts<handle_type><<<warpcore::detail::grid_for(num_in, 1, 8), block_size(), 0, stream>>>(
            reinterpret_cast<const std::uint64_t*>(begin_offsets_out), num_in, end_offsets_out);
        warpcore::detail::device_scan(end_offsets_out, end_offsets_out, num_in, scan_parts_, true, stream);
        kernels::bl_walk<view_type><<<warpcore::detail::grid_for(num_in, 1, 32), block_size(), 0, stream>>>(
            num_in, begin_offsets_out, end_offsets_out, values_out, view());
        num_out = 0;
        (void)hipMemcpyAsync(&num_out, end_offsets_out + (num_in - 1), sizeof(index_type), hipMemcpyDeviceToHost,
                             stream);
        (void)hipStreamSynchronize(stream);
    }

    // :460-488; num_values_out is index_type& (see MultiValueHashTable)
    WC_H void retrieve_all(key_type* keys_out, index_type& num_keys_out, index_type* begin_offsets_out,
                           index_type* end_offsets_out, value_type* values_out, index_type& num_values_out,
                           hipStream_t stream = 0) const noexcept {
        retrieve_all_keys(keys_out, num_keys_out, stream);
        if (keys_out != nullptr)
            retrieve(keys_out, num_keys_out, begin_offsets_out, end_offsets_out, values_out, num_values_out,
                     stream);
        if (stream == 0) (void)hipStreamSynchronize(stream);
    }

    // :495-527; order unspecified
    WC_H void retrieve_all_keys(key_type* keys_out, index_type& num_out, hipStream_t stream = 0) const noexcept {
        if (!hash_table_.is_initialized_) return;
        if (keys_out == nullptr) {
            num_out = hash_table_.size(stream);
            return;
        }
        index_type* tmp = hash_table_.temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        kernels::retrieve_all<typename hash_table_type::view_type>
            <<<warpcore::detail::grid_for(key_capacity(), 1, 8), block_size(), 0, stream>>>(
                keys_out, static_cast<handle_type*>(nullptr), tmp, hash_table_.view());
        (void)hipMemcpyAsync(&num_out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
    }

    // :769-803
    template<class StatusHandler = defaults::status_handler_t>
    WC_H void num_values(const key_type* keys_in, index_type num_in, index_type& num_out,
                         index_type* num_per_key_out = nullptr, hipStream_t stream = 0,
                         index_type probing_length = defaults::probing_length(),
                         typename StatusHandler::base_type* status_out = nullptr) const noexcept {
        static_assert(checks::is_status_handler<StatusHandler>(), "not a valid status handler type");
        if (!hash_table_.is_initialized_) return;
        index_type* tmp = hash_table_.temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        if (num_in != 0)
            kernels::bl_num_values<view_type, StatusHandler>
                <<<warpcore::detail::grid_for(num_in, cg_size()), block_size(), 0, stream>>>(
                    keys_in, num_in, tmp, num_per_key_out, view(), probing_length, status_out);
        (void)hipMemcpyAsync(&num_out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
    }

    // :805-822
    WC_H index_type num_values(hipStream_t stream = 0) const noexcept {
        if (!hash_table_.is_initialized_) return 0;
        index_type* tmp = hash_table_.temp_.get();
        (void)hipMemsetAsync(tmp, 0, sizeof(index_type), stream);
        kernels::bl_total_values<view_type>
            <<<warpcore::detail::grid_for(key_capacity(), 1, 8), block_size(), 0, stream>>>(tmp, view());
        index_type out = 0;
        (void)hipMemcpyAsync(&out, tmp, sizeof(index_type), hipMemcpyDeviceToHost, stream);
        (void)hipStreamSynchronize(stream);
        return out;
    }

    WC_H float key_load_factor(hipStream_t stream = 0) const noexcept { return hash_table_.load_factor(stream); }
    WC_H float value_load_factor(hipStream_t stream = 0) const noexcept { return value_store_.load_factor(stream); }
    WC_H index_type bytes_total() const noexcept { return hash_table_.bytes_total() + value_store_.bytes_total(); }
    WC_H index_type bytes_keys(hipStream_t stream = 0) const noexcept { return num_keys(stream) * sizeof(key_type); }
    WC_H index_type bytes_values(hipStream_t stream = 0) const noexcept {
        return num_values(stream) * sizeof(value_type);
    }
    WC_H index_type bytes_payload(hipStream_t stream = 0) const noexcept {
        return bytes_keys(stream) + bytes_values(stream);
    }
    WC_H float storage_density(hipStream_t stream = 0) const noexcept {
        return float(bytes_payload(stream)) / float(bytes_total());
    }
    WC_H float relative_storage_density(hipStream_t stream = 0) const noexcept {
        const float bytes_hash_table = float(hash_table_.capacity() * (sizeof(key_type) + sizeof(handle_type)));
        const float bytes_value_store = float(value_store_.bytes_occupied(stream));
        return float(bytes_payload(stream)) / (bytes_value_store + bytes_hash_table);
    }
    WC_HD bool is_initialized() const noexcept { return hash_table_.is_initialized(); }
    WC_H status_type peek_status(hipStream_t stream = 0) const noexcept { return hash_table_.peek_status(stream); }
    WC_H status_type pop_status(hipStream_t stream = 0) noexcept { return hash_table_.pop_status(stream); }
    WC_HD index_type key_capacity() const noexcept { return hash_table_.capacity(); }
    WC_HD index_type value_capacity() const noexcept { return value_store_.capacity(); }
    WC_H index_type num_keys(hipStream_t stream = 0) const noexcept { return hash_table_.size(stream); }
    WC_H index_type value_cells_used(hipStream_t stream = 0) const noexcept { return value_store_.next_free(stream); }
    WC_HD index_type max_values_per_key() const noexcept { return max_values_per_key_; }
    WC_HD bool is_copy() const noexcept { return is_copy_; }

    WC_HD view_type view() const noexcept {
        return view_type{hash_table_.view(), value_store_.view(), max_values_per_key_};
    }

private:
    WC_H void join_status(status_type status, hipStream_t stream = 0) const noexcept {
        hash_table_.join_status(status, stream);
    }
    hash_table_type hash_table_;
    value_store_type value_store_;
    index_type max_values_per_key_;
    bool is_copy_;
    index_type* scan_parts_;   // scratch of the two-pass retrieve's prefix sum
    // scratch of the two-pass insert (per pair of a chunk: its key's slot and
    // its key status), grown on demand; copies of the table do not own it
    std::uint32_t* scratch_ = nullptr;
    index_type scratch_cap_ = 0;
    WC_H bool reserve_scratch(index_type n) noexcept {
        if (n <= scratch_cap_) return true;
        if (is_copy_) return false;
        if (scratch_ != nullptr) (void)hipFree(scratch_);
        scratch_ = nullptr;
        scratch_cap_ = 0;
        if (hipMalloc(&scratch_, 2 * sizeof(std::uint32_t) * n) != hipSuccess) {
            scratch_ = nullptr;
            return false;
        }
        scratch_cap_ = n;
        return true;
    }
};

}  // namespace warpcore
