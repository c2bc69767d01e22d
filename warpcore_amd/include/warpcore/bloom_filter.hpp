// Pattern-blocked Bloom filter (drop-in for include/warpcore/bloom_filter.cuh).
//
// The filter is num_slots words of `Slot` bits grouped into blocks of CG
// words. A key selects one block, block = H(key + seed) % num_blocks, and
// sets k bits inside it: for j < k, h_j = H(key + seed + j) % (CG * bits of
// Slot) names word h_j / bits and bit h_j % bits of the block
// (bloom_filter.cuh:172-186). Lane r of a CG-lane group owns word r of the
// block; with CG = 1 a lane owns a whole key. Insert ORs each word's
// pattern in with an atomicOr, skipped when the bits are already set
// (:188-191); a query is true iff every word holds its pattern (:236). The
// bit array is the OR of all patterns, so it does not depend on the order of
// inserts.
//
// fpr(n) evaluates the same blocked model as the reference (:341-408), a
// binomial mixture of per-block load, in log space: the reference's direct
// product of factorials overflows to +inf once n reaches a few hundred
// thousand keys (for 2^27 keys in 2^30 bits it returns inf), this stays
// finite and equals the direct product wherever that one is finite.
#pragma once

#include "base.hpp"
#include "capacity.hpp"
#include "defaults.hpp"
#include "engine.hpp"
#include "fastmod.hpp"
#include "hashers.hpp"
#include "status.hpp"
#include "tile.hpp"

#include <cmath>

namespace warpcore {

namespace detail {

template<class Key, class Hasher, class Slot, index_t CGSize>
struct BloomView {
    using key_type = Key;
    using slot_type = Slot;
    static constexpr index_t slot_bits = sizeof(Slot) * CHAR_BIT;
    static constexpr index_t block_bits = slot_bits * CGSize;

    Slot* words;
    index_t num_slots;
    FastMod mod_blocks;
    FastMod mod_slots;
    index_t k;
    Key seed;

    // word of `rank` in the key's block and the bits the key sets in it
    WC_D void locate(Key key, unsigned rank, index_t& slot_index, Slot& pattern) const noexcept {
        const Key seeded = Key(key + seed);
        const index_t block = mod_blocks.mod(std::uint64_t(Hasher::hash(seeded)));
        slot_index = block * CGSize + rank;
        if constexpr (CGSize > 1) slot_index = mod_slots.mod(slot_index);
        pattern = 0;
        for (index_t j = 0; j < k; ++j) {
            const index_t h = index_t(Hasher::hash(Key(seeded + Key(j)))) & (block_bits - 1);
            if (h / slot_bits == rank) pattern |= Slot(1) << (h % slot_bits);
        }
    }

    WC_D void insert(Key key, const tile<CGSize>& group) const noexcept {
        index_t i;
        Slot pattern;
        locate(key, group.thread_rank(), i, pattern);
        if ((pattern & words[i]) != pattern) atomic_or(words + i, pattern);
    }
    WC_D bool retrieve(Key key, const tile<CGSize>& group) const noexcept {
        index_t i;
        Slot pattern;
        locate(key, group.thread_rank(), i, pattern);
        return group.all((pattern & words[i]) == pattern);
    }
    // bloom_filter.cuh:263-297: insert and report whether all bits were set before
    WC_D bool insert_and_query(Key key, const tile<CGSize>& group) const noexcept {
        index_t i;
        Slot pattern;
        locate(key, group.thread_rank(), i, pattern);
        bool was_set = true;
        if ((pattern & words[i]) != pattern) {
            const Slot old = atomic_or(words + i, pattern);
            was_set = (old & pattern) == pattern;
        }
        return group.all(was_set);
    }

    WC_D static std::uint32_t atomic_or(std::uint32_t* p, std::uint32_t v) noexcept { return atomicOr(p, v); }
    WC_D static std::uint64_t atomic_or(std::uint64_t* p, std::uint64_t v) noexcept {
        return atomicOr(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
    }
};

}  // namespace detail

namespace kernels {

// Insert / query with U keys per group in flight (gpu_engine.cuh:127-174).
template<class View, int U>
__global__ __launch_bounds__(256) void bloom_insert(const typename View::key_type* __restrict__ keys_in,
                                                    index_t num_in, View view) {
    using K = typename View::key_type;
    using S = typename View::slot_type;
    constexpr unsigned CG = unsigned(View::block_bits / View::slot_bits);
    const unsigned rank = lane_id() & (CG - 1);
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    for (index_t base = gid; base < num_in; base += groups * U) {
        index_t at[U];
        S pat[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const index_t i = base + index_t(u) * groups;
            const K key = keys_in[i < num_in ? i : num_in - 1];
            view.locate(key, rank, at[u], pat[u]);
        }
        // OR without reading the word first: a non-returning atomic that
        // nothing waits for (the per-key device insert reads first and skips
        // words that already hold the pattern; the bit array is the same)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const index_t i = base + index_t(u) * groups;
            if (i < num_in && pat[u] != 0)   // a lane's word of the block may get no bit
                (void)__hip_atomic_fetch_or(view.words + at[u], pat[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template<class View, int U>
__global__ __launch_bounds__(256) void bloom_retrieve(const typename View::key_type* __restrict__ keys_in,
                                                      index_t num_in, bool* __restrict__ flags_out, View view) {
    using K = typename View::key_type;
    using S = typename View::slot_type;
    constexpr unsigned CG = unsigned(View::block_bits / View::slot_bits);
    const tile<CG> group;
    const unsigned rank = group.thread_rank();
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    for (index_t base = gid; base < num_in; base += groups * U) {
        index_t at[U];
        S pat[U], cur[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const index_t i = base + index_t(u) * groups;
            const K key = keys_in[i < num_in ? i : num_in - 1];
            view.locate(key, rank, at[u], pat[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) cur[u] = view.words[at[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const index_t i = base + index_t(u) * groups;
            const bool all = group.all((pat[u] & cur[u]) == pat[u]);
            if (i < num_in && rank == 0) flags_out[i] = all;
        }
    }
}

// Bulk form of the device member insert_and_query (bloom_filter.cuh:263-297):
// one lane per key (CG = 1); flag = all of the key's bits were set before
// its own atomicOr.
template<class View>
__global__ __launch_bounds__(256) void bloom_insert_and_query(const typename View::key_type* __restrict__ keys_in,
                                                              index_t num_in, bool* __restrict__ flags_out,
                                                              View view) {
    static_assert(View::block_bits == View::slot_bits, "insert_and_query needs one lane per key");
    const tile<1> group;
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num_in; i += stride)
        flags_out[i] = view.insert_and_query(keys_in[i], group);
}

}  // namespace kernels

template<class Key, class Hasher = defaults::hasher_t<Key>, class Slot = std::uint64_t, index_t CGSize = 1>
class BloomFilter {
    static_assert(checks::is_valid_key_type<Key>(), "invalid key type");
    static_assert(checks::is_hasher<Hasher>(), "not a valid hasher type");
    static_assert(checks::is_valid_slot_type<Slot>(), "invalid slot type");
    static_assert(checks::is_valid_cg_size(CGSize), "invalid cooperative group size");

public:
    using key_type = Key;
    using value_type = bool;
    using index_type = index_t;
    using slot_type = Slot;
    using status_type = Status;
    using view_type = detail::BloomView<Key, Hasher, Slot, CGSize>;

    WC_HD static constexpr index_type cg_size() noexcept { return CGSize; }
    WC_HD static constexpr index_type slot_bits() noexcept { return sizeof(slot_type) * CHAR_BIT; }
    WC_HD static constexpr index_type block_bits() noexcept { return slot_bits() * cg_size(); }

    // bloom_filter.cuh:77-93
    WC_H explicit BloomFilter(index_type num_bits, index_type k, key_type seed = defaults::seed<key_type>()) noexcept
        : bloom_filter_(nullptr), num_bits_(num_bits), num_slots_(SDIV(num_bits, slot_bits())),
          num_blocks_(SDIV(num_slots_, cg_size())), mod_blocks_(num_blocks_ ? num_blocks_ : 1),
          mod_slots_(num_slots_ ? num_slots_ : 1), k_(k), seed_(seed), is_copy_(false) {
        if (num_slots_ != 0 && hipMalloc(&bloom_filter_, sizeof(slot_type) * num_slots_) != hipSuccess)
            bloom_filter_ = nullptr;
        init();
    }
    WC_HD BloomFilter(const BloomFilter& o) noexcept
        : bloom_filter_(o.bloom_filter_), num_bits_(o.num_bits_), num_slots_(o.num_slots_),
          num_blocks_(o.num_blocks_), mod_blocks_(o.mod_blocks_), mod_slots_(o.mod_slots_), k_(o.k_),
          seed_(o.seed_), is_copy_(true) {}
    WC_H BloomFilter(BloomFilter&& o) noexcept
        : bloom_filter_(o.bloom_filter_), num_bits_(o.num_bits_), num_slots_(o.num_slots_),
          num_blocks_(o.num_blocks_), mod_blocks_(o.mod_blocks_), mod_slots_(o.mod_slots_), k_(o.k_),
          seed_(o.seed_), is_copy_(o.is_copy_) {
        o.is_copy_ = true;
    }
    WC_HD ~BloomFilter() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_ && bloom_filter_ != nullptr) (void)hipFree(bloom_filter_);
#endif
    }

    // bloom_filter.cuh:142-161
    WC_H void init(key_type seed, hipStream_t stream = 0) noexcept {
        seed_ = seed;
        if (bloom_filter_ == nullptr) return;
        kernels::memset<slot_type><<<detail::grid_for(num_slots_, 1, 8), block_size(), 0, stream>>>(
            bloom_filter_, num_slots_, slot_type(0));
    }
    WC_H void init(hipStream_t stream = 0) noexcept { init(seed_, stream); }

    WC_D void insert(key_type key, const tile<cg_size()>& group) const noexcept { view().insert(key, group); }
    WC_D bool retrieve(key_type key, const tile<cg_size()>& group) const noexcept { return view().retrieve(key, group); }
    template<index_type C = cg_size(), class = std::enable_if_t<C == 1>>
    WC_D bool insert_and_query(key_type key, const tile<cg_size()>& group) const noexcept {
        return view().insert_and_query(key, group);
    }

    // bloom_filter.cuh:199-208
    WC_H void insert(const key_type* keys_in, index_type num_in, hipStream_t stream = 0) noexcept {
        if (bloom_filter_ == nullptr || num_in == 0) return;
        kernels::bloom_insert<view_type, 4>
            <<<detail::grid_for(SDIV(num_in, 4), cg_size()), block_size(), 0, stream>>>(keys_in, num_in, view());
    }
    // bloom_filter.cuh:245-255
    WC_H void retrieve(const key_type* keys_in, index_type num_in, bool* flags_out,
                       hipStream_t stream = 0) const noexcept {
        if (bloom_filter_ == nullptr || num_in == 0) return;
        kernels::bloom_retrieve<view_type, 4>
            <<<detail::grid_for(SDIV(num_in, 4), cg_size()), block_size(), 0, stream>>>(keys_in, num_in, flags_out,
                                                                                      view());
    }

    WC_HD index_type num_bits() const noexcept { return num_bits_; }
    WC_HD index_type num_slots() const noexcept { return num_slots_; }
    WC_HD index_type num_blocks() const noexcept { return num_blocks_; }
    WC_HD index_type k() const noexcept { return k_; }
    WC_HD key_type seed() const noexcept { return seed_; }
    WC_HD slot_type* data() const noexcept { return bloom_filter_; }
    WC_H index_type bytes_total() const noexcept { return num_slots_ * sizeof(slot_type); }
    WC_HD bool is_copy() const noexcept { return is_copy_; }

    // False-positive rate after n inserts under the blocked model of
    // bloom_filter.cuh:341-408, summed in log space.
    WC_H double fpr(index_type n) const noexcept {
        const index_type blocks = num_bits_ / block_bits();
        if (blocks == 0) return 1.0;
        const double b = double(blocks);
        const index_type m = index_type(double(num_bits_) / b);
        const index_type imax = 5 * n / blocks;
        const double p = 1.0 / b;
        const double lp = std::log(p), lq = std::log1p(-p), lfn = std::lgamma(double(n) + 1.0);
        double res = 0.0;
        for (index_type i = 0; i < imax && i <= n; ++i) {
            const double lbin = lfn - std::lgamma(double(i) + 1.0) - std::lgamma(double(n - i) + 1.0) +
                                double(i) * lp + double(n - i) * lq;
            const double base = std::pow(1.0 - std::pow(1.0 - 1.0 / double(m), double(i * k_)), double(k_));
            res += std::exp(lbin) * base;
        }
        return res;
    }

    WC_HD view_type view() const noexcept {
        return view_type{bloom_filter_, num_slots_, mod_blocks_, mod_slots_, k_, seed_};
    }

private:
    slot_type* bloom_filter_;
    index_type num_bits_;
    index_type num_slots_;
    index_type num_blocks_;
    detail::FastMod mod_blocks_;
    detail::FastMod mod_slots_;
    index_type k_;
    key_type seed_;
    bool is_copy_;
};

}  // namespace warpcore
