// Device building blocks shared by the tables: slot loads, atomics, status
// joining, and the generic batch kernels (the counterpart of
// include/warpcore/gpu_engine.cuh). Batch kernels are grid-stride: a grid
// sized to fill every CU walks the input, CG lanes per item, so no launch is
// limited by the 2^32-thread grid bound the reference's one-group-per-item
// grids hit at 2^28 keys x CG 16.
#pragma once

#include "base.hpp"
#include "status.hpp"
#include "tile.hpp"

namespace warpcore {
namespace detail {

// ---------------------------------------------------------------- atomics --
WC_D std::uint32_t atomic_cas(std::uint32_t* p, std::uint32_t e, std::uint32_t d) noexcept { return atomicCAS(p, e, d); }
WC_D std::uint64_t atomic_cas(std::uint64_t* p, std::uint64_t e, std::uint64_t d) noexcept {
    return atomicCAS(reinterpret_cast<unsigned long long*>(p), (unsigned long long)e, (unsigned long long)d);
}
WC_D int atomic_add(int* p, int v) noexcept { return atomicAdd(p, v); }
WC_D std::uint32_t atomic_add(std::uint32_t* p, std::uint32_t v) noexcept { return atomicAdd(p, v); }
WC_D std::uint64_t atomic_add(std::uint64_t* p, std::uint64_t v) noexcept {
    return atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}
WC_D int atomic_exch(int* p, int v) noexcept { return atomicExch(p, v); }
WC_D std::uint32_t atomic_exch(std::uint32_t* p, std::uint32_t v) noexcept { return atomicExch(p, v); }
WC_D std::uint64_t atomic_exch(std::uint64_t* p, std::uint64_t v) noexcept {
    return atomicExch(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}
template<class T> WC_D T atomic_load_relaxed(const T* p) noexcept {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// --------------------------------------------------------------- slot i/o --
// One vector load per slot: dwordx4 for 16-byte slots, dwordx2 for 8-byte.
template<class Slot>
WC_D Slot load_slot(const Slot* p) noexcept {
    Slot s;
    if constexpr (sizeof(Slot) == 16 && alignof(Slot) == 16) {
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        __builtin_memcpy(&s, &v, 16);
    } else if constexpr (sizeof(Slot) == 8 && alignof(Slot) == 8) {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        __builtin_memcpy(&s, &v, 8);
    } else {
        s = *p;
    }
    return s;
}

// u32 key + u32 value in one 8-byte slot: key and value publish together
// with one 64-bit compare-and-swap.
template<class Slot>
constexpr bool is_packed_pair() noexcept {
    return sizeof(Slot) == 8 && sizeof(Slot::key) == 4 && sizeof(Slot::value) == 4 &&
           std::is_trivially_copyable<decltype(Slot::value)>::value;
}

template<class Slot>
WC_D std::uint64_t slot_bits(const Slot& s) noexcept { std::uint64_t b; __builtin_memcpy(&b, &s, 8); return b; }
template<class Slot>
WC_D Slot slot_from_bits(std::uint64_t b) noexcept { Slot s; __builtin_memcpy(&s, &b, 8); return s; }

// ---------------------------------------------------------- status joining --
// Device API calls join table flags at once (status.cuh:187-193).
struct ImmediateJoin {
    Status* word;
    WC_D void operator()(Status s) const noexcept { if (word != nullptr) word->atomic_join(s); }
};

// Batch kernels OR flags into a register and join them once per wavefront
// when the kernel ends; the table word is only observed after the kernel,
// so the result is the same as joining per key, minus up to 2^28 atomics.
struct DeferredJoin {
    status_base_t acc = 0;
    WC_D void operator()(Status s) noexcept { acc |= s.base(); }
    // Must be called by all 64 lanes of the wavefront.
    WC_D void flush(Status* word) noexcept {
        const status_base_t all = tile<64>::reduce_or(acc);
        if (lane_id() == 0 && all != 0 && word != nullptr) word->atomic_join(Status::from_base(all));
    }
};

// Grid-stride iteration over items, one CG per item.
template<unsigned CG>
struct GroupRange {
    index_t first, stride;
    WC_D GroupRange() noexcept
        : first((index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG),
          stride((index_t(gridDim.x) * blockDim.x) / CG) {}
};

}  // namespace detail

namespace kernels {

template<class T>
__global__ __launch_bounds__(256) void memset(T* __restrict__ arr, index_t num, T value) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < num; i += stride) arr[i] = value;
}

// insert(keys, values) -> status per key (gpu_engine.cuh:203-228)
template<class Core, class StatusHandler>
__global__ __launch_bounds__(256) void insert(const typename Core::key_type* __restrict__ keys_in,
                                              const typename Core::value_type* __restrict__ values_in,
                                              index_t num_in, Core core, index_t probing_length,
                                              typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    const detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        const Status st = core.insert(keys_in[gid], values_in[gid], group, probing_length, join);
        if (group.thread_rank() == 0) StatusHandler::handle(st, status_out, gid);
    }
    join.flush(core.status_word());
}

// insert(keys) for key-only tables (gpu_engine.cuh:177-201)
template<class Core, class StatusHandler>
__global__ __launch_bounds__(256) void insert(const typename Core::key_type* __restrict__ keys_in, index_t num_in,
                                              Core core, index_t probing_length,
                                              typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    const detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        const Status st = core.insert(keys_in[gid], group, probing_length, join);
        if (group.thread_rank() == 0) StatusHandler::handle(st, status_out, gid);
    }
    join.flush(core.status_word());
}

// retrieve(keys) -> values; a value is written only when the status is none
// (gpu_engine.cuh:230-262)
template<class Core, class StatusHandler>
__global__ __launch_bounds__(256) void retrieve(const typename Core::key_type* __restrict__ keys_in, index_t num_in,
                                                typename Core::value_type* __restrict__ values_out, Core core,
                                                index_t probing_length,
                                                typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    const detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        typename Core::value_type v;
        const Status st = core.retrieve(keys_in[gid], v, group, probing_length, join);
        if (group.thread_rank() == 0) {
            if (!st.has_any()) values_out[gid] = v;
            StatusHandler::handle(st, status_out, gid);
        }
    }
    join.flush(core.status_word());
}

// erase(keys) (gpu_engine.cuh:352-376)
template<class Core, class StatusHandler>
__global__ __launch_bounds__(256) void erase(const typename Core::key_type* __restrict__ keys_in, index_t num_in,
                                             Core core, index_t probing_length,
                                             typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    const tile<CG> group;
    detail::DeferredJoin join;
    const detail::GroupRange<CG> r;
    for (index_t gid = r.first; gid < num_in; gid += r.stride) {
        const Status st = core.erase(keys_in[gid], group, probing_length, join);
        if (group.thread_rank() == 0) StatusHandler::handle(st, status_out, gid);
    }
    join.flush(core.status_word());
}

// Number of slots holding a valid key: per-lane counts, one LDS sum per
// block and one atomicAdd per block (gpu_engine.cuh:378-417).
template<class Core>
__global__ __launch_bounds__(256) void size(index_t* __restrict__ num_out, Core core, bool claimed = false) {
    __shared__ index_t partial[256 / 64];
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    index_t count = 0;
    // valid keys, or (claimed) every slot that is not EMPTY: keys and tombstones
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < core.capacity(); i += stride) {
        const auto k = core.key_at(i);
        count += claimed ? !Core::is_empty_key(k) : core.is_valid_key(k);
    }
    count = tile<64>::reduce_add(count);
    if (lane_id() == 0) partial[threadIdx.x / 64] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
        index_t s = 0;
        for (unsigned w = 0; w < blockDim.x / 64; ++w) s += partial[w];
        if (s) detail::atomic_add(num_out, s);
    }
}

// f(key, value) for every slot with a valid key (gpu_engine.cuh:28-46)
template<class Func, class Core>
__global__ __launch_bounds__(256) void for_each(Func f, Core core) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < core.capacity(); i += stride) {
        const auto s = core.slot_at(i);
        if (core.is_valid_key(s.key)) f(s.key, s.value);
    }
}

// Compact all valid (key, value) pairs; one atomicAdd per wavefront
// (single_value_hash_table.cuh:390-417, order unspecified).
template<class Core>
__global__ __launch_bounds__(256) void retrieve_all(typename Core::key_type* __restrict__ keys_out,
                                                    typename Core::value_type* __restrict__ values_out,
                                                    index_t* __restrict__ num_out, Core core) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    const index_t cap = core.capacity();
    // every lane of a wave iterates the same number of times
    const index_t first = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const index_t wave_first = first - lane_id();
    for (index_t base = wave_first; base < cap; base += stride) {
        const index_t i = base + lane_id();
        bool valid = false;
        typename Core::key_type k{};
        typename Core::value_type v{};
        if (i < cap) {
            const auto s = core.slot_at(i);
            k = s.key;
            v = s.value;
            valid = core.is_valid_key(k);
        }
        const index_t out = atomic_agg_inc(num_out, valid);
        if (valid) {
            if (keys_out != nullptr) keys_out[out] = k;
            if (values_out != nullptr) values_out[out] = v;
        }
    }
}

}  // namespace kernels
}  // namespace warpcore
