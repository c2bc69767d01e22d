// Cooperative probing groups on 64-lane CDNA wavefronts.
//
// The reference probes with cg::thread_block_tile<CG> (CG <= 32, 32-bit
// ballot masks; include/warpcore/checks.cuh:13-28). Here a tile is an aligned
// run of N lanes (N = 1..64) of one wavefront: ballots are 64-bit
// (v_cmp + s_ballot), shuffles are ds_bpermute within the wave, and a tile of
// 64 is the whole wavefront. Lanes of a tile execute in lockstep, so no
// explicit tile barrier exists; callers keep tile-uniform control flow.
#pragma once

#include "base.hpp"

namespace warpcore {

WC_D unsigned lane_id() noexcept {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Lowest set bit index + 1 (0 if none), like __ffs on a 64-bit mask.
WC_HD int ffs64(std::uint64_t m) noexcept { return m ? __builtin_ctzll(m) + 1 : 0; }

template<unsigned N>
class tile {
    static_assert(N >= 1 && N <= 64 && (N & (N - 1)) == 0, "tile size must be a power of two <= 64");

public:
    using mask_type = std::uint64_t;

    WC_HD static constexpr unsigned size() noexcept { return N; }
    WC_HD static constexpr unsigned num_threads() noexcept { return N; }
    WC_HD static constexpr mask_type full_mask() noexcept { return N == 64 ? ~mask_type(0) : ((mask_type(1) << N) - 1); }

    WC_D static unsigned thread_rank() noexcept { return lane_id() & (N - 1); }
    WC_D static unsigned base_lane() noexcept { return lane_id() & ~(N - 1); }

    // Bit r of the result is set iff lane r of this tile is active and pred holds.
    WC_D static mask_type ballot(bool pred) noexcept {
        const mask_type m = __builtin_amdgcn_ballot_w64(pred);
        if constexpr (N == 64) return m;
        else return (m >> base_lane()) & full_mask();
    }
    WC_D static bool any(bool pred) noexcept { return ballot(pred) != 0; }
    WC_D static bool all(bool pred) noexcept { return ballot(!pred) == 0; }

    // Value of `v` held by tile lane `src`.
    template<class T>
    WC_D static T shfl(T v, unsigned src) noexcept {
        static_assert(std::is_trivially_copyable<T>::value, "shfl needs a trivially copyable type");
        if constexpr (N == 1) {
            return v;
        } else {
            const int addr = int((base_lane() + (src & (N - 1))) << 2);
            constexpr int words = (sizeof(T) + 3) / 4;
            int in[words] = {};
            int out[words];
            __builtin_memcpy(in, &v, sizeof(T));
#pragma unroll
            for (int i = 0; i < words; ++i) out[i] = __builtin_amdgcn_ds_bpermute(addr, in[i]);
            T r;
            __builtin_memcpy(&r, out, sizeof(T));
            return r;
        }
    }

    // Bitwise OR of `v` over the active lanes of this tile.
    WC_D static std::uint32_t reduce_or(std::uint32_t v) noexcept {
#pragma unroll
        for (unsigned d = 1; d < N; d <<= 1) v |= std::uint32_t(__shfl_xor(int(v), int(d), 64));
        return v;
    }
    // Sum of `v` over the lanes of this tile (all lanes must be active).
    WC_D static std::uint64_t reduce_add(std::uint64_t v) noexcept {
#pragma unroll
        for (unsigned d = 1; d < N; d <<= 1) {
            const std::uint32_t lo = std::uint32_t(__shfl_xor(int(std::uint32_t(v)), int(d), 64));
            const std::uint32_t hi = std::uint32_t(__shfl_xor(int(std::uint32_t(v >> 32)), int(d), 64));
            v += (std::uint64_t(hi) << 32) | lo;
        }
        return v;
    }

    // Lanes of this tile below the calling lane, as a tile-relative mask.
    WC_D static mask_type lanemask_lt() noexcept { return (mask_type(1) << thread_rank()) - 1; }
};

// Aggregated atomic increment: one atomicAdd per wavefront for all lanes
// with `pred` set; returns each such lane's unique slot (hpc_helpers'
// atomicAggInc re-derived for a 64-bit ballot).
WC_D index_t atomic_agg_inc(index_t* counter, bool pred = true) noexcept {
    const std::uint64_t m = __builtin_amdgcn_ballot_w64(pred);
    if (!pred) return 0;
    const unsigned lane = lane_id();
    const int leader = __builtin_ctzll(m);
    unsigned long long base = 0;
    if (int(lane) == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)__builtin_popcountll(m));
    const std::uint32_t lo = std::uint32_t(__builtin_amdgcn_readlane(int(std::uint32_t(base)), leader));
    const std::uint32_t hi = std::uint32_t(__builtin_amdgcn_readlane(int(std::uint32_t(base >> 32)), leader));
    const std::uint64_t b = (std::uint64_t(hi) << 32) | lo;
    const std::uint64_t below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
    return b + __builtin_popcountll(below);
}

}  // namespace warpcore
