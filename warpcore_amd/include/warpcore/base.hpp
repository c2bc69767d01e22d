// Core definitions of the MI355X-native warpcore hot path.
//
// Counterpart of include/warpcore/base.cuh (index_t, status_base_t,
// get_valid_capacity) and the pieces of hpc_helpers the reference relies on
// (SDIV, thread ids). Everything here is HIP for gfx950; there is no CUDA
// path and no compatibility header.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#define WC_HD __host__ __device__ __forceinline__
#define WC_D __device__ __forceinline__
#define WC_H __host__ inline

namespace warpcore {

using index_t = std::uint64_t;        // include/warpcore/base.cuh:21
using status_base_t = std::uint32_t;  // include/warpcore/base.cuh:22

// Ceiling division (hpc_helpers' SDIV).
WC_HD constexpr index_t SDIV(index_t a, index_t b) noexcept { return (a + b - 1) / b; }

// Threads per block of every launch in this library: four 64-lane waves.
constexpr index_t block_size() noexcept { return 256; }
// Lanes of one wavefront on CDNA.
constexpr index_t wave_size() noexcept { return 64; }

namespace detail {

// Number of compute units of the current device (cached per device).
WC_H int num_compute_units() noexcept {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

// Grid for a grid-stride launch that assigns `groups` cooperative groups of
// `cg` lanes: enough waves to fill every CU several times over, never more
// than the work needs.
WC_H unsigned grid_for(index_t groups, index_t cg, index_t waves_per_cu = 32) noexcept {
    const index_t threads = groups * cg;
    const index_t needed = SDIV(threads, block_size());
    const index_t cap = index_t(num_compute_units()) * waves_per_cu * wave_size() / block_size();
    index_t g = std::min(needed, cap);
    return unsigned(g == 0 ? 1 : g);
}

// Keys each cooperative group keeps in flight in the batched probing kernels.
#ifndef WARPCORE_KEYS_PER_GROUP
#define WARPCORE_KEYS_PER_GROUP 4
#endif
constexpr int keys_per_group() noexcept { return WARPCORE_KEYS_PER_GROUP; }

// Grid of a batched probing launch: one wavefront takes 64 / cg groups of
// keys_per_group() keys per round; at most 8 blocks of 256 per CU.
WC_H unsigned grid_for_batched(index_t items, index_t cg) noexcept {
    const index_t per_wave = (wave_size() / cg) * index_t(keys_per_group());
    const index_t waves = SDIV(items, per_wave);
    const index_t blocks = SDIV(waves, block_size() / wave_size());
    const index_t cap = index_t(num_compute_units()) * 8;
    const index_t g = std::min(blocks, cap);
    return unsigned(g == 0 ? 1 : g);
}

// Keys each lane keeps in flight in the lane-per-key probing kernels.
#ifndef WARPCORE_KEYS_PER_LANE
#define WARPCORE_KEYS_PER_LANE 2
#endif
constexpr int keys_per_lane() noexcept { return WARPCORE_KEYS_PER_LANE; }

// Largest table (bytes of slots) whose lookups are probed lane-per-key, and
// the same for inserts. One lane reading a whole 128-byte window issues 8
// loads, each translating 64 addresses, so on larger tables the address
// translation caches miss and cooperative windows (one translation per
// line) are faster. Measured on MI355X at load factor 0.9: lane lookups win
// up to about 75 MB of 16-byte slots (2^22 keys: 25.2 vs 24.7 Gops/s; 2^20:
// 29.4 vs 25.9) and lose from 300 MB on (2^24: 14.9 vs 19.8); lane inserts
// lose at every size (2^20: 8.6 vs 11.9 Gops/s). Both limits are process-
// wide settings (set_lane_table_bytes, set_lane_insert_table_bytes) so a
// tuner or a test can move them.
#ifndef WARPCORE_LANE_TABLE_BYTES
#define WARPCORE_LANE_TABLE_BYTES (std::uint64_t(1) << 27)
#endif
#ifndef WARPCORE_LANE_INSERT_TABLE_BYTES
#define WARPCORE_LANE_INSERT_TABLE_BYTES std::uint64_t(0)
#endif
inline std::atomic<std::uint64_t>& lane_table_bytes_setting() noexcept {
    static std::atomic<std::uint64_t> bytes{WARPCORE_LANE_TABLE_BYTES};
    return bytes;
}
inline std::atomic<std::uint64_t>& lane_insert_table_bytes_setting() noexcept {
    static std::atomic<std::uint64_t> bytes{WARPCORE_LANE_INSERT_TABLE_BYTES};
    return bytes;
}
WC_H std::uint64_t lane_table_bytes() noexcept { return lane_table_bytes_setting().load(std::memory_order_relaxed); }
WC_H void set_lane_table_bytes(std::uint64_t bytes) noexcept {
    lane_table_bytes_setting().store(bytes, std::memory_order_relaxed);
}
WC_H std::uint64_t lane_insert_table_bytes() noexcept {
    return lane_insert_table_bytes_setting().load(std::memory_order_relaxed);
}
WC_H void set_lane_insert_table_bytes(std::uint64_t bytes) noexcept {
    lane_insert_table_bytes_setting().store(bytes, std::memory_order_relaxed);
}

// Grid of a lane-per-key launch: keys_per_lane() keys per lane, at most 8
// blocks of 256 per CU.
WC_H unsigned grid_for_lanes(index_t items) noexcept {
    const index_t lanes = SDIV(items, index_t(keys_per_lane()));
    const index_t blocks = SDIV(lanes, block_size());
    const index_t cap = index_t(num_compute_units()) * 8;
    const index_t g = std::min(blocks, cap);
    return unsigned(g == 0 ? 1 : g);
}

// Available device memory (hpc_helpers' available_gpu_memory).
WC_H std::size_t available_gpu_memory() noexcept {
    std::size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    return free_b;
}

}  // namespace detail

}  // namespace warpcore
