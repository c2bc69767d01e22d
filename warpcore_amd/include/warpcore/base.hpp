// Core definitions of the MI355X-native warpcore hot path.
//
// Counterpart of include/warpcore/base.cuh (index_t, status_base_t,
// get_valid_capacity) and the pieces of hpc_helpers the reference relies on
// (SDIV, thread ids). Everything here is HIP for gfx950; there is no CUDA
// path and no compatibility header.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
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

// Available device memory (hpc_helpers' available_gpu_memory).
WC_H std::size_t available_gpu_memory() noexcept {
    std::size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    return free_b;
}

}  // namespace detail

}  // namespace warpcore
