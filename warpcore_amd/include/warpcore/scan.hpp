// Device-wide prefix sum of integer counts (summed as index_t), used by the two-pass
// multi-value retrieve (the role CUB's DeviceScan::InclusiveSum plays in
// multi_value_hash_table.cuh:431-475 and bucket_list_hash_table.cuh:414-457).
//
// Reduce-then-scan over a fixed number of contiguous chunks: pass 1 sums
// each chunk, pass 2 scans the (at most kScanParts) chunk sums in one
// workgroup, pass 3 rescans each chunk from its offset. Counts are read
// twice and written once (24 bytes per element), the scratch is
// kScanParts words whatever n is, and in == out is allowed.
#pragma once

#include "base.hpp"
#include "tile.hpp"

namespace warpcore {
namespace detail {

constexpr index_t kScanParts = 1024;
constexpr unsigned kScanThreads = 256;
constexpr unsigned kScanItems = 8;   // per thread and tile: 2048 counts per tile

WC_H index_t scan_parts_for(index_t n) noexcept {
    const index_t tiles = SDIV(n, index_t(kScanThreads) * kScanItems);
    return tiles < kScanParts ? (tiles == 0 ? 1 : tiles) : kScanParts;
}

// inclusive scan over the 64 lanes of a wavefront
WC_D index_t wave_inclusive_scan(index_t v) noexcept {
    const unsigned lane = lane_id();
#pragma unroll
    for (unsigned d = 1; d < 64; d <<= 1) {
        const index_t u = (index_t)__shfl_up((unsigned long long)v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Inclusive scan of one value per thread over the whole workgroup of
// kScanThreads; `total` gets the sum of all values.
WC_D index_t block_inclusive_scan(index_t v, index_t& total) noexcept {
    __shared__ index_t wave_sums[kScanThreads / 64];
    const index_t incl = wave_inclusive_scan(v);
    const unsigned wave = threadIdx.x / 64;
    if (lane_id() == 63) wave_sums[wave] = incl;
    __syncthreads();
    index_t before = 0;
    total = 0;
#pragma unroll
    for (unsigned w = 0; w < kScanThreads / 64; ++w) {
        const index_t s = wave_sums[w];
        if (w < wave) before += s;
        total += s;
    }
    __syncthreads();
    return incl + before;
}

}  // namespace detail

namespace kernels {

// pass 1: parts[b] = sum of chunk b
template<class T>
__global__ __launch_bounds__(256) void scan_reduce(const T* __restrict__ in, index_t n, index_t chunk,
                                                   index_t* __restrict__ parts) {
    const index_t lo = index_t(blockIdx.x) * chunk;
    const index_t hi = lo + chunk < n ? lo + chunk : n;
    index_t s = 0;
    for (index_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s += in[i];
    index_t total;
    (void)detail::block_inclusive_scan(s, total);
    if (threadIdx.x == 0) parts[blockIdx.x] = total;
}

// pass 2: exclusive scan of the part sums in place (one workgroup)
template<class T>
__global__ __launch_bounds__(256) void scan_parts(index_t* __restrict__ parts, index_t num_parts) {
    index_t carry = 0;
    for (index_t base = 0; base < num_parts; base += blockDim.x) {
        const index_t i = base + threadIdx.x;
        const index_t v = i < num_parts ? parts[i] : 0;
        index_t total;
        const index_t incl = detail::block_inclusive_scan(v, total);
        if (i < num_parts) parts[i] = carry + incl - v;
        carry += total;
    }
}

// pass 3: scan chunk b from offset parts[b]; tiles of 2048 counts staged in
// LDS so that loads and stores stay coalesced
template<class T>
__global__ __launch_bounds__(256) void scan_apply(const T* in, index_t n, index_t chunk,
                                                  const index_t* __restrict__ parts, T* out, bool inclusive) {
    constexpr unsigned NT = detail::kScanThreads, I = detail::kScanItems;
    __shared__ index_t tile_buf[NT * I];
    const index_t lo = index_t(blockIdx.x) * chunk;
    const index_t hi = lo + chunk < n ? lo + chunk : n;
    index_t carry = parts[blockIdx.x];
    for (index_t base = lo; base < hi; base += NT * I) {
#pragma unroll
        for (unsigned j = 0; j < I; ++j) {
            const index_t i = base + j * NT + threadIdx.x;
            tile_buf[j * NT + threadIdx.x] = i < hi ? index_t(in[i]) : 0;
        }
        __syncthreads();
        index_t mine[I];
        index_t sum = 0;
#pragma unroll
        for (unsigned j = 0; j < I; ++j) {
            mine[j] = tile_buf[threadIdx.x * I + j];
            sum += mine[j];
        }
        index_t total;
        const index_t incl = detail::block_inclusive_scan(sum, total);
        index_t run = carry + incl - sum;
#pragma unroll
        for (unsigned j = 0; j < I; ++j) {
            const index_t v = mine[j];
            tile_buf[threadIdx.x * I + j] = inclusive ? run + v : run;
            run += v;
        }
        __syncthreads();
#pragma unroll
        for (unsigned j = 0; j < I; ++j) {
            const index_t i = base + j * NT + threadIdx.x;
            if (i < hi) out[i] = T(tile_buf[j * NT + threadIdx.x]);
        }
        __syncthreads();
        carry += total;
    }
}

}  // namespace kernels

namespace detail {

// Prefix sum of n counts (in == out allowed). `parts` must hold
// kScanParts words. Three launches on `stream`, no host synchronisation.
template<class T>
WC_H void device_scan(const T* in, T* out, index_t n, index_t* parts, bool inclusive, hipStream_t stream) noexcept {
    if (n == 0) return;
    const index_t num_parts = scan_parts_for(n);
    const index_t chunk = SDIV(n, num_parts);
    const unsigned grid = unsigned(SDIV(n, chunk));
    kernels::scan_reduce<T><<<grid, kScanThreads, 0, stream>>>(in, n, chunk, parts);
    kernels::scan_parts<T><<<1, kScanThreads, 0, stream>>>(parts, index_t(grid));
    kernels::scan_apply<T><<<grid, kScanThreads, 0, stream>>>(in, n, chunk, parts, out, inclusive);
}

}  // namespace detail
}  // namespace warpcore
