// Measurement probes of the memory system (declared in warpcore_probes.h):
// a streaming copy of known traffic (to calibrate traffic counters), random
// whole-line reads and random line claims (the practical ceilings of the
// probing kernels). Built into libwarpcore_probes.so for bench.py; not part
// of the table library.
#include <warpcore/base.hpp>
#include <warpcore/fastmod.hpp>
#include <warpcore/hashers.hpp>

#include <hip/hip_runtime.h>

#include <cstdint>

using namespace warpcore;

namespace {

constexpr int kBadArgument = -2;

hipStream_t as_stream(void* s) noexcept { return reinterpret_cast<hipStream_t>(s); }
int last_error() noexcept { return int(hipGetLastError()); }
unsigned grid1(index_t n) { return detail::grid_for(n, 1, 8); }

// Contiguous copy, 16 bytes per lane and access: a plain streaming kernel
// of known traffic (reads and writes `n16` * 16 bytes once).
__global__ __launch_bounds__(256) void copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, index_t n16) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}
// Random whole-line gather: group g of 8 lanes reads `per_group` lines in
// rounds of U, each line chosen by a hash of (seed, index), 16 bytes per
// lane; the XOR of what was read goes to *sink only if it matches a
// constant, so the loads cannot be removed. Measures the rate at which the
// memory system serves random 128-byte lines for a given footprint.
template<int U>
__global__ __launch_bounds__(256) void gather_lines(const uint4* __restrict__ table, detail::FastMod mod,
                                                    index_t total, std::uint64_t seed, std::uint32_t* sink) {
    const unsigned rank = threadIdx.x & 7u;
    const index_t group = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / 8;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / 8;
    std::uint32_t acc = 0;
    for (index_t base = group * U; base < total; base += groups * U) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const index_t line = mod.mod(hashers::MurmurHash<std::uint64_t>::hash(seed ^ (base + u)));
            v[u] = table[line * 8 + rank];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc ^= v[u].x ^ v[u].w;
    }
    if (acc == 0x9e3779b9u) *sink = acc;
}
// Random line claim: as gather_lines, and then one lane per line swaps the
// 8 bytes it read for that value plus one with a returning compare-and-swap,
// the lane chosen by the data read, so the CAS waits for the line. The
// memory pattern of an insert (read the window, claim a slot of it) without
// the probing: the rate of line reads each followed by a dependent atomic.
template<int U>
__global__ __launch_bounds__(256) void claim_lines(uint4* __restrict__ table, detail::FastMod mod, index_t total,
                                                   std::uint64_t seed, std::uint32_t* sink) {
    const unsigned rank = threadIdx.x & 7u;
    const index_t group = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / 8;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / 8;
    std::uint32_t acc = 0;
    for (index_t base = group * U; base < total; base += groups * U) {
        index_t line[U];
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            line[u] = mod.mod(hashers::MurmurHash<std::uint64_t>::hash(seed ^ (base + u)));
            v[u] = table[line[u] * 8 + rank];
        }
        unsigned long long old[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            old[u] = 0;
            if (rank == (v[u].x & 7u)) {
                const unsigned long long seen = (unsigned long long)v[u].y << 32 | v[u].x;
                old[u] = atomicCAS(reinterpret_cast<unsigned long long*>(table + line[u] * 8 + rank), seen, seen + 1);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc ^= std::uint32_t(old[u]) ^ v[u].w;
    }
    if (acc == 0x9e3779b9u) *sink = acc;
}
}  // namespace

extern "C" int wc_util_claim_lines(void* table, uint64_t table_bytes, uint64_t lines, uint64_t seed, uint32_t* sink,
                                   void* stream) {
    const uint64_t num_lines = table_bytes / 128;
    if (!table || !sink || num_lines == 0 || reinterpret_cast<std::uintptr_t>(table) % 16 != 0)
        return kBadArgument;
    const unsigned grid = unsigned(detail::num_compute_units() * 8);
    claim_lines<8><<<grid, block_size(), 0, as_stream(stream)>>>(static_cast<uint4*>(table), detail::FastMod(num_lines),
                                                                 lines, seed, sink);
    return last_error();
}

extern "C" int wc_util_gather_lines(const void* table, uint64_t table_bytes, uint64_t lines, uint64_t seed,
                                    uint32_t* sink, void* stream) {
    const uint64_t num_lines = table_bytes / 128;
    if (!table || !sink || num_lines == 0 || reinterpret_cast<std::uintptr_t>(table) % 16 != 0)
        return kBadArgument;
    const unsigned grid = unsigned(detail::num_compute_units() * 8);
    gather_lines<8><<<grid, block_size(), 0, as_stream(stream)>>>(static_cast<const uint4*>(table),
                                                                 detail::FastMod(num_lines), lines, seed, sink);
    return last_error();
}

extern "C" int wc_util_copy(void* dst, const void* src, uint64_t bytes, void* stream) {
    if (bytes == 0) return 0;
    if (!dst || !src || bytes % 16 != 0 || reinterpret_cast<std::uintptr_t>(dst) % 16 != 0 ||
        reinterpret_cast<std::uintptr_t>(src) % 16 != 0)
        return kBadArgument;
    copy16<<<grid1(bytes / 16), block_size(), 0, as_stream(stream)>>>(static_cast<uint4*>(dst),
                                                                      static_cast<const uint4*>(src), bytes / 16);
    return last_error();
}

