// extern "C" generators of synthetic inputs (declared in include/warpcore_c.h).
#include "capi_common.hpp"

#include <warpcore/fastmod.hpp>
#include <warpcore/hashers.hpp>
#include <warpcore/io.hpp>
#include <warpcore/random_distributions.hpp>

#include <algorithm>

using namespace warpcore;
using wc_capi::as_stream;
using wc_capi::last_error;

namespace {
unsigned grid1(index_t n) { return detail::grid_for(n, 1, 8); }

}  // namespace

extern "C" int wc_gen_unique_u32(uint32_t* out, uint64_t n, uint64_t offset, void* stream) {
    if (n == 0) return 0;
    if (n > (uint64_t(1) << 32) - 4) return wc_capi::kBadArgument;
    kernels::generate_unique<uint32_t><<<grid1(n), block_size(), 0, as_stream(stream)>>>(out, n, offset);
    return last_error();
}

extern "C" int wc_gen_unique_u64(uint64_t* out, uint64_t n, uint64_t offset, void* stream) {
    if (n == 0) return 0;
    kernels::generate_unique<uint64_t><<<grid1(n), block_size(), 0, as_stream(stream)>>>(out, n, offset);
    return last_error();
}

extern "C" int wc_gen_uniform_u32(uint32_t* out, uint64_t n, uint64_t seed, void* stream) {
    if (n == 0) return 0;
    kernels::generate_uniform<uint32_t><<<grid1(n), block_size(), 0, as_stream(stream)>>>(out, n, seed);
    return last_error();
}

extern "C" int wc_gen_uniform_u64(uint64_t* out, uint64_t n, uint64_t seed, void* stream) {
    if (n == 0) return 0;
    kernels::generate_uniform<uint64_t><<<grid1(n), block_size(), 0, as_stream(stream)>>>(out, n, seed);
    return last_error();
}

extern "C" int wc_gen_zipf_u64(uint64_t* out, uint64_t n, uint64_t num_distinct, double s, uint64_t seed,
                               uint64_t key_offset, void* stream) {
    if (n == 0) return 0;
    if (num_distinct == 0 || !(s > 0.0)) return wc_capi::kBadArgument;
    kernels::generate_zipf<uint64_t><<<grid1(n), block_size(), 0, as_stream(stream)>>>(
        out, n, random::ZipfSampler(num_distinct, s), seed, key_offset);
    return last_error();
}

extern "C" int wc_gen_zipf_u32(uint32_t* out, uint64_t n, uint64_t num_distinct, double s, uint64_t seed,
                               uint64_t key_offset, void* stream) {
    if (n == 0) return 0;
    if (num_distinct == 0 || !(s > 0.0)) return wc_capi::kBadArgument;
    kernels::generate_zipf<uint32_t><<<grid1(n), block_size(), 0, as_stream(stream)>>>(
        out, n, random::ZipfSampler(num_distinct, s), seed, key_offset);
    return last_error();
}

extern "C" int wc_gen_interleave_u64(uint64_t* out, const uint64_t* a, const uint64_t* b, uint64_t half,
                                     uint64_t seed, void* stream) {
    if (half == 0) return 0;
    kernels::interleave_halves<uint64_t><<<grid1(half), block_size(), 0, as_stream(stream)>>>(out, a, b, half, seed);
    return last_error();
}

extern "C" int wc_gen_interleave_u32(uint32_t* out, const uint32_t* a, const uint32_t* b, uint64_t half,
                                     uint64_t seed, void* stream) {
    if (half == 0) return 0;
    kernels::interleave_halves<uint32_t><<<grid1(half), block_size(), 0, as_stream(stream)>>>(out, a, b, half, seed);
    return last_error();
}

// Host-side helpers shared with the tests: the capacity rule and the unique
// key map evaluated on the CPU (no device needed).
extern "C" uint64_t wc_get_valid_capacity(uint64_t min_capacity, uint64_t cg_size) {
    return detail::get_valid_capacity(min_capacity, cg_size);
}
extern "C" uint64_t wc_unique_key_u64(uint64_t offset, uint64_t i) { return random::unique_key<uint64_t>(offset, i); }
extern "C" uint32_t wc_unique_key_u32(uint64_t offset, uint64_t i) { return random::unique_key<uint32_t>(offset, i); }

// Kernel-shape policy: tables of at most this many bytes of slots are probed
// one lane per key, larger ones by cooperative groups.
extern "C" void wc_set_lane_table_bytes(uint64_t bytes) { detail::set_lane_table_bytes(bytes); }
extern "C" uint64_t wc_get_lane_table_bytes(void) { return detail::lane_table_bytes(); }
extern "C" void wc_set_lane_insert_table_bytes(uint64_t bytes) { detail::set_lane_insert_table_bytes(bytes); }
extern "C" uint64_t wc_get_lane_insert_table_bytes(void) { return detail::lane_insert_table_bytes(); }

// Kernel used for bulk lookups probed by cooperative groups: 0 = streaming
// state machine, 1 = lockstep batches.
extern "C" void wc_set_group_kernel(int kernel) { detail::set_group_kernel(kernel); }
extern "C" int wc_get_group_kernel(void) { return detail::group_kernel(); }

// CountingHashTable: combine equal keys of a batch before inserting
// (1, default) or insert every occurrence on its own (0).
extern "C" void wc_set_counting_combine(int mode) { detail::set_counting_combine(mode); }
extern "C" int wc_get_counting_combine(void) { return detail::counting_combine(); }

// Key files (warpcore/io.hpp): raw little-endian key arrays.
#define WC_KEYFILE_API(SFX, K)                                                                                    \
    extern "C" int64_t wc_load_binary_##SFX(const char* path, uint64_t max_keys, K* keys_out) {                  \
        std::vector<K> keys;                                                                                      \
        if (!keys_out && max_keys) return -1;                                                                     \
        if (!io::load_binary<K>(path, max_keys, keys)) return -1;                                                 \
        std::copy(keys.begin(), keys.end(), keys_out);                                                            \
        return int64_t(keys.size());                                                                              \
    }                                                                                                             \
    extern "C" int wc_load_keys_##SFX(const char* path, uint64_t max_keys, K* keys_d, uint64_t* num_out,         \
                                     void* stream) {                                                              \
        if (!keys_d && max_keys) return wc_capi::kBadArgument;                                                    \
        const hipError_t e = io::load_keys<K>(path, max_keys, keys_d, num_out, as_stream(stream));                \
        return e == hipErrorInvalidValue ? wc_capi::kBadArgument : int(e);                                        \
    }

WC_KEYFILE_API(u32, uint32_t)
WC_KEYFILE_API(u64, uint64_t)
