// Key files: raw little-endian arrays of sizeof(K)-byte keys, the input
// format of the reference's benchmarks (benchmarks/include/
// benchmark_common.cuh:62-75, load_keys over helpers::load_binary).
//
// load_binary reads at most max_keys keys from the start of the file (fewer
// when the file is shorter; a trailing partial key is ignored). load_keys
// copies them to a device buffer; unlike the reference's load_keys, which
// copies max_keys keys whatever the file held, it copies only what was read
// and reports the count.
#pragma once

#include "base.hpp"

#include <cstdint>
#include <cstdio>
#include <vector>

namespace warpcore {
namespace io {

// Returns false when the file cannot be opened or read.
template<class K>
WC_H bool load_binary(const char* path, std::uint64_t max_keys, std::vector<K>& out) {
    out.clear();
    if (path == nullptr) return false;
    std::FILE* f = std::fopen(path, "rb");
    if (f == nullptr) return false;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long bytes = ok ? std::ftell(f) : -1;
    ok = ok && bytes >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        const std::uint64_t in_file = std::uint64_t(bytes) / sizeof(K);
        const std::uint64_t n = in_file < max_keys ? in_file : max_keys;
        out.resize(n);
        ok = n == 0 || std::fread(out.data(), sizeof(K), n, f) == n;
        if (!ok) out.clear();
    }
    std::fclose(f);
    return ok;
}

// Reads at most max_keys keys into the device buffer keys_d (room for
// max_keys keys); *num_out = keys read. Returns hipSuccess, or
// hipErrorInvalidValue when the file cannot be read.
template<class K>
WC_H hipError_t load_keys(const char* path, std::uint64_t max_keys, K* keys_d, std::uint64_t* num_out,
                          hipStream_t stream = 0) {
    std::vector<K> keys;
    if (num_out) *num_out = 0;
    if (!load_binary<K>(path, max_keys, keys)) return hipErrorInvalidValue;
    if (!keys.empty()) {
        const hipError_t e = hipMemcpyAsync(keys_d, keys.data(), sizeof(K) * keys.size(), hipMemcpyHostToDevice,
                                            stream);
        if (e != hipSuccess) return e;
        const hipError_t s = hipStreamSynchronize(stream);   // keys is a host temporary
        if (s != hipSuccess) return s;
    }
    if (num_out) *num_out = keys.size();
    return hipSuccess;
}

}  // namespace io
}  // namespace warpcore
