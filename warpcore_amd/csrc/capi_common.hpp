// Shared helpers of the extern "C" layer: stream casts, error capture and
// exception fences. No exception crosses the C boundary.
#pragma once

#include <warpcore/warpcore.hpp>

#include <new>

namespace wc_capi {

inline hipStream_t as_stream(void* s) noexcept { return reinterpret_cast<hipStream_t>(s); }

// hipError_t of the most recent launch on this thread (0 = success).
inline int last_error() noexcept { return int(hipGetLastError()); }

constexpr int kBadHandle = -1;
constexpr int kBadArgument = -2;
constexpr int kNoMemory = -3;
constexpr int kTooLarge = -4;    // input larger than a fixed-size buffer of the call
constexpr int kNoScratch = -5;   // device address of a static scratch array unavailable

template<class F>
inline int guarded(F&& f) noexcept {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return kNoMemory;
    } catch (...) {
        return kBadArgument;
    }
}

}  // namespace wc_capi
