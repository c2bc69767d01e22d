// Integer hash functions (include/warpcore/hashers.cuh:14-175). MurmurHash3's
// fmix32/fmix64 finalizers are the defaults and must be bit-identical; the
// constants and shifts below are the published MurmurHash3 finalizers.
#pragma once

#include "base.hpp"
#include "tags.hpp"

namespace warpcore {
namespace hashers {

// hashers.cuh:14-46 (u32 only, not a permutation tag)
class NvidiaHash {
public:
    using key_type = std::uint32_t;
    using hash_type = std::uint32_t;
    using tag = tags::hasher;
    template<class T> WC_HD static hash_type hash(T) = delete;
    WC_HD static hash_type hash(key_type x) noexcept {
        x = (x + 0x7ed55d16u) + (x << 12);
        x = (x ^ 0xc761c23cu) ^ (x >> 19);
        x = (x + 0x165667b1u) + (x << 5);
        x = (x + 0xd3a2646cu) ^ (x << 9);
        x = (x + 0xfd7046c5u) + (x << 3);
        x = (x ^ 0xb55a4f09u) ^ (x >> 16);
        return x;
    }
};

// hashers.cuh:50-79 (u32 only)
class MuellerHash {
public:
    using key_type = std::uint32_t;
    using hash_type = std::uint32_t;
    using tag = tags::true_permutation_hasher;
    template<class T> WC_HD static hash_type hash(T) = delete;
    WC_HD static hash_type hash(key_type x) noexcept {
        x = ((x >> 16) ^ x) * 0x45d9f3bu;
        x = ((x >> 16) ^ x) * 0x45d9f3bu;
        x = ((x >> 16) ^ x);
        return x;
    }
};

// MurmurHash3 finalizer, hashers.cuh:85-141
template<class K>
class MurmurHash {
public:
    using key_type = K;
    using hash_type = K;
    using tag = tags::true_permutation_hasher;

    template<class T>
    WC_HD static T hash(T x) noexcept {
        static_assert(std::is_same<T, key_type>::value, "invalid key type");
        return mix(x);
    }

private:
    WC_HD static std::uint32_t mix(std::uint32_t x) noexcept {
        x ^= x >> 16;
        x *= 0x85ebca6bu;
        x ^= x >> 13;
        x *= 0xc2b2ae35u;
        x ^= x >> 16;
        return x;
    }
    WC_HD static std::uint64_t mix(std::uint64_t x) noexcept {
        x ^= x >> 33;
        x *= 0xff51afd7ed558ccdull;
        x ^= x >> 33;
        x *= 0xc4ceb9fe1a85ec53ull;
        x ^= x >> 33;
        return x;
    }
};

// hashers.cuh:147-175
template<class K, class H = std::uint32_t>
class IdentityMap {
public:
    using key_type = K;
    using hash_type = H;
    using tag = tags::true_permutation_hasher;
    static_assert(std::is_same<H, std::uint32_t>::value || std::is_same<H, std::uint64_t>::value,
                  "invalid hash type");
    static_assert(std::is_convertible<K, H>::value, "key type not convertible to hash type");
    WC_HD static hash_type hash(key_type x) noexcept { return hash_type(x); }
};

}  // namespace hashers
}  // namespace warpcore
