// wc_svht_*_create_custom: SingleValueHashTables with a probing scheme,
// hasher or slot storage other than the defaults (declared in
// include/warpcore_c.h). The handle is the one of wc_svht_*_create, so every
// other wc_svht_* entry point works on it. The compiled combinations, all
// with cooperative groups of 8 lanes:
//   DoubleHashing<H, H, 8>   H = MurmurHash, MuellerHash, NvidiaHash, IdentityMap
//   LinearProbing<H, 8>      H = MurmurHash, IdentityMap
//   QuadraticProbing<H, 8>   H = MurmurHash, IdentityMap
//   DoubleHashing<MurmurHash, MurmurHash, 8> over SoAStore
// MuellerHash and NvidiaHash exist for 32-bit keys only
// (include/warpcore/hashers.cuh:14-79).
#include "capi_svht.hpp"

namespace {

using namespace warpcore;
using wc_capi::SvhtApi;
using wc_capi::SvhtImpl;

enum : int { kDouble = 0, kLinear = 1, kQuadratic = 2 };
enum : int { kMurmur = 0, kMueller = 1, kNvidia = 2, kIdentity = 3 };
enum : int { kAoS = 0, kSoA = 1 };

template<class K, class V, class Scheme, template<class, class> class Store = storage::key_value::AoSStore>
using Table = SingleValueHashTable<K, V, defaults::empty_key<K>(), defaults::tombstone_key<K>(), Scheme, Store<K, V>>;

template<class K, class V, class H>
SvhtApi<K, V>* make_with_hasher(index_t min_cap, K seed, bool no_init, int probing) {
    using namespace probing_schemes;
    switch (probing) {
    case kDouble: return new SvhtImpl<Table<K, V, DoubleHashing<H, H, 8>>>(min_cap, seed, no_init);
    case kLinear: return new SvhtImpl<Table<K, V, LinearProbing<H, 8>>>(min_cap, seed, no_init);
    case kQuadratic: return new SvhtImpl<Table<K, V, QuadraticProbing<H, 8>>>(min_cap, seed, no_init);
    default: return nullptr;
    }
}

template<class K, class V>
SvhtApi<K, V>* make_custom(index_t min_cap, K seed, bool no_init, unsigned cg, int probing, int hasher1, int hasher2,
                           int storage) {
    using namespace probing_schemes;
    using Murmur = hashers::MurmurHash<K>;
    if (cg != 0 && cg != 8) return nullptr;
    if (hasher1 != hasher2) return nullptr;
    if (storage == kSoA) {
        if (probing != kDouble || hasher1 != kMurmur) return nullptr;
        return new SvhtImpl<Table<K, V, DoubleHashing<Murmur, Murmur, 8>, storage::key_value::SoAStore>>(min_cap, seed,
                                                                                                       no_init);
    }
    if (storage != kAoS) return nullptr;
    switch (hasher1) {
    case kMurmur: return make_with_hasher<K, V, Murmur>(min_cap, seed, no_init, probing);
    case kIdentity:
        if (probing == kDouble)
            return new SvhtImpl<Table<K, V, DoubleHashing<hashers::IdentityMap<K, K>, hashers::IdentityMap<K, K>, 8>>>(
                min_cap, seed, no_init);
        return make_with_hasher<K, V, hashers::IdentityMap<K, K>>(min_cap, seed, no_init, probing);
    case kMueller:
    case kNvidia:
        if constexpr (std::is_same<K, std::uint32_t>::value) {
            if (probing != kDouble) return nullptr;
            if (hasher1 == kMueller)
                return new SvhtImpl<Table<K, V, DoubleHashing<hashers::MuellerHash, hashers::MuellerHash, 8>>>(
                    min_cap, seed, no_init);
            return new SvhtImpl<Table<K, V, DoubleHashing<hashers::NvidiaHash, hashers::NvidiaHash, 8>>>(
                min_cap, seed, no_init);
        }
        return nullptr;
    default: return nullptr;
    }
}

}  // namespace

extern "C" void* wc_svht_u32u32_create_custom(uint64_t min_capacity, uint32_t seed, int no_init, uint32_t cg_size,
                                              int probing, int hasher1, int hasher2, int storage) {
    try {
        return make_custom<std::uint32_t, std::uint32_t>(min_capacity, seed, no_init != 0, cg_size, probing, hasher1,
                                                         hasher2, storage);
    } catch (...) {
        return nullptr;
    }
}

extern "C" void* wc_svht_u64u64_create_custom(uint64_t min_capacity, uint64_t seed, int no_init, uint32_t cg_size,
                                              int probing, int hasher1, int hasher2, int storage) {
    try {
        return make_custom<std::uint64_t, std::uint64_t>(min_capacity, seed, no_init != 0, cg_size, probing, hasher1,
                                                         hasher2, storage);
    } catch (...) {
        return nullptr;
    }
}
