// extern "C" entry points of SingleValueHashTable<uint64_t, uint64_t> and
// SingleValueHashTable<uint64_t, uint32_t> (declared in include/warpcore_c.h). One
// handle type per (key, value) pair; the cooperative group size is chosen
// at creation among the compiled sizes.
#include "capi_svht.hpp"

namespace {

using wc_capi::DefaultSvht;
using wc_capi::SvhtApi;
using wc_capi::SvhtImpl;
using warpcore::index_t;

template<class K, class V>
SvhtApi<K, V>* make_svht(index_t min_cap, K seed, bool no_init, unsigned cg) {
    switch (cg) {
    case 0:
    case 8: return new SvhtImpl<DefaultSvht<K, V, 8>>(min_cap, seed, no_init);
    case 1: return new SvhtImpl<DefaultSvht<K, V, 1>>(min_cap, seed, no_init);
    case 2: return new SvhtImpl<DefaultSvht<K, V, 2>>(min_cap, seed, no_init);
    case 4: return new SvhtImpl<DefaultSvht<K, V, 4>>(min_cap, seed, no_init);
    case 16: return new SvhtImpl<DefaultSvht<K, V, 16>>(min_cap, seed, no_init);
    case 32: return new SvhtImpl<DefaultSvht<K, V, 32>>(min_cap, seed, no_init);
    case 64: return new SvhtImpl<DefaultSvht<K, V, 64>>(min_cap, seed, no_init);
    default: return nullptr;
    }
}

template<class K, class V>
SvhtApi<K, V>* make_svht_default_cg(index_t min_cap, K seed, bool no_init, unsigned cg) {
    if (cg != 0 && cg != 8) return nullptr;
    return new SvhtImpl<DefaultSvht<K, V, 8>>(min_cap, seed, no_init);
}

}  // namespace

WC_SVHT_API(u64u64, uint64_t, uint64_t, make_svht)
WC_SVHT_API(u64u32, uint64_t, uint32_t, make_svht_default_cg)
