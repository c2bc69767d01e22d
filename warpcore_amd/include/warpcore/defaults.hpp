// Default sentinels and policy types (include/warpcore/defaults.cuh:11-42):
// EmptyKey 0, TombstoneKey ~0, seed 0x5ad0ded, 2048 B of temp words,
// probing length ~0 - 1024, MurmurHash, DoubleHashing, ReturnNothing, AoS.
#pragma once

#include "base.hpp"
#include "hashers.hpp"
#include "probing_schemes.hpp"
#include "status.hpp"
#include "storage.hpp"

namespace warpcore {
namespace defaults {

template<class Key> WC_HD constexpr Key empty_key() noexcept { return Key{0}; }
template<class Key> WC_HD constexpr Key tombstone_key() noexcept { return ~Key{0}; }
template<class Key> WC_HD constexpr Key seed() noexcept { return Key{0x5ad0ded}; }
WC_HD constexpr index_t temp_memory_bytes() noexcept { return 2048; }
WC_HD constexpr index_t probing_length() noexcept { return ~index_t(0) - 1024; }

template<class Key> using hasher_t = hashers::MurmurHash<Key>;

template<class Key, index_t CGSize>
using probing_scheme_t = probing_schemes::DoubleHashing<hasher_t<Key>, hasher_t<Key>, CGSize>;

using status_handler_t = status_handlers::ReturnNothing;

template<class Key, class Value> using table_storage_t = storage::key_value::AoSStore<Key, Value>;

}  // namespace defaults
}  // namespace warpcore
