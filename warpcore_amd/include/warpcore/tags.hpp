// Compile-time concept tags and checks (include/warpcore/tags.cuh:12-19,
// include/warpcore/checks.cuh:13-118). Cooperative groups may span a whole
// 64-lane wavefront here, so cg sizes 1..64 are valid.
#pragma once

#include "base.hpp"

namespace warpcore {

namespace tags {
struct hasher {};
struct true_permutation_hasher {};
struct probing_scheme {};
struct cycle_free_probing_scheme {};
struct key_value_storage {};
struct static_value_storage {};
struct dynamic_value_storage {};
struct status_handler {};
}  // namespace tags

namespace checks {

WC_HD constexpr bool is_valid_cg_size(index_t cgsize) noexcept {
    return cgsize == 1 || cgsize == 2 || cgsize == 4 || cgsize == 8 || cgsize == 16 || cgsize == 32 ||
           cgsize == 64;
}
template<class T> WC_HD constexpr bool is_valid_key_type() noexcept {
    return std::is_same<T, std::uint32_t>::value || std::is_same<T, std::uint64_t>::value;
}
template<class T> WC_HD constexpr bool is_valid_slot_type() noexcept { return is_valid_key_type<T>(); }
template<class T> WC_HD constexpr bool is_valid_value_type() noexcept { return std::is_trivially_copyable<T>::value; }
template<class T> WC_HD constexpr bool is_valid_counter_type() noexcept {
    return std::is_same<T, int>::value || std::is_same<T, std::uint32_t>::value ||
           std::is_same<T, std::uint64_t>::value;
}
template<class T> WC_HD constexpr bool is_hasher() noexcept {
    return std::is_same<typename T::tag, tags::hasher>::value ||
           std::is_same<typename T::tag, tags::true_permutation_hasher>::value;
}
template<class T> WC_HD constexpr bool is_true_permutation_hasher() noexcept {
    return std::is_same<typename T::tag, tags::true_permutation_hasher>::value;
}
template<class T> WC_HD constexpr bool is_probing_scheme() noexcept {
    return std::is_same<typename T::tag, tags::probing_scheme>::value ||
           std::is_same<typename T::tag, tags::cycle_free_probing_scheme>::value;
}
template<class T> WC_HD constexpr bool is_cycle_free_probing_scheme() noexcept {
    return std::is_same<typename T::tag, tags::cycle_free_probing_scheme>::value;
}
template<class T> WC_HD constexpr bool is_key_value_storage() noexcept {
    return std::is_same<typename T::tag, tags::key_value_storage>::value;
}
template<class T> WC_HD constexpr bool is_value_storage() noexcept {
    return std::is_same<typename T::tag, tags::static_value_storage>::value ||
           std::is_same<typename T::tag, tags::dynamic_value_storage>::value;
}
template<class T> WC_HD constexpr bool is_status_handler() noexcept {
    return std::is_same<typename T::tag, tags::status_handler>::value;
}

}  // namespace checks

}  // namespace warpcore
