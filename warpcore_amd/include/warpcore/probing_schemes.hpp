// Cooperative probing schemes over cache-line aligned windows.
//
// The reference iterators (include/warpcore/probing_schemes.cuh:22-302) give
// lane r of a CG the slot (h1 + r + i*step) % capacity, so a window of CG
// consecutive slots starts anywhere and usually straddles two 128-byte lines.
// Here capacity = p * CG (the capacity rule already guarantees it) and the
// table is viewed as p windows of CG slots: probe i reads window w_i and lane
// r takes slot w_i * CG + r. With 16-byte slots and CG = 8, or 8-byte slots
// and CG = 16, a window is exactly one aligned 128-byte line.
//
//   DoubleHashing:    w_0 = h1 % p, w_{i+1} = (w_i + s) % p, s = h2 % (p-1) + 1
//   LinearProbing:    w_{i+1} = (w_i + 1) % p
//   QuadraticProbing: w_{i+1} = (w_i + i + 1) % p
//
// with h1 = Hasher1(key + seed), h2 = Hasher2(key + seed + 1), the same hash
// inputs and step range as the reference. p is prime, so double and linear
// probing visit every window once per cycle (the cycle-free property the
// reference tests, tests/src/probing_schemes_test.cu:44-68). Only slot
// placement differs from the reference; membership, values and statuses do
// not depend on it.
//
// Probe count: max(1, SDIV(probing_length, CG)) windows as in the reference
// (begin() is unchecked, next() stops at SDIV(pl, CG)*CG probes), capped at
// one full cycle of p windows, after which nothing new can be found. The cap
// turns the reference's endless probe of a table with no EMPTY slot (its
// default probing length is ~2^64) into probing_length_exceeded.
#pragma once

#include "base.hpp"
#include "fastmod.hpp"
#include "hashers.hpp"
#include "tile.hpp"

namespace warpcore {
namespace probing_schemes {

// Window geometry of one table, computed on the host.
struct Geometry {
    index_t capacity = 0;   // slots = windows * cg
    index_t windows = 0;    // p
    detail::FastMod mod_windows;
    detail::FastMod mod_windows_minus_one;

    WC_H static Geometry make(index_t capacity, index_t cg) noexcept {
        Geometry g;
        g.capacity = capacity;
        g.windows = cg ? capacity / cg : 0;
        g.mod_windows = detail::FastMod(g.windows ? g.windows : 1);
        g.mod_windows_minus_one = detail::FastMod(g.windows > 1 ? g.windows - 1 : 1);
        return g;
    }
};

namespace detail {
WC_D index_t window_budget(index_t probing_length, index_t cg, index_t windows) noexcept {
    const index_t w = probing_length > ~index_t(0) - cg ? ~index_t(0) / cg + 1 : SDIV(probing_length, cg);
    const index_t n = w == 0 ? 1 : w;
    return n < windows ? n : windows;
}
}  // namespace detail

template<class Hasher1, class Hasher2, index_t CGSize = 1>
class DoubleHashing {
    static_assert(checks::is_valid_cg_size(CGSize), "invalid cooperative group size");
    static_assert(std::is_same<typename Hasher1::key_type, typename Hasher2::key_type>::value,
                  "key types of both hashers must be the same");

public:
    using key_type = typename Hasher1::key_type;
    using index_type = index_t;
    using tag = typename std::conditional<checks::is_true_permutation_hasher<Hasher1>() &&
                                              checks::is_true_permutation_hasher<Hasher2>(),
                                          tags::cycle_free_probing_scheme, tags::probing_scheme>::type;

    WC_HD static constexpr index_type cg_size() noexcept { return CGSize; }

    WC_D explicit DoubleHashing(const Geometry& g, index_type probing_length, const tile<CGSize>&) noexcept
        : geo_(g), left_(detail::window_budget(probing_length, CGSize, g.windows)) {}

    // First slot of this lane for `key`.
    WC_D index_type begin(key_type key, key_type seed = 0) noexcept {
        const key_type h1 = Hasher1::hash(key_type(key + seed));
        const key_type h2 = Hasher2::hash(key_type(key + seed + key_type(1)));
        w_ = geo_.mod_windows.mod(std::uint64_t(h1));
        step_ = geo_.mod_windows_minus_one.mod(std::uint64_t(h2)) + 1;
        return w_ * CGSize + tile<CGSize>::thread_rank();
    }
    // Next slot of this lane, or end() once the budget is spent.
    WC_D index_type next() noexcept {
        if (--left_ == 0) return end();
        w_ += step_;
        if (w_ >= geo_.windows) w_ -= geo_.windows;
        return w_ * CGSize + tile<CGSize>::thread_rank();
    }
    WC_D index_type window() const noexcept { return w_; }
    WC_HD static constexpr index_type end() noexcept { return ~index_type(0); }

private:
    Geometry geo_;
    index_type left_;
    index_type w_ = 0;
    index_type step_ = 1;
};

template<class Hasher, index_t CGSize = 1>
class LinearProbing {
    static_assert(checks::is_valid_cg_size(CGSize), "invalid cooperative group size");

public:
    using key_type = typename Hasher::key_type;
    using index_type = index_t;
    using tag = tags::cycle_free_probing_scheme;

    WC_HD static constexpr index_type cg_size() noexcept { return CGSize; }

    WC_D explicit LinearProbing(const Geometry& g, index_type probing_length, const tile<CGSize>&) noexcept
        : geo_(g), left_(detail::window_budget(probing_length, CGSize, g.windows)) {}

    WC_D index_type begin(key_type key, key_type seed = 0) noexcept {
        w_ = geo_.mod_windows.mod(std::uint64_t(Hasher::hash(key_type(key + seed))));
        return w_ * CGSize + tile<CGSize>::thread_rank();
    }
    WC_D index_type next() noexcept {
        if (--left_ == 0) return end();
        if (++w_ == geo_.windows) w_ = 0;
        return w_ * CGSize + tile<CGSize>::thread_rank();
    }
    WC_D index_type window() const noexcept { return w_; }
    WC_HD static constexpr index_type end() noexcept { return ~index_type(0); }

private:
    Geometry geo_;
    index_type left_;
    index_type w_ = 0;
};

template<class Hasher, index_t CGSize = 1>
class QuadraticProbing {
    static_assert(checks::is_valid_cg_size(CGSize), "invalid cooperative group size");

public:
    using key_type = typename Hasher::key_type;
    using index_type = index_t;
    using tag = tags::probing_scheme;

    WC_HD static constexpr index_type cg_size() noexcept { return CGSize; }

    WC_D explicit QuadraticProbing(const Geometry& g, index_type probing_length, const tile<CGSize>&) noexcept
        : geo_(g), left_(detail::window_budget(probing_length, CGSize, g.windows)) {}

    WC_D index_type begin(key_type key, key_type seed = 0) noexcept {
        w_ = geo_.mod_windows.mod(std::uint64_t(Hasher::hash(key_type(key + seed))));
        step_ = 1;
        return w_ * CGSize + tile<CGSize>::thread_rank();
    }
    WC_D index_type next() noexcept {
        if (--left_ == 0) return end();
        w_ = geo_.mod_windows.mod(w_ + step_);
        ++step_;
        return w_ * CGSize + tile<CGSize>::thread_rank();
    }
    WC_D index_type window() const noexcept { return w_; }
    WC_HD static constexpr index_type end() noexcept { return ~index_type(0); }

private:
    Geometry geo_;
    index_type left_;
    index_type w_ = 0;
    index_type step_ = 1;
};

}  // namespace probing_schemes
}  // namespace warpcore
