// Capacity rule: capacity = p * cg with p the smallest entry >= SDIV(min, cg)
// of the prime table, or 0 past its end (include/warpcore/base.cuh:27-34).
//
// The table is the one include/warpcore/primes.hpp stores as 131036
// constants; its generating rule (utils/generate_primes_hpp.py:20-43) is
// "first prime >= 131072, then the next prime >= previous + 131072, up to
// 17177772037". Instead of shipping 131k constants, this header extends that
// chain lazily with a deterministic Miller-Rabin test; a request for a
// capacity of C slots walks about C / (131072 * cg) links once per process.
#pragma once

#include "base.hpp"

namespace warpcore {
namespace detail {

struct prime_rule {
    static constexpr std::uint64_t min = 131072;
    static constexpr std::uint64_t step = 131072;
    static constexpr std::uint64_t max = 17177772037ull;
};

WC_H std::uint64_t mulmod_u64(std::uint64_t a, std::uint64_t b, std::uint64_t m) noexcept {
    return std::uint64_t((unsigned __int128)a * b % m);
}

// Deterministic for all 64-bit n (bases cover n < 3.3e24).
WC_H bool is_prime_u64(std::uint64_t n) noexcept {
    if (n < 2) return false;
    static constexpr std::uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41};
    for (auto p : bases) if (n % p == 0) return n == p;
    std::uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (auto a : bases) {
        std::uint64_t x = 1, b = a % n, e = d;
        while (e) { if (e & 1) x = mulmod_u64(x, b, n); b = mulmod_u64(b, b, n); e >>= 1; }
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int r = 1; r < s && composite; ++r) { x = mulmod_u64(x, x, n); composite = x != n - 1; }
        if (composite) return false;
    }
    return true;
}

WC_H std::uint64_t next_prime_at_least(std::uint64_t x) noexcept {
    if (x <= 2) return 2;
    if ((x & 1) == 0) ++x;
    while (!is_prime_u64(x)) x += 2;
    return x;
}

class PrimeChain {
public:
    static PrimeChain& instance() noexcept { static PrimeChain c; return c; }

    // Smallest chain entry >= x, or 0 if x is beyond the last entry.
    std::uint64_t lower_bound(std::uint64_t x) noexcept {
        std::lock_guard<std::mutex> lock(mu_);
        while (!complete_ && chain_.back() < x) extend();
        auto it = std::lower_bound(chain_.begin(), chain_.end(), x);
        return it == chain_.end() ? 0 : *it;
    }
    std::uint64_t size() noexcept { complete(); return chain_.size(); }
    std::uint64_t at(std::uint64_t i) noexcept { complete(); return i < chain_.size() ? chain_[i] : 0; }

private:
    PrimeChain() { chain_.push_back(next_prime_at_least(prime_rule::min)); }
    void extend() noexcept {
        const std::uint64_t p = next_prime_at_least(chain_.back() + prime_rule::step);
        if (p > prime_rule::max) complete_ = true;
        else chain_.push_back(p);
    }
    void complete() noexcept {
        std::lock_guard<std::mutex> lock(mu_);
        while (!complete_) extend();
    }
    std::mutex mu_;
    std::vector<std::uint64_t> chain_;
    bool complete_ = false;
};

// include/warpcore/base.cuh:27-34
WC_H index_t get_valid_capacity(index_t min_capacity, index_t cg_size) noexcept {
    const index_t x = SDIV(min_capacity, cg_size);
    const std::uint64_t p = PrimeChain::instance().lower_bound(x);
    return p * cg_size;
}

}  // namespace detail
}  // namespace warpcore
