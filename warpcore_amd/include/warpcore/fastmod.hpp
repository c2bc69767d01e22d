// Exact 64-bit remainder by a runtime divisor without a division
// instruction: Barrett reduction with m = floor(2^64 / d), which leaves the
// quotient estimate at most one short, so one conditional subtraction gives
// x % d exactly for every 64-bit x. CDNA has no 64-bit integer divide; the
// compiler's generic `%` is a long call-free sequence per probe, this is
// one 64x64 high multiply plus a few ALU ops. Results equal `x % d`
// bit for bit (checked against `%` in the tests).
#pragma once

#include "base.hpp"

namespace warpcore {
namespace detail {

class FastMod {
public:
    WC_HD FastMod() noexcept : d_(1), m_(0) {}
    WC_H explicit FastMod(std::uint64_t d) noexcept : d_(d ? d : 1), m_(0) {
        if (d_ > 1) m_ = std::uint64_t(((unsigned __int128)1 << 64) / d_);
    }
    WC_HD std::uint64_t divisor() const noexcept { return d_; }
    WC_HD std::uint64_t magic() const noexcept { return m_; }

    WC_HD std::uint64_t mod(std::uint64_t x) const noexcept {
        if (d_ == 1) return 0;
        const std::uint64_t q = mulhi(x, m_);
        std::uint64_t r = x - q * d_;
        return r >= d_ ? r - d_ : r;
    }

    WC_HD static std::uint64_t mulhi(std::uint64_t a, std::uint64_t b) noexcept {
#if defined(__HIP_DEVICE_COMPILE__)
        return __umul64hi(a, b);
#else
        return std::uint64_t(((unsigned __int128)a * b) >> 64);
#endif
    }

private:
    std::uint64_t d_;
    std::uint64_t m_;
};

}  // namespace detail
}  // namespace warpcore
