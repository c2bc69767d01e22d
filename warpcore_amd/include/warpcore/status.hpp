// Status flags and per-key status handlers.
//
// Same 13 flags, bit positions, error/warning masks and algebra as
// include/warpcore/status.cuh:16-214 (flags 39-68, masks 71-86, operator~
// 146-150, atomic_join 187-193, mask 207), and the same three handlers
// (ReturnNothing 278-292, ReturnStatus 296-312, ReturnBoolean 316-333). The
// bit layout is an interface: callers read the words these tables write, so
// it is kept identical, and so is the algebra, quirks included:
//   - `mask` (what all() returns) is 0xFFF, so invalid_value (bit 12) is in
//     neither all() nor warning_mask();
//   - ~s is (~s) ^ 0xFFF: the low 12 bits keep their value and bits 12..31
//     flip, so ~Status::none() is 0xFFFFF000;
//   - therefore ReturnBoolean<Ignore> writes has_any(~Status(Ignore)): with
//     the default Ignore = none it is true only for invalid_value, and an
//     Ignore flag below bit 12 makes that flag count instead of be ignored.
// ReturnAnyFlag<Ignore> (no reference counterpart) is the handler the
// reference's comment describes: true iff any of the 13 flags outside
// Ignore is set. Callers opt into it explicitly.
#pragma once

#include "base.hpp"
#include "tags.hpp"

#include <ostream>
#include <string>
#include <vector>

namespace warpcore {

namespace status_handlers { template<status_base_t Ignore> class ReturnBoolean; }

class Status {
public:
    using base_type = status_base_t;

    Status() noexcept = default;
    constexpr Status(const Status&) noexcept = default;
    constexpr Status(Status&&) noexcept = default;
    constexpr Status& operator=(const Status&) noexcept = default;

    WC_HD constexpr operator base_type() const noexcept { return status_; }
    WC_HD constexpr base_type base() const noexcept { return status_; }

    WC_HD static constexpr Status none() noexcept { return Status(base_type(0)); }
    WC_HD static constexpr Status all() noexcept { return Status(mask); }
    WC_HD static constexpr Status unknown_error() noexcept { return Status(one << 0); }
    WC_HD static constexpr Status probing_length_exceeded() noexcept { return Status(one << 1); }
    WC_HD static constexpr Status invalid_configuration() noexcept { return Status(one << 2); }
    WC_HD static constexpr Status invalid_key() noexcept { return Status(one << 3); }
    WC_HD static constexpr Status duplicate_key() noexcept { return Status(one << 4); }
    WC_HD static constexpr Status key_not_found() noexcept { return Status(one << 5); }
    WC_HD static constexpr Status index_overflow() noexcept { return Status(one << 6); }
    WC_HD static constexpr Status out_of_memory() noexcept { return Status(one << 7); }
    WC_HD static constexpr Status not_initialized() noexcept { return Status(one << 8); }
    WC_HD static constexpr Status dry_run() noexcept { return Status(one << 9); }
    WC_HD static constexpr Status invalid_phase_overlap() noexcept { return Status(one << 10); }
    WC_HD static constexpr Status max_values_for_key_reached() noexcept { return Status(one << 11); }
    WC_HD static constexpr Status invalid_value() noexcept { return Status(one << 12); }

    WC_HD static constexpr Status error_mask() noexcept {
        return Status(unknown_error().status_ | invalid_configuration().status_ | index_overflow().status_ |
                      out_of_memory().status_ | not_initialized().status_ | invalid_phase_overlap().status_);
    }
    WC_HD static constexpr Status warning_mask() noexcept { return Status(mask & ~error_mask().status_); }

    WC_HD constexpr Status get_errors() const noexcept { return Status(status_ & error_mask().status_); }
    WC_HD constexpr Status get_warnings() const noexcept { return Status(status_ & warning_mask().status_); }

    WC_HD constexpr bool has_any(const Status& s = all()) const noexcept { return (status_ & s.status_) != 0; }
    WC_HD constexpr bool has_any_errors() const noexcept { return has_any(error_mask()); }
    WC_HD constexpr bool has_any_warnings() const noexcept { return has_any(warning_mask()); }
    WC_HD constexpr bool has_all(const Status& s = all()) const noexcept { return (status_ & s.status_) == s.status_; }

    WC_HD constexpr bool has_unknown_error() const noexcept { return has_any(unknown_error()); }
    WC_HD constexpr bool has_probing_length_exceeded() const noexcept { return has_any(probing_length_exceeded()); }
    WC_HD constexpr bool has_invalid_configuration() const noexcept { return has_any(invalid_configuration()); }
    WC_HD constexpr bool has_invalid_key() const noexcept { return has_any(invalid_key()); }
    WC_HD constexpr bool has_duplicate_key() const noexcept { return has_any(duplicate_key()); }
    WC_HD constexpr bool has_key_not_found() const noexcept { return has_any(key_not_found()); }
    WC_HD constexpr bool has_index_overflow() const noexcept { return has_any(index_overflow()); }
    WC_HD constexpr bool has_out_of_memory() const noexcept { return has_any(out_of_memory()); }
    WC_HD constexpr bool has_not_initialized() const noexcept { return has_any(not_initialized()); }
    WC_HD constexpr bool has_dry_run() const noexcept { return has_any(dry_run()); }
    WC_HD constexpr bool has_invalid_phase_overlap() const noexcept { return has_any(invalid_phase_overlap()); }
    WC_HD constexpr bool has_max_values_for_key_reached() const noexcept { return has_any(max_values_for_key_reached()); }
    WC_HD constexpr bool has_invalid_value() const noexcept { return has_any(invalid_value()); }

    // (~s) ^ mask (status.cuh:146-150): bits 0..11 unchanged, 12..31 flipped
    WC_HD constexpr Status operator~() const noexcept { return Status((~status_) ^ mask); }
    WC_HD constexpr Status& operator+=(const Status& a) noexcept { status_ |= a.status_; return *this; }
    WC_HD constexpr Status operator+(const Status& a) const noexcept { return Status(status_ | a.status_); }
    WC_HD constexpr Status operator-(const Status& a) const noexcept { return Status(status_ & ~a.status_); }
    WC_HD constexpr Status& operator-=(const Status& a) noexcept { status_ &= ~a.status_; return *this; }
    WC_HD constexpr bool operator==(const Status& a) const noexcept { return status_ == a.status_; }
    WC_HD constexpr bool operator!=(const Status& a) const noexcept { return status_ != a.status_; }

    // Join flags into a status word in device memory; skips the atomic when
    // every flag is already set (status.cuh:187-193). The read goes to L2 so
    // a stale L1 copy cannot force needless atomics.
    WC_D void atomic_join(Status s) noexcept {
        const base_type cur = __hip_atomic_load(&status_, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((cur & s.status_) != s.status_) atomicOr(&status_, s.status_);
    }
    WC_D void atomic_assign(Status s) noexcept { atomicExch(&status_, s.status_); }

    WC_HD static constexpr Status from_base(base_type s) noexcept { return Status(s); }

    // Every defined flag (bits 0..12) except those of `s`; unlike ~s this is
    // the set-theoretic complement within the 13 flags. No reference
    // counterpart.
    WC_HD static constexpr Status flags_except(Status s) noexcept { return Status(defined_flags & ~s.status_); }

private:
    WC_HD explicit constexpr Status(base_type s) noexcept : status_(s) {}

    static constexpr base_type one = 1;
    static constexpr base_type mask = 0xFFF;            // status.cuh:207
    static constexpr base_type defined_flags = 0x1FFF;  // bits 0..12

    base_type status_;

    template<base_type> friend class status_handlers::ReturnBoolean;
};

static_assert(sizeof(Status) == sizeof(status_base_t), "Status must be one status word");
static_assert(std::is_trivially_copyable<Status>::value, "Status must be trivially copyable");

// Names of the set flags, e.g. "[duplicate key, key not found]" (status.cuh:216-269)
inline std::string to_string(Status status) {
    std::vector<std::string> msg;
    if (status.has_duplicate_key()) msg.push_back("duplicate key");
    if (status.has_unknown_error()) msg.push_back("unknown error");
    if (status.has_probing_length_exceeded()) msg.push_back("probing length exceeded");
    if (status.has_invalid_configuration()) msg.push_back("invalid configuration");
    if (status.has_invalid_key()) msg.push_back("invalid key");
    if (status.has_key_not_found()) msg.push_back("key not found");
    if (status.has_index_overflow()) msg.push_back("index overflow");
    if (status.has_out_of_memory()) msg.push_back("out of memory");
    if (status.has_not_initialized()) msg.push_back("not initialized");
    if (status.has_dry_run()) msg.push_back("dry run");
    if (status.has_invalid_phase_overlap()) msg.push_back("invalid phase overlap");
    if (status.has_max_values_for_key_reached()) msg.push_back("max values for key reached");
    if (status.has_invalid_value()) msg.push_back("invalid value");
    std::string out = "[";
    for (std::size_t i = 0; i < msg.size(); ++i) out += (i ? ", " : "") + msg[i];
    return out + "]";
}

inline std::ostream& operator<<(std::ostream& os, Status status) { return os << to_string(status); }

namespace status_handlers {

// No per-key output (status.cuh:278-292).
class ReturnNothing {
public:
    using base_type = void;
    using tag = tags::status_handler;
    WC_D static void handle(Status, base_type*, index_t) noexcept {}
};

// One Status word per key (status.cuh:296-312).
class ReturnStatus {
public:
    using base_type = Status;
    using tag = tags::status_handler;
    WC_D static void handle(Status status, base_type* out, index_t index) noexcept {
        if (out != nullptr) out[index] = status;
    }
};

// One bool per key: status.has_any(~Status(Ignore)) with the reference's
// operator~ (status.cuh:316-333), bit for bit. With Ignore = none only
// invalid_value (and bits above the 13 flags) make it true.
template<Status::base_type Ignore = 0>
class ReturnBoolean {
public:
    using base_type = bool;
    using tag = tags::status_handler;
    WC_D static void handle(Status status, base_type* out, index_t index) noexcept {
        if (out != nullptr) out[index] = status.has_any(~Status(Ignore));
    }
};

// One bool per key: true iff any of the 13 flags outside `Ignore` is set.
// No reference counterpart; the opt-in alternative to ReturnBoolean<>.
template<Status::base_type Ignore = 0>
class ReturnAnyFlag {
public:
    using base_type = bool;
    using tag = tags::status_handler;
    WC_D static void handle(Status status, base_type* out, index_t index) noexcept {
        if (out != nullptr) out[index] = status.has_any(Status::flags_except(Status::from_base(Ignore)));
    }
};

}  // namespace status_handlers

}  // namespace warpcore
