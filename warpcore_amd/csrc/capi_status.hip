// extern "C" host evaluation of the Status algebra (declared in
// include/warpcore_c.h): the same constexpr members the kernels use, so a
// harness in another language reads per-key and table status words with the
// reference's semantics (include/warpcore/status.cuh:16-333).
#include "capi_common.hpp"

#include <cstring>

using warpcore::Status;

namespace {
Status st(uint32_t s) noexcept { return Status::from_base(s); }
}  // namespace

extern "C" uint32_t wc_status_all(void) { return Status::all().base(); }
extern "C" uint32_t wc_status_error_mask(void) { return Status::error_mask().base(); }
extern "C" uint32_t wc_status_warning_mask(void) { return Status::warning_mask().base(); }
extern "C" uint32_t wc_status_complement(uint32_t s) { return (~st(s)).base(); }
extern "C" uint32_t wc_status_join(uint32_t a, uint32_t b) { return (st(a) + st(b)).base(); }
extern "C" uint32_t wc_status_remove(uint32_t a, uint32_t b) { return (st(a) - st(b)).base(); }
extern "C" int wc_status_has_any(uint32_t s, uint32_t flags) { return st(s).has_any(st(flags)) ? 1 : 0; }
extern "C" int wc_status_has_all(uint32_t s, uint32_t flags) { return st(s).has_all(st(flags)) ? 1 : 0; }
extern "C" int wc_status_has_any_errors(uint32_t s) { return st(s).has_any_errors() ? 1 : 0; }
extern "C" int wc_status_has_any_warnings(uint32_t s) { return st(s).has_any_warnings() ? 1 : 0; }
extern "C" uint32_t wc_status_get_errors(uint32_t s) { return st(s).get_errors().base(); }
extern "C" uint32_t wc_status_get_warnings(uint32_t s) { return st(s).get_warnings().base(); }
extern "C" int wc_status_return_boolean(uint32_t s, uint32_t ignore) {
    return st(s).has_any(~st(ignore)) ? 1 : 0;   // status_handlers::ReturnBoolean<ignore>
}
extern "C" int wc_status_return_any_flag(uint32_t s, uint32_t ignore) {
    return st(s).has_any(Status::flags_except(st(ignore))) ? 1 : 0;   // status_handlers::ReturnAnyFlag<ignore>
}
extern "C" int wc_status_to_string(uint32_t s, char* buf, uint64_t len) {
    if (buf == nullptr || len == 0) return wc_capi::kBadArgument;
    return wc_capi::guarded([&] {
        const std::string text = warpcore::to_string(st(s));
        const std::size_t n = text.size() < len - 1 ? text.size() : std::size_t(len - 1);
        std::memcpy(buf, text.data(), n);
        buf[n] = '\0';
        return int(text.size());
    });
}
