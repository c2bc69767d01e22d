// Streaming probing: each cooperative group keeps U keys in flight.
//
// A one-key-per-group loop leaves one outstanding line request per group
// and serializes the key load, the window load and the CAS of every key, so
// a wavefront waits out a full memory latency per key. Here each group owns
// U slots. In one round every slot that needs a window issues its load and
// every slot that owns an EMPTY line lane issues its CAS, all before any
// result is consumed; then each slot moves on according to what it saw.
// A slot whose key is finished writes its result and takes the group's next
// key at once, so a slow key (a miss at load factor 0.9 that walks several
// windows) does not hold up the other slots: U real requests stay in flight
// per group.
//
// Input keys come through a per-group buffer of CG keys, one per lane: each
// lane hashes its own key and computes its first window and step (two fmix
// and two Barrett reductions per key, 64 keys per wavefront instruction
// instead of 64/CG), and a slot that takes buffer entry j receives key,
// window and step from lane j by shuffle. The next buffer's keys are
// re-read every round so that no load sits under a branch.
//
// Per key the sequence of reads and CASes is exactly the one-key protocol
// of the reference (single_value_hash_table.cuh:742-823 for insert,
// :267-307 for retrieve): the same windows in the same order, a hit means
// duplicate, the EMPTY lanes of a window are tried lowest first, a CAS that
// returns the key itself is a duplicate. Keys of one group interleave like
// keys of different groups do in the reference, so results are unchanged;
// two copies of one key in the same batch issue their CASes on the same
// slot in order from the same lane, and the second sees the first.
#pragma once

#include "engine.hpp"
#include "tile.hpp"

namespace warpcore {
namespace detail {

// Cores of multi-value tables (static constexpr bool multi_value = true)
// change the insert protocol: every pair takes a new slot, the key's pairs
// met on the way are counted (multi_value_hash_table.cuh:200-289).
template<class C, class = void> struct is_multi_value_core : std::false_type {};
template<class C>
struct is_multi_value_core<C, std::void_t<decltype(C::multi_value)>> : std::bool_constant<C::multi_value> {};

// Overflow filter of a table (cores with a member `filter`: one byte per
// probe window, a bit set for every key whose probing passed the window
// full, the bit chosen by the key's window step). A lookup at a full window
// without the key whose bit is clear knows the key is in no later window.
// Cores without a filter get no-ops here.
template<class C, class = void> struct has_overflow_filter : std::false_type {};
template<class C>
struct has_overflow_filter<C, std::void_t<decltype(std::declval<const C&>().filter)>> : std::true_type {};

WC_HD constexpr std::uint32_t overflow_bit(index_t w, index_t step) noexcept {
    return (std::uint32_t(1) << (step & 7)) << (8 * (w & 3));
}

// Marks window w as passed by the key of window step `step` (one
// non-returning atomic OR).
template<class Core>
WC_D void mark_overflow(const Core& core, index_t w, index_t step) noexcept {
    if constexpr (has_overflow_filter<Core>::value) {
        if (core.filter != nullptr)
            (void)__hip_atomic_fetch_or(core.filter + (w >> 2), overflow_bit(w, step), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Window iterator state of one key: current window, step (double hashing:
// fixed; quadratic: growing; linear: 1) and windows left.
template<class ProbingScheme>
struct WindowCursor;

template<class H1, class H2, index_t CG>
struct WindowCursor<probing_schemes::DoubleHashing<H1, H2, CG>> {
    index_t w, step, left;
    template<class K>
    WC_D void begin(K key, K seed, const probing_schemes::Geometry& g) noexcept {
        const K h1 = H1::hash(K(key + seed));
        const K h2 = H2::hash(K(key + seed + K(1)));
        w = g.mod_windows.mod(std::uint64_t(h1));
        step = g.mod_windows_minus_one.mod(std::uint64_t(h2)) + 1;
    }
    WC_D bool advance(const probing_schemes::Geometry& g) noexcept {
        if (--left == 0) return false;
        w += step;
        if (w >= g.windows) w -= g.windows;
        return true;
    }
};

template<class H, index_t CG>
struct WindowCursor<probing_schemes::LinearProbing<H, CG>> {
    index_t w, step, left;
    template<class K>
    WC_D void begin(K key, K seed, const probing_schemes::Geometry& g) noexcept {
        w = g.mod_windows.mod(std::uint64_t(H::hash(K(key + seed))));
        step = 1;
    }
    WC_D bool advance(const probing_schemes::Geometry& g) noexcept {
        if (--left == 0) return false;
        if (++w == g.windows) w = 0;
        return true;
    }
};

template<class H, index_t CG>
struct WindowCursor<probing_schemes::QuadraticProbing<H, CG>> {
    index_t w, step, left;
    template<class K>
    WC_D void begin(K key, K seed, const probing_schemes::Geometry& g) noexcept {
        w = g.mod_windows.mod(std::uint64_t(H::hash(K(key + seed))));
        step = 1;
    }
    WC_D bool advance(const probing_schemes::Geometry& g) noexcept {
        if (--left == 0) return false;
        w = g.mod_windows.mod(w + step);
        ++step;
        return true;
    }
};

// Per-slot phase: kDone no key left; kLoad window read pending; kCas claim
// pending on the lowest untried EMPTY lane; kEmit result ready to write;
// kWait waiting for the next input buffer.
enum : std::uint32_t { kDone = 0, kLoad = 1, kCas = 2, kEmit = 3, kWait = 4 };

// State of a buffer entry as prepared by its owning lane.
enum : std::uint32_t { kKeyOk = 0, kKeyInvalid = 1, kNotReady = 2, kPastEnd = 3 };

// The per-group input buffer of the streaming kernels: lane r holds input
// entry base + r and its prepared cursor.
template<class Core, unsigned CG, bool HasValues>
struct KeyBuffer {
    using K = typename Core::key_type;
    using V = typename Core::value_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    index_t base, stride;
    K key, next_key;
    V val, next_val;
    Cursor cur;
    std::uint32_t state;
    unsigned taken;       // entries handed out from this buffer
    bool next_ready;      // next_key was re-read this round and not consumed

    WC_D void prepare(const Core& core, index_t num_in, unsigned rank) noexcept {
        const index_t i = base + rank;
        cur.w = 0;
        cur.step = 1;
        if (i >= num_in) state = kPastEnd;
        else if (!core.ready()) state = kNotReady;
        else if (!core.is_valid_key(key)) state = kKeyInvalid;
        else { state = kKeyOk; cur.begin(key, core.hash_seed(), core.geometry()); }
    }
    // re-read the next buffer (once per round, unconditionally)
    WC_D void reload(const typename Core::key_type* keys_in, const V* values_in, index_t num_in,
                     unsigned rank) noexcept {
        const index_t i = base + stride + rank;
        const index_t at = i < num_in ? i : num_in - 1;
        next_key = keys_in[at];
        if constexpr (HasValues) next_val = values_in[at];
        next_ready = true;
    }
    // switch to the next buffer if this one is used up; false if the next
    // keys are not there yet
    WC_D bool ensure(const Core& core, index_t num_in, unsigned rank) noexcept {
        if (taken < CG) return true;
        if (!next_ready) return false;
        base += stride;
        key = next_key;
        if constexpr (HasValues) val = next_val;
        next_ready = false;
        taken = 0;
        prepare(core, num_in, rank);
        return true;
    }
};

}  // namespace detail

namespace kernels {

// Lookup with U keys per group in flight. `Core` (a table view) provides
// the probed element type probe_slot_type, slot_ptr(), key_of() / value_of()
// of a probed element, ready(), hash_seed(), geometry(), is_valid_key(),
// is_empty_key(), status_word(), found(raw value) -> reported value and
// missing(out) -> status of a window with an EMPTY slot.
template<class Core, class StatusHandler, int U>
__global__ __launch_bounds__(256) void retrieve_batched(const typename Core::key_type* __restrict__ keys_in,
                                                        index_t num_in,
                                                        typename Core::out_type* __restrict__ values_out, Core core,
                                                        index_t probing_length,
                                                        typename StatusHandler::base_type* __restrict__ status_out) {
    using K = typename Core::key_type;
    using O = typename Core::out_type;
    using Slot = typename Core::probe_slot_type;
    using Cursor = detail::WindowCursor<typename Core::probing_scheme_type>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    constexpr std::uint64_t full = tile<CG>::full_mask();
    const unsigned rank = lane_id() & (CG - 1);
    const unsigned shift = lane_id() & ~(CG - 1);
    auto gballot = [&](bool p) -> std::uint64_t {
        const std::uint64_t m = __builtin_amdgcn_ballot_w64(p);
        if constexpr (CG == 64) return m;
        else return (m >> shift) & full;
    };
    const tile<CG> group;
    detail::DeferredJoin join;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    const probing_schemes::Geometry& geo = core.geometry();
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, geo.windows);
    const Slot* __restrict__ slots = core.slot_ptr();

    detail::KeyBuffer<Core, CG, false> buf;
    buf.base = gid * CG;
    buf.stride = groups * CG;
    buf.taken = 0;
    buf.key = keys_in[buf.base + rank < num_in ? buf.base + rank : num_in - 1];
    buf.reload(keys_in, nullptr, num_in, rank);
    buf.prepare(core, num_in, rank);

    index_t idx[U];
    K key[U];
    Cursor cur[U];
    std::uint32_t phase[U];
    Status st[U];
    O out[U];

    // hand the next buffered key to slot u (or finish it at once)
    auto take = [&](int u) {
        out[u] = O{};
        st[u] = Status::none();
        if (!buf.ensure(core, num_in, rank)) { phase[u] = detail::kWait; return; }
        const unsigned j = buf.taken++;
        idx[u] = buf.base + j;
        const std::uint32_t state = group.shfl(buf.state, j);
        key[u] = group.shfl(buf.key, j);
        cur[u].w = group.shfl(buf.cur.w, j);
        cur[u].step = group.shfl(buf.cur.step, j);
        cur[u].left = budget;
        if (state == detail::kKeyOk) {
            phase[u] = detail::kLoad;
        } else if (state == detail::kPastEnd) {
            phase[u] = detail::kDone;
        } else {
            st[u] = state == detail::kNotReady ? Status::not_initialized() : Status::invalid_key();
            if (state == detail::kKeyInvalid) join(st[u]);
            phase[u] = detail::kEmit;
        }
    };
#pragma unroll
    for (int u = 0; u < U; ++u) {
        cur[u].w = 0;
        take(u);
    }

    while (true) {
        bool busy = false;
#pragma unroll
        for (int u = 0; u < U; ++u) busy |= phase[u] != detail::kDone;
        if (!busy) break;
        // every slot loads (idle ones re-read their last window, an L2 hit)
        // and the next buffer is re-read: loads under a branch would make
        // the compiler wait for all outstanding loads before the next one
        Slot s[U];
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = detail::load_slot(slots + cur[u].w * CG + rank);
        buf.reload(keys_in, nullptr, num_in, rank);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (phase[u] == detail::kLoad) {
                const auto hit = gballot(Core::key_of(s[u]) == key[u]);
                if (hit) {
                    out[u] = core.found(group.shfl(Core::value_of(s[u]), unsigned(ffs64(hit) - 1)));
                    phase[u] = detail::kEmit;
                } else if (gballot(Core::is_empty_key(Core::key_of(s[u])))) {
                    st[u] = core.missing(out[u]);
                    join(st[u]);
                    phase[u] = detail::kEmit;
                } else if (!cur[u].advance(geo)) {
                    st[u] = Status::probing_length_exceeded();
                    join(st[u]);
                    phase[u] = detail::kEmit;
                }
            }
            while (phase[u] == detail::kEmit || phase[u] == detail::kWait) {
                if (phase[u] == detail::kEmit && rank == 0) {
                    if (!st[u].has_any()) values_out[idx[u]] = out[u];
                    StatusHandler::handle(st[u], status_out, idx[u]);
                }
                take(u);
                if (phase[u] == detail::kWait) break;
            }
        }
    }
    join.flush(core.status_word());
}

// Insert with U keys per group in flight. `Core` additionally provides the
// part of a slot insert reads (insert_probe_type, load_insert_probe,
// probe_key), the claiming CAS split into cas_issue() / cas_outcome(),
// publishes_value, and finish_insert() (value store or counter update once
// the key's slot is known).
template<class Core, class StatusHandler, int U, bool HasValues>
__global__ __launch_bounds__(256) void insert_batched(const typename Core::key_type* __restrict__ keys_in,
                                                      const typename Core::value_type* __restrict__ values_in,
                                                      index_t num_in, Core core, index_t probing_length,
                                                      typename StatusHandler::base_type* __restrict__ status_out) {
    using K = typename Core::key_type;
    using V = typename Core::value_type;
    using Probe = typename Core::insert_probe_type;
    using Cursor = detail::WindowCursor<typename Core::probing_scheme_type>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    constexpr std::uint64_t full = tile<CG>::full_mask();
    const unsigned rank = lane_id() & (CG - 1);
    const unsigned shift = lane_id() & ~(CG - 1);
    auto gballot = [&](bool p) -> std::uint64_t {
        const std::uint64_t m = __builtin_amdgcn_ballot_w64(p);
        if constexpr (CG == 64) return m;
        else return (m >> shift) & full;
    };
    const tile<CG> group;
    detail::DeferredJoin join;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    const probing_schemes::Geometry& geo = core.geometry();
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, geo.windows);
    auto* __restrict__ slots = core.slot_ptr();

    detail::KeyBuffer<Core, CG, HasValues> buf;
    buf.base = gid * CG;
    buf.stride = groups * CG;
    buf.taken = 0;
    {
        const index_t at = buf.base + rank < num_in ? buf.base + rank : num_in - 1;
        buf.key = keys_in[at];
        if constexpr (HasValues) buf.val = values_in[at];
        else { buf.val = V{}; buf.next_val = V{}; }
    }
    buf.reload(keys_in, values_in, num_in, rank);
    buf.prepare(core, num_in, rank);

    index_t idx[U];
    K key[U];
    V val[U];
    Cursor cur[U];
    std::uint32_t phase[U];
    std::uint64_t empty[U];   // EMPTY lanes of the current window not tried yet
    Probe seen[U];            // what this lane read of its slot in the current window
    index_t where[U];         // slot of the key once known
    Status st[U];
    bool published[U];

    auto take = [&](int u) {
        st[u] = Status::none();
        empty[u] = 0;
        where[u] = ~index_t(0);
        published[u] = false;
        if (!buf.ensure(core, num_in, rank)) { phase[u] = detail::kWait; return; }
        const unsigned j = buf.taken++;
        idx[u] = buf.base + j;
        const std::uint32_t state = group.shfl(buf.state, j);
        key[u] = group.shfl(buf.key, j);
        if constexpr (HasValues) val[u] = group.shfl(buf.val, j);
        else val[u] = V{};
        cur[u].w = group.shfl(buf.cur.w, j);
        cur[u].step = group.shfl(buf.cur.step, j);
        cur[u].left = budget;
        if (state == detail::kKeyOk) {
            phase[u] = detail::kLoad;
        } else if (state == detail::kPastEnd) {
            phase[u] = detail::kDone;
        } else {
            st[u] = state == detail::kNotReady ? Status::not_initialized() : Status::invalid_key();
            if (state == detail::kKeyInvalid) join(st[u]);
            phase[u] = detail::kEmit;
        }
    };
#pragma unroll
    for (int u = 0; u < U; ++u) {
        cur[u].w = 0;
        seen[u] = Probe{};
        take(u);
    }

    while (true) {
        bool busy = false;
#pragma unroll
        for (int u = 0; u < U; ++u) busy |= phase[u] != detail::kDone;
        if (!busy) break;
        Probe fresh[U];
        typename Core::cas_word old[U];
        bool lead[U];
        // unconditional window loads and next-buffer reads (see
        // retrieve_batched), then the claiming CASes; no result is read
        // before all of them are issued
#pragma unroll
        for (int u = 0; u < U; ++u) fresh[u] = Core::load_insert_probe(slots + cur[u].w * CG + rank);
        buf.reload(keys_in, values_in, num_in, rank);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            lead[u] = phase[u] == detail::kCas && rank == unsigned(ffs64(empty[u]) - 1);
            old[u] = typename Core::cas_word{};
            if (lead[u]) old[u] = core.cas_issue(slots + cur[u].w * CG + rank, seen[u], key[u], val[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (phase[u] == detail::kLoad) {
                seen[u] = fresh[u];
                const K tk = Core::probe_key(seen[u]);
                const auto hit = gballot(tk == key[u]);
                if (hit) {
                    st[u] = Status::duplicate_key();
                    join(st[u]);
                    where[u] = cur[u].w * CG + unsigned(ffs64(hit) - 1);
                    phase[u] = detail::kEmit;
                } else {
                    empty[u] = gballot(Core::is_empty_key(tk));
                    if (empty[u]) {
                        phase[u] = detail::kCas;   // its CAS goes out next round
                    } else if (!cur[u].advance(geo)) {
                        st[u] = Status::probing_length_exceeded();
                        join(st[u]);
                        phase[u] = detail::kEmit;
                    }
                }
            } else if (phase[u] == detail::kCas) {
                const unsigned leader = unsigned(ffs64(empty[u]) - 1);
                const int outcome = lead[u] ? core.cas_outcome(old[u], seen[u], key[u]) : -1;
                if (gballot(outcome == 2)) {
                    st[u] = Status::duplicate_key();
                    join(st[u]);
                    where[u] = cur[u].w * CG + leader;
                    phase[u] = detail::kEmit;
                } else if (gballot(outcome == 1)) {
                    where[u] = cur[u].w * CG + leader;
                    published[u] = Core::publishes_value;
                    phase[u] = detail::kEmit;
                } else if (!gballot(outcome == 3)) {
                    empty[u] ^= std::uint64_t(1) << leader;
                    if (!empty[u]) {
                        if (cur[u].advance(geo)) {
                            phase[u] = detail::kLoad;
                        } else {
                            st[u] = Status::probing_length_exceeded();
                            join(st[u]);
                            phase[u] = detail::kEmit;
                        }
                    }
                }
            }
            while (phase[u] == detail::kEmit || phase[u] == detail::kWait) {
                if (phase[u] == detail::kEmit && rank == 0) {
                    if (where[u] != ~index_t(0)) core.finish_insert(where[u], val[u], published[u], st[u], join);
                    StatusHandler::handle(st[u], status_out, idx[u]);
                }
                take(u);
                if (phase[u] == detail::kWait) break;
            }
        }
    }
    join.flush(core.status_word());
}

}  // namespace kernels
}  // namespace warpcore
