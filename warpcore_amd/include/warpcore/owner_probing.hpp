// Lane-owner probing: in a group of CG lanes (CG = slots per probe window,
// one 128-byte line for 8- and 16-byte slots) every lane owns one key. Each
// round the group reads the windows of its CG keys one after another, all
// lanes of the group reading one window per load instruction (one address
// translation per line), so a wavefront has 64 keys in flight. A key's
// state (key, value, window, step, phase) lives in its owner lane only:
// per round and window the wavefront spends one LDS broadcast read (the
// owner's key and window, published in LDS), one load, two compares with
// 64-lane ballots and a bit-field extract; the owner then decides alone,
// and claims a slot with its own CAS. Free lanes take the wave's next key
// (prefetched one key ahead).
//
// The windows of a key are read in the order of the one-key protocols
// (single_value_hash_table.cuh:742-823 for insert, :267-307 for retrieve):
// a hit in a window is a duplicate (insert: the value is overwritten) or
// the key's value (retrieve); an insert claims the lowest EMPTY lane of the
// first window that has one and tries that window's other EMPTY lanes in
// order when a CAS loses; a lookup ends as missing at a window with an
// EMPTY lane; the window budget ends a key with probing_length_exceeded.
// Results and statuses are those of the other kernel families; keys of one
// batch interleave as keys of different groups do in the reference.
//
// SPL (slots per lane) 2, for 8-byte packed slots: a group of CG/2 lanes
// reads a window of CG slots, each lane its two adjacent slots with one
// 16-byte load. Window bit b of the hit/EMPTY masks is slot 2*b of the
// window for b < CG/2 and slot 2*(b - CG/2) + 1 above, so an insert tries
// the window's even slots before its odd ones: the lane order inside a
// window is slot placement only, which results and statuses do not see.
#pragma once

#include "batched_probing.hpp"
#include "engine.hpp"
#include "tile.hpp"

namespace warpcore {
namespace detail {

// What an owner publishes for its group each round: its key, its current
// window (kNoWindow: not reading a window this round) and (lookups) its
// position in the range.
template<class K, class W>
struct alignas(16) OwnerEntry {
    K key;
    W w;
    std::uint32_t pos;
};

template<class W> constexpr W kNoWindow = W(~W(0));

WC_D unsigned lanes_below(std::uint64_t m) noexcept {
    return __builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u));
}

// Bits of the caller's group in the 64-lane mask m.
template<unsigned CG>
WC_D std::uint64_t group_field(std::uint64_t m, unsigned shift) noexcept {
    if constexpr (CG == 64) return m;
    else return (m >> shift) & ((std::uint64_t(1) << CG) - 1);
}

// Cores that want each key's slot (found or claimed) by input position
// provide on_slot(slot, position, claimed): claimed when this key's own
// compare-and-swap took the slot.
template<class C, class = void>
struct has_slot_hook : std::false_type {};
template<class C>
struct has_slot_hook<C, std::void_t<decltype(std::declval<const C&>().on_slot(index_t(0), index_t(0), false))>>
    : std::true_type {};
template<class Core>
WC_D void record_slot(const Core& core, index_t slot, index_t pos, bool claimed) noexcept {
    if constexpr (has_slot_hook<Core>::value) core.on_slot(slot, pos, claimed);
}

// Slot in its window of window-mask bit b (groups of G lanes, SPL slots each).
template<unsigned G, unsigned SPL>
WC_D unsigned owner_slot_of_bit(unsigned b) noexcept {
    if constexpr (SPL == 1) return b;
    else return (b & (G - 1)) * 2 + b / G;
}

// The SPL slots of one lane (SPL 2: one 16-byte load of two 8-byte slots).
template<unsigned SPL, class T>
WC_D void owner_load_pair(const T* p, T (&out)[SPL]) noexcept {
    if constexpr (SPL == 1) {
        out[0] = load_slot(p);
    } else {
        static_assert(SPL == 2 && sizeof(T) == 8, "two slots per lane: 8-byte slots");
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        __builtin_memcpy(&out[0], &v, 8);
        __builtin_memcpy(&out[1], reinterpret_cast<const char*>(&v) + 8, 8);
    }
}

// Inserts of keys_in[first .. end) by the calling wavefront (all 64 lanes
// call this with the same range). `lds` holds 64 entries of this wave.
// Single-value cores only (insert_refill_range covers multi-value cores).
template<class Core, class StatusHandler, class W, bool HasValues, unsigned SPL = 1>
WC_D void insert_owner_range(const typename Core::key_type* __restrict__ keys_in,
                             const typename Core::value_type* __restrict__ values_in, index_t first, index_t end,
                             const Core& core, index_t budget,
                             typename StatusHandler::base_type* __restrict__ status_out,
                             OwnerEntry<typename Core::key_type, W>* lds, DeferredJoin& join) {
    using K = typename Core::key_type;
    using V = typename Core::value_type;
    using Probe = typename Core::insert_probe_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    using Entry = OwnerEntry<K, W>;
    using Bits = typename std::conditional<(Core::cg_size() <= 32), std::uint32_t, std::uint64_t>::type;
    constexpr unsigned WIN = unsigned(Core::cg_size());   // slots per window
    constexpr unsigned CG = WIN / SPL;                     // lanes per group
    static_assert(CG >= 2 && CG <= 64 && (CG & (CG - 1)) == 0, "groups of 2..64 lanes");
    constexpr bool multi = is_multi_value_core<Core>::value;
    static_assert(SPL == 1 || (SPL == 2 && !multi && sizeof(Probe) == 8 && !std::is_same<Probe, K>::value),
                  "two slots per lane: packed 8-byte slots of single-value cores");
    constexpr bool keep_seen = !std::is_same<Probe, K>::value;
    constexpr unsigned kFree = 0, kLoad = 1, kClaim = 2;
    const unsigned lane = lane_id();
    const unsigned rank = lane & (CG - 1);
    const unsigned shift = lane & ~(CG - 1);
    auto ballot = [](bool p) -> std::uint64_t { return __builtin_amdgcn_ballot_w64(p); };

    const probing_schemes::Geometry& geo = core.geometry();
    const W p = W(geo.windows);
    auto* __restrict__ slots = core.slot_ptr();
    if (end <= first) return;
    if (!core.ready()) {
        for (index_t i = first + lane; i < end; i += 64)
            StatusHandler::handle(Status::not_initialized(), status_out, i);
        return;
    }

    K key = K{};
    V val = V{};
    W w = 0, step = 1;
    std::uint32_t rnd = 0;
    index_t pos = 0;
    unsigned phase = kFree;
    Bits emp = 0;
    Probe seen = Core::empty_probe();
    index_t count = 0;      // multi-value: values of the key met so far
    index_t new_keys = 0;   // multi-value: keys this lane stored a first value of
    // the lane's next key, loaded one key ahead
    index_t cursor = first;
    K nk = K{};
    V nv = V{};
    index_t npos = cursor + lane;
    bool nhas = npos < end;
    if (nhas) {
        nk = keys_in[npos];
        if constexpr (HasValues) nv = values_in[npos];
    }
    cursor += 64;

    auto advance = [&]() {   // to the key's next window, or out of windows
        mark_overflow(core, w, step);
        if (rnd + 1 >= budget) {
            Status st = Status::probing_length_exceeded();
            if constexpr (multi)
                if (count > 0) st = st + Status::duplicate_key();
            join(st);
            StatusHandler::handle(st, status_out, pos);
            phase = kFree;
        } else {
            const W nw = w + step;
            w = nw >= p ? W(nw - p) : nw;
            rnd += 1;
            phase = kLoad;
        }
    };

    while (true) {
        // a free lane takes its prefetched key and prefetches the wave's next
        const bool take = phase == kFree && nhas;
        if (take) {
            key = nk;
            val = nv;
            pos = npos;
            if (core.is_valid_key(key)) {
                Cursor c;
                c.begin(key, core.hash_seed(), geo);
                w = W(c.w);
                step = W(c.step);
                rnd = 0;
                count = 0;
                phase = kLoad;
            } else {
                join(Status::invalid_key());
                StatusHandler::handle(Status::invalid_key(), status_out, pos);
            }
        }
        const std::uint64_t taken = ballot(take);
        if (taken != 0) {
            if (take) {
                npos = cursor + lanes_below(taken);
                nhas = npos < end;
                if (nhas) {
                    nk = keys_in[npos];
                    if constexpr (HasValues) nv = values_in[npos];
                }
            }
            cursor += index_t(__builtin_popcountll(taken));
        }
        if (ballot(phase != kFree || nhas) == 0) break;

        // publish key and window; every group reads its owners' windows
        Entry mine;
        mine.key = key;
        mine.w = phase == kLoad ? w : kNoWindow<W>;
        mine.pos = 0;
        lds[lane] = mine;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const unsigned ph = phase;   // this round's phase (the claim outcome below may change it)
        // claiming owners: CAS on the lowest untried EMPTY slot, issued ahead
        // of the window loads, so that its outcome (and the value store that
        // follows a claim) waits for the CAS alone
        const unsigned lead = owner_slot_of_bit<CG, SPL>(unsigned(ffs64(std::uint64_t(emp)) - 1));
        typename Core::cas_word old{};
        if (ph == kClaim) old = core.cas_issue(slots + index_t(w) * WIN + lead, seen, key, val);
        // only the loaded words stay in registers while the loads are in
        // flight; the owners' entries are read again from LDS to compare
        Probe probe[CG][SPL];
#pragma unroll
        for (unsigned j = 0; j < CG; ++j) {
            const W ew = lds[shift + j].w;
            const W at = ew != kNoWindow<W> ? ew : W(0);
            if constexpr (SPL == 1) probe[j][0] = Core::load_insert_probe(slots + index_t(at) * WIN + rank);
            else owner_load_pair<SPL>(slots + index_t(at) * WIN + rank * SPL, probe[j]);
        }
        if constexpr (multi) {
            if (ph == kClaim) {
                // multi_value_hash_table.cuh:200-289 per key: a claim stores one
                // more value; a slot that took this key meanwhile counts as one
                // of its values
                const int code = core.cas_outcome(old, seen, key);
                if (code == 1) {
                    const Status st = count > 0 ? Status::duplicate_key() : Status::none();
                    Status fin = st;
                    core.finish_insert(index_t(w) * WIN + lead, val, Core::publishes_value, fin, join);
                    if (count == 0) ++new_keys;
                    StatusHandler::handle(st, status_out, pos);
                    phase = kFree;
                } else if (code == 0 || code == 2) {
                    if (code == 2) count += 1;
                    if (count >= index_t(core.max_values)) {
                        const Status full = Status::duplicate_key() + Status::max_values_for_key_reached();
                        join(full);
                        StatusHandler::handle(full, status_out, pos);
                        phase = kFree;
                    } else {
                        emp &= Bits(emp - 1);
                        if (emp == 0) advance();
                    }
                }
            }
        } else {
            if (ph == kClaim) {
                const int code = core.cas_outcome(old, seen, key);
                if (code == 1 || code == 2) {   // claimed, or the key was already there
                    Status st = code == 1 ? Status::none() : Status::duplicate_key();
                    core.finish_insert(index_t(w) * WIN + lead, val, code == 1 && Core::publishes_value, st, join);
                    record_slot(core, index_t(w) * WIN + lead, pos, code == 1);
                    if (code == 2) join(Status::duplicate_key());
                    StatusHandler::handle(st, status_out, pos);
                    phase = kFree;
                } else if (code == 0) {   // another key took the slot: the window's next EMPTY slot
                    emp &= Bits(emp - 1);
                    if (emp == 0) advance();
                }                         // code 3: the same slot again, with its refreshed word
            }
        }
        __builtin_amdgcn_wave_barrier();

        // each window's hits and EMPTY slots, extracted by its owner
        Bits hit = 0, empty = 0;
#pragma unroll
        for (unsigned j = 0; j < CG; ++j) {
            const Entry e = lds[shift + j];
            const bool on = e.w != kNoWindow<W>;
#pragma unroll
            for (unsigned h = 0; h < SPL; ++h) {
                const K tk = Core::probe_key(probe[j][h]);
                const std::uint64_t hm = ballot(on && tk == e.key);
                const std::uint64_t em = ballot(on && Core::is_empty_key(tk));
                if (rank == j) {
                    hit |= Bits(group_field<CG>(hm, shift) << (h * CG));
                    empty |= Bits(group_field<CG>(em, shift) << (h * CG));
                }
            }
        }

        if constexpr (multi) {
            if (ph == kLoad) {
                count += index_t(__builtin_popcountll(std::uint64_t(hit)));
                if (count >= index_t(core.max_values)) {
                    const Status full = Status::duplicate_key() + Status::max_values_for_key_reached();
                    join(full);
                    StatusHandler::handle(full, status_out, pos);
                    phase = kFree;
                } else if (empty != 0) {
                    emp = empty;
                    if constexpr (keep_seen) seen = Core::empty_probe();
                    phase = kClaim;
                } else {
                    advance();
                }
            }
        } else {
            if (ph == kLoad) {
                if (hit != 0) {   // duplicate: the value is overwritten in place
                    Status st = Status::duplicate_key();
                    const index_t at = index_t(w) * WIN + owner_slot_of_bit<CG, SPL>(unsigned(ffs64(std::uint64_t(hit)) - 1));
                    core.finish_insert(at, val, false, st, join);
                    record_slot(core, at, pos, false);
                    join(Status::duplicate_key());
                    StatusHandler::handle(st, status_out, pos);
                    phase = kFree;
                } else if (empty != 0) {
                    emp = empty;
                    if constexpr (keep_seen) seen = Core::empty_probe();
                    phase = kClaim;
                } else {
                    advance();
                }
            }
        }
    }
    if constexpr (multi) core.add_keys(new_keys);
}

// Lookups of keys_in[first .. end) by the calling wavefront. A hit is
// written by the lane holding the slot; the owner writes the status.
//
// Erase (single_value_hash_table.cuh:425-468 per key): the same walk; at a
// hit the lane holding the slot writes TOMBSTONE over the key and the key
// reports none; key_not_found and probing_length_exceeded are reported but
// not joined into the table status (invalid keys are); values_out is not
// used.
//
// Overflow filter (cores with one, see mark_overflow): a key whose window
// is full and holds no copy of it reads the window's filter byte (next
// round, from the MALL-sized filter instead of another window) and ends as
// missing when its bit is clear: no key with that bit ever moved past the
// window, so the key is in no later window. This gives the status the full
// walk would: it is used only when the window budget covers the whole
// cycle of windows and the wavefront has seen a window with an EMPTY slot
// (the table is static during the call), where the walk would end at an
// EMPTY slot with key_not_found.
template<class Core, class StatusHandler, class W, unsigned SPL = 1, bool Erase = false>
WC_D void retrieve_owner_range(const typename Core::key_type* __restrict__ keys_in, index_t first, index_t end,
                               typename Core::out_type* __restrict__ values_out, const Core& core, index_t budget,
                               typename StatusHandler::base_type* __restrict__ status_out,
                               OwnerEntry<typename Core::key_type, W>* lds, DeferredJoin& join) {
    using K = typename Core::key_type;
    using O = typename Core::out_type;
    using Slot = typename Core::probe_slot_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    using Entry = OwnerEntry<K, W>;
    constexpr unsigned WIN = unsigned(Core::cg_size());   // slots per window
    constexpr unsigned CG = WIN / SPL;                     // lanes per group
    static_assert(CG >= 2 && CG <= 64 && (CG & (CG - 1)) == 0, "groups of 2..64 lanes");
    static_assert(!is_multi_value_core<Core>::value, "single-value cores only");
    static_assert(SPL == 1 || (SPL == 2 && sizeof(Slot) == 8), "two slots per lane: 8-byte slots");
    constexpr unsigned kFree = 0, kLoad = 1, kFilter = 2;
    const unsigned lane = lane_id();
    const unsigned rank = lane & (CG - 1);
    const unsigned shift = lane & ~(CG - 1);
    auto ballot = [](bool p) -> std::uint64_t { return __builtin_amdgcn_ballot_w64(p); };

    const probing_schemes::Geometry& geo = core.geometry();
    const W p = W(geo.windows);
    const Slot* __restrict__ lane_slots = core.slot_ptr() + rank * SPL;
    const std::uint32_t* filter = nullptr;
    if constexpr (has_overflow_filter<Core>::value)
        if (budget >= index_t(geo.windows)) filter = core.filter;
    bool seen_empty = false;   // some window of the table has an EMPTY slot
    if (end <= first) return;
    if (!core.ready()) {
        for (index_t i = first + lane; i < end; i += 64)
            StatusHandler::handle(Status::not_initialized(), status_out, i);
        return;
    }
    O miss_value{};
    const Status miss_status = Erase ? Status::key_not_found() : core.missing(miss_value);

    K key = K{};
    W w = 0, step = 1;
    std::uint32_t rnd = 0;
    index_t pos = 0;
    unsigned phase = kFree;
    index_t cursor = first;
    K nk = K{};
    index_t npos = cursor + lane;
    bool nhas = npos < end;
    if (nhas) nk = keys_in[npos];
    cursor += 64;

    while (true) {
        const bool take = phase == kFree && nhas;
        if (take) {
            key = nk;
            pos = npos;
            if (core.is_valid_key(key)) {
                Cursor c;
                c.begin(key, core.hash_seed(), geo);
                w = W(c.w);
                step = W(c.step);
                rnd = 0;
                phase = kLoad;
            } else {
                join(Status::invalid_key());
                StatusHandler::handle(Status::invalid_key(), status_out, pos);
            }
        }
        const std::uint64_t taken = ballot(take);
        if (taken != 0) {
            if (take) {
                npos = cursor + lanes_below(taken);
                nhas = npos < end;
                if (nhas) nk = keys_in[npos];
            }
            cursor += index_t(__builtin_popcountll(taken));
        }
        if (ballot(phase != kFree || nhas) == 0) break;

        Entry mine;
        mine.key = key;
        mine.w = phase == kLoad ? w : kNoWindow<W>;
        mine.pos = std::uint32_t(pos - first);
        lds[lane] = mine;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // only the loaded slots stay in registers while the loads are in
        // flight; the owners' entries are read again from LDS to compare
        Slot s[CG][SPL];
#pragma unroll
        for (unsigned j = 0; j < CG; ++j) {
            const W ew = lds[shift + j].w;
            if (ew != kNoWindow<W>) {   // no load for an owner without a window this round
                owner_load_pair<SPL>(lane_slots + index_t(ew) * WIN, s[j]);
            } else {
#pragma unroll
                for (unsigned x = 0; x < SPL; ++x) s[j][x] = Slot{};
            }
        }
        std::uint32_t fword = 0;
        if (phase == kFilter) fword = filter[w >> 2];
        __builtin_amdgcn_wave_barrier();
        bool hit = false, empty = false;
        std::uint64_t any_empty = 0;
#pragma unroll
        for (unsigned j = 0; j < CG; ++j) {
            const Entry e = lds[shift + j];
            const bool on = e.w != kNoWindow<W>;
#pragma unroll
            for (unsigned x = 0; x < SPL; ++x) {
                const K tk = Core::key_of(s[j][x]);
                const bool h = on && tk == e.key;
                if (h) {   // the slot's lane writes the hit
                    if constexpr (Erase) core.erase_slot(index_t(e.w) * WIN + rank * SPL + x);
                    else values_out[first + e.pos] = core.found(Core::value_of(s[j][x]));
                }
                const std::uint64_t hm = ballot(h);
                const std::uint64_t em = ballot(on && Core::is_empty_key(tk));
                any_empty |= em;
                if (rank == j) {
                    hit = hit || group_field<CG>(hm, shift) != 0;
                    empty = empty || group_field<CG>(em, shift) != 0;
                }
            }
        }
        seen_empty = seen_empty || any_empty != 0;
        auto miss = [&]() {
            if constexpr (Erase) {
                StatusHandler::handle(Status::key_not_found(), status_out, pos);
            } else {
                join(miss_status);
                StatusHandler::handle(miss_status, status_out, pos);
                if (!miss_status.has_any()) values_out[pos] = miss_value;
            }
            phase = kFree;
        };
        auto next_window = [&]() {
            const W nw = w + step;
            w = nw >= p ? W(nw - p) : nw;
            rnd += 1;
            phase = kLoad;
        };
        if (phase == kLoad) {
            if (hit) {
                StatusHandler::handle(Status::none(), status_out, pos);
                phase = kFree;
            } else if (empty) {
                miss();
            } else if (rnd + 1 >= budget) {
                if constexpr (!Erase) join(Status::probing_length_exceeded());
                StatusHandler::handle(Status::probing_length_exceeded(), status_out, pos);
                phase = kFree;
            } else if (filter != nullptr && seen_empty) {
                phase = kFilter;   // read the window's filter byte next round
            } else {
                next_window();
            }
        } else if (phase == kFilter) {
            if ((fword & overflow_bit(w, step)) == 0) miss();
            else next_window();
        }
    }
}

}  // namespace detail

namespace kernels {

// Bulk inserts / lookups by lane owners: wavefront v of the grid takes keys
// [v * per, (v + 1) * per).
template<class Core, class StatusHandler, class W, bool HasValues, unsigned SPL = 1>
__global__ __launch_bounds__(256) void insert_owner(const typename Core::key_type* __restrict__ keys_in,
                                                    const typename Core::value_type* __restrict__ values_in,
                                                    index_t num_in, Core core, index_t probing_length,
                                                    typename StatusHandler::base_type* __restrict__ status_out,
                                                    const index_t* __restrict__ num_dev) {
    __shared__ detail::OwnerEntry<typename Core::key_type, W> lds[256];
    detail::DeferredJoin join;
    if (num_dev != nullptr && *num_dev < num_in) num_in = *num_dev;
    const index_t wave = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const index_t waves = (index_t(gridDim.x) * blockDim.x) / 64;
    const index_t per = SDIV(num_in, waves);
    const index_t first = wave * per < num_in ? wave * per : num_in;
    const index_t end = first + per < num_in ? first + per : num_in;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, Core::cg_size(),
                                                                  core.geometry().windows);
    detail::insert_owner_range<Core, StatusHandler, W, HasValues, SPL>(keys_in, values_in, first, end, core, budget,
                                                                  status_out, lds + (threadIdx.x & ~63u), join);
    join.flush(core.status_word());
}

template<class Core, class StatusHandler, class W, unsigned SPL = 1>
__global__ __launch_bounds__(256) void retrieve_owner(const typename Core::key_type* __restrict__ keys_in,
                                                      index_t num_in, typename Core::out_type* __restrict__ values_out,
                                                      Core core, index_t probing_length,
                                                      typename StatusHandler::base_type* __restrict__ status_out) {
    __shared__ detail::OwnerEntry<typename Core::key_type, W> lds[256];
    detail::DeferredJoin join;
    const index_t wave = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const index_t waves = (index_t(gridDim.x) * blockDim.x) / 64;
    const index_t per = SDIV(num_in, waves);
    const index_t first = wave * per < num_in ? wave * per : num_in;
    const index_t end = first + per < num_in ? first + per : num_in;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, Core::cg_size(),
                                                                  core.geometry().windows);
    detail::retrieve_owner_range<Core, StatusHandler, W, SPL>(keys_in, first, end, values_out, core, budget, status_out,
                                                         lds + (threadIdx.x & ~63u), join);
    join.flush(core.status_word());
}

template<class Core, class StatusHandler, class W, unsigned SPL = 1>
__global__ __launch_bounds__(256) void erase_owner(const typename Core::key_type* __restrict__ keys_in,
                                                   index_t num_in, Core core, index_t probing_length,
                                                   typename StatusHandler::base_type* __restrict__ status_out) {
    __shared__ detail::OwnerEntry<typename Core::key_type, W> lds[256];
    detail::DeferredJoin join;
    const index_t wave = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const index_t waves = (index_t(gridDim.x) * blockDim.x) / 64;
    const index_t per = SDIV(num_in, waves);
    const index_t first = wave * per < num_in ? wave * per : num_in;
    const index_t end = first + per < num_in ? first + per : num_in;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, Core::cg_size(),
                                                                  core.geometry().windows);
    detail::retrieve_owner_range<Core, StatusHandler, W, SPL, true>(keys_in, first, end, nullptr, core, budget,
                                                                  status_out, lds + (threadIdx.x & ~63u), join);
    join.flush(core.status_word());
}

}  // namespace kernels
}  // namespace warpcore
