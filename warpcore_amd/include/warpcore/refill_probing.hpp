// Refill probing for lookups: a group of CG lanes keeps U keys in flight and
// walks their windows one round at a time, like the lockstep kernel, but a
// key that finishes is replaced by the group's next key at the start of the
// following round, so one long probe sequence (a miss at load factor 0.9
// can walk a dozen windows) no longer idles the other slots of its group or
// the other groups of its wavefront.
//
// What is live is kept as one 64-bit lane mask per slot (bit l: the group of
// lane l has a key in slot u), which the compiler holds in scalar registers:
// the decisions of a round are scalar mask arithmetic, a lane's bit is read
// with inverse_ballot (the mask used as the condition of a select), and the
// "any lane of my group" of a ballot is one AND with the lane's group mask
// and one 64-bit compare. Per slot and round a wavefront spends two key
// compares, two such reductions, the window address and the window advance.
//
// Keys enter through a per-group LDS stage of CG entries: lane j of the group
// hashes key `next + j` of the group's range (first window and step), and a
// slot being refilled reads the lowest untaken entry (an LDS broadcast).
// Invalid keys are reported at staging and never take a slot.
//
// Per key the windows are read in the same order as the one-key protocol
// (single_value_hash_table.cuh:267-307): a hit ends the key with its value,
// a window with an EMPTY lane and no hit ends it as missing, the window
// budget ends it with probing_length_exceeded. Results and statuses are
// those of the other kernel families.
#pragma once

#include "batched_probing.hpp"
#include "engine.hpp"
#include "owner_probing.hpp"
#include "tile.hpp"

namespace warpcore {
namespace detail {

// Slots per group of the refill kernels (pipelined lookups: their own
// count, since every slot's window stays in registers for a whole pass).
#ifndef WARPCORE_REFILL_PIPE_KEYS
#define WARPCORE_REFILL_PIPE_KEYS 8
#endif
#ifndef WARPCORE_REFILL_KEYS
#define WARPCORE_REFILL_KEYS 8
#endif
template<unsigned CG> constexpr int refill_keys() noexcept {
    return CG < 2 ? 2 : WARPCORE_REFILL_KEYS;   // slots are not tied to lanes
}
#ifndef WARPCORE_REFILL_INSERT_KEYS
#define WARPCORE_REFILL_INSERT_KEYS 4
#endif
template<unsigned CG> constexpr int refill_insert_keys() noexcept {
    return CG < WARPCORE_REFILL_INSERT_KEYS ? int(CG) : WARPCORE_REFILL_INSERT_KEYS;
}


// One staged key: the key, its first window and its window step.
template<class K, class W>
struct alignas(16) RefillEntry {
    K key;
    W w, step;
};

// Lookups of keys_in[first .. end) by the calling group (all lanes of the
// wavefront call this together; each group passes its own range). W: window
// index type (std::uint32_t below 2^31 windows). `stage` holds CG entries of
// this group.
//
// Multi-value cores (multi_value_hash_table.cuh:605-661 per key): when a
// slot takes key i it asks core.range_of(i, base, room) for the key's
// output range (read while the first window is in flight); a hit does not
// end the key, its lanes holding the key pass each value to
// core.emit(base, room, j, value) (the j-th value of the key in probe
// order); the key ends at a window with an EMPTY lane or once
// core.max_values values were met, with none, or key_not_found (joined)
// if it has no value, and core.finish(i, count, room, status, join,
// *tally) sees the key's value count.
template<class Core, class StatusHandler, int U, class W, bool Pipe = false>
WC_D void retrieve_refill_range(const typename Core::key_type* __restrict__ keys_in, index_t first, index_t end,
                                typename Core::out_type* __restrict__ values_out, const Core& core, index_t budget,
                                typename StatusHandler::base_type* __restrict__ status_out,
                                RefillEntry<typename Core::key_type, W>* stage, DeferredJoin& join,
                                index_t* tally = nullptr) {
    using K = typename Core::key_type;
    using O = typename Core::out_type;
    using Slot = typename Core::probe_slot_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    using Entry = RefillEntry<K, W>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    static_assert(CG >= 1 && CG <= 64 && (CG & (CG - 1)) == 0, "groups of 1..64 lanes");
    static_assert(U >= 1 && U <= 16, "slots per group");
    const unsigned lane = lane_id();
    const unsigned rank = lane & (CG - 1);
    const unsigned shift = lane & ~(CG - 1);
    constexpr std::uint64_t full = tile<CG>::full_mask();
    const std::uint64_t gm = CG == 64 ? ~std::uint64_t(0) : (full << shift);   // lanes of my group
    auto ballot = [](bool p) -> std::uint64_t { return __builtin_amdgcn_ballot_w64(p); };
    auto mine = [](std::uint64_t m) -> bool { return __builtin_amdgcn_inverse_ballot_w64(m); };
    // lanes of every group that has a lane in m
    auto spread = [&](std::uint64_t m) -> std::uint64_t { return ballot((m & gm) != 0); };
    auto group_bits = [&](std::uint64_t m) -> std::uint64_t { return CG == 64 ? m : ((m >> shift) & full); };
    constexpr bool multi = is_multi_value_core<Core>::value;

    const probing_schemes::Geometry& geo = core.geometry();
    const W p = W(geo.windows);
    const Slot* __restrict__ lane_slots = core.slot_ptr() + rank;
    const index_t len = end > first ? end - first : 0;
    O* __restrict__ out = values_out != nullptr ? values_out + first : nullptr;

    if (!core.ready()) {   // uniform: every key reports not_initialized
        for (index_t i = rank; i < len; i += CG) StatusHandler::handle(Status::not_initialized(), status_out, first + i);
        return;
    }
    O miss_value{};
    const Status miss_status = core.missing(miss_value);
    // groups that saw a miss / ran out of windows (scalar masks, joined per
    // group at the end: groups of one wavefront may probe different tables)
    std::uint64_t missed = 0, exhausted = 0;

    K key[U];
    W w[U], step[U];
    std::uint32_t pos[U], rnd[U];
    std::uint64_t live[U];
    Slot s[U];   // the window each slot read (pipelined: in flight from the previous pass)
    index_t count[multi ? U : 1];   // multi-value: values of the key met so far
    index_t base[multi ? U : 1], room[multi ? U : 1];   // multi-value: the key's output range
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if constexpr (multi) {
            count[u] = 0;
            base[u] = 0;
            room[u] = 0;
        }
        key[u] = K{};
        w[u] = 0;
        step[u] = 1;
        pos[u] = 0;
        rnd[u] = 0;
        live[u] = 0;
        s[u] = Slot{};
    }
    index_t next = 0;          // keys of the range staged so far (group-uniform)
    std::uint64_t avail = 0;   // stage entries holding a key not yet taken (group-uniform)
    std::uint32_t stage_pos = 0;
    // the keys of the next stage are loaded one stage ahead, so staging
    // never waits for memory
    K ahead = K{};
    if (len > 0) ahead = keys_in[first + (rank < len ? rank : 0)];
    if constexpr (Pipe) {   // every slot has a load in flight (window 0) from the start, issued
                            // after `ahead`, so the first staging need not wait for the slots
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = load_slot(lane_slots);
    }

    while (true) {
        // stage the next CG keys of the range once the stage is used up
        if (avail == 0 && next < len) {
            const bool in = next + rank < len;
            const index_t i = first + next + (in ? rank : 0);
            const K k = ahead;
            const bool valid = in && core.is_valid_key(k);
            if (valid) {
                Cursor c;
                c.begin(k, core.hash_seed(), geo);
                Entry e;
                e.key = k;
                e.w = W(c.w);
                e.step = W(c.step);
                stage[rank] = e;
            }
            if (in && !valid) {
                join(Status::invalid_key());
                StatusHandler::handle(Status::invalid_key(), status_out, i);
            }
            const std::uint64_t vb = ballot(valid);
            avail = CG == 64 ? vb : ((vb >> shift) & full);
            {   // the next stage's keys, loaded once this stage's are used (the
                // load then writes `ahead` in place) and unconditionally (the
                // last key again past the end): a load that had to be copied
                // into `ahead` would be waited for at once
                const index_t nn = next + CG + rank;
                ahead = keys_in[first + (nn < len ? nn : len - 1)];
            }
            stage_pos = std::uint32_t(next);
            next += CG;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // refill slot u from the stage, lowest entry first
        auto take_slot = [&](int u) {
            const bool take = !mine(live[u]) && avail != 0;
            if (take) {
                const unsigned j = unsigned(ffs64(avail) - 1);
                avail &= avail - 1;
                const Entry e = stage[j];
                key[u] = e.key;
                w[u] = e.w;
                step[u] = e.step;
                pos[u] = stage_pos + j;
                rnd[u] = 0;
                if constexpr (multi) {
                    count[u] = 0;
                    core.range_of(first + pos[u], base[u], room[u]);
                }
            }
            live[u] |= ballot(take);
        };
        // a free slot reads window 0: one line every free slot shares
        auto issue_slot = [&](int u) {
            const W at = mine(live[u]) ? w[u] : W(0);
            s[u] = load_slot(lane_slots + index_t(at) * CG);
        };
        // slot u's window has arrived: finish or advance its key
        auto process_slot = [&](int u) {
            const K tk = Core::key_of(s[u]);
            const std::uint64_t hit = ballot(tk == key[u]) & live[u];
            if constexpr (multi) {
                const std::uint64_t gh = group_bits(hit);
                if (mine(hit))
                    core.emit(base[u], room[u],
                              count[u] + index_t(__builtin_popcountll(gh & ((std::uint64_t(1) << rank) - 1))),
                              Core::value_of(s[u]));
                count[u] += index_t(__builtin_popcountll(gh));
                const std::uint64_t stop =
                    spread((ballot(Core::is_empty_key(tk)) | ballot(count[u] >= core.max_values)) & live[u]);
                const std::uint64_t out_of_windows = live[u] & ~stop & ballot(rnd[u] + 1 >= budget);
                const std::uint64_t missing = stop & ballot(count[u] == 0);
                missed |= missing;
                exhausted |= out_of_windows;
                if (rank == 0 && mine(stop | out_of_windows)) {
                    Status st = mine(out_of_windows) ? Status::probing_length_exceeded()
                                                     : (mine(missing) ? Status::key_not_found() : Status::none());
                    core.finish(first + pos[u], count[u], room[u], st, join, tally);
                    StatusHandler::handle(st, status_out, first + pos[u]);
                }
                const std::uint64_t cont = live[u] & ~stop & ~out_of_windows;
                const W nw = w[u] + step[u];
                w[u] = mine(cont) ? (nw >= p ? W(nw - p) : nw) : w[u];
                rnd[u] += 1;
                live[u] = cont;
            } else {
                const std::uint64_t stop = spread(hit | (ballot(Core::is_empty_key(tk)) & live[u]));
                const std::uint64_t ghit = spread(hit);
                const std::uint64_t out_of_windows = live[u] & ~stop & ballot(rnd[u] + 1 >= budget);
                const std::uint64_t missing = stop & ~ghit;
                if (mine(hit)) out[pos[u]] = core.found(Core::value_of(s[u]));
                missed |= missing;
                exhausted |= out_of_windows;
                if (rank == 0 && mine(stop | out_of_windows)) {
                    const Status st = mine(ghit) ? Status::none()
                                                 : (mine(missing) ? miss_status : Status::probing_length_exceeded());
                    StatusHandler::handle(st, status_out, first + pos[u]);
                    if (mine(missing) && !miss_status.has_any()) out[pos[u]] = miss_value;
                }
                const std::uint64_t cont = live[u] & ~stop & ~out_of_windows;
                const W nw = w[u] + step[u];
                w[u] = mine(cont) ? (nw >= p ? W(nw - p) : nw) : w[u];
                rnd[u] += 1;
                live[u] = cont;
            }
        };
        std::uint64_t any_live = 0;
        if constexpr (!Pipe) {
            // one round: refill, every slot's load, then every slot's result
#pragma unroll
            for (int u = 0; u < U; ++u) {
                take_slot(u);
                any_live |= live[u];
            }
            if (any_live == 0 && ballot(avail != 0 || next < len) == 0) break;
            // staging reads and LDS reads of this round are done before the
            // stage can be overwritten at the top of the next round
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < U; ++u) issue_slot(u);
#pragma unroll
            for (int u = 0; u < U; ++u) process_slot(u);
        } else {
            // pipelined: slot u's window (issued one pass earlier) is
            // consumed and slot u's next load issued before slot u + 1 is
            // looked at, so a pass waits only for the oldest load in flight
            // (loads return in issue order) while the other slots' loads
            // stay outstanding
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (live[u] != 0) process_slot(u);
                take_slot(u);
                any_live |= live[u];
                issue_slot(u);
            }
            __builtin_amdgcn_wave_barrier();
            if (any_live == 0 && ballot(avail != 0 || next < len) == 0) break;
        }
    }
    if (mine(missed)) join(miss_status);
    if (mine(exhausted)) join(Status::probing_length_exceeded());
}

// One staged key of an insert: the key, its first window, its step and the
// value (or increment) that goes with it.
template<class K, class W, class V>
struct alignas(16) RefillInsertEntry {
    K key;
    W w, step;
    V value;
};

// Inserts of keys_in[first .. end) by the calling group, with U keys in
// flight (single_value_hash_table.cuh:742-823 per key). A slot is in one of
// two phases: loading (read the key's current window) or claiming (its
// window had EMPTY lanes; the lowest untried one issues the CAS). Each
// round issues the loads of loading slots and the CASes of claiming slots,
// then moves every slot on; a finished slot writes its result (value store
// or counter update through finish_insert, then the status) and is
// refilled the next round. `Core` provides the hooks of insert_batched.
//
// Multi-value cores: a window's lanes holding the key add to the key's
// count instead of ending the insert, a claiming CAS that returns the key
// counts one more and moves to the next EMPTY lane, and the insert ends
// when a CAS claims a slot (none if no value of the key was met, else
// duplicate_key; not joined) or when the count reaches core.max_values
// (duplicate_key + max_values_for_key_reached, joined). A claim with no
// earlier value adds one to *new_keys.
template<class Core, class StatusHandler, int U, class W, bool HasValues>
WC_D void insert_refill_range(const typename Core::key_type* __restrict__ keys_in,
                              const typename Core::value_type* __restrict__ values_in, index_t first, index_t end,
                              const Core& core, index_t budget,
                              typename StatusHandler::base_type* __restrict__ status_out,
                              RefillInsertEntry<typename Core::key_type, W, typename Core::value_type>* stage,
                              DeferredJoin& join, index_t* new_keys = nullptr) {
    using K = typename Core::key_type;
    using V = typename Core::value_type;
    using Probe = typename Core::insert_probe_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    using Entry = RefillInsertEntry<K, W, V>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    static_assert(CG >= 1 && CG <= 64 && (CG & (CG - 1)) == 0, "groups of 1..64 lanes");
    static_assert(U >= 1 && U <= 16, "slots per group");
    const unsigned lane = lane_id();
    const unsigned rank = lane & (CG - 1);
    const unsigned shift = lane & ~(CG - 1);
    constexpr std::uint64_t full = tile<CG>::full_mask();
    const std::uint64_t gm = CG == 64 ? ~std::uint64_t(0) : (full << shift);
    auto ballot = [](bool p) -> std::uint64_t { return __builtin_amdgcn_ballot_w64(p); };
    auto mine = [](std::uint64_t m) -> bool { return __builtin_amdgcn_inverse_ballot_w64(m); };
    auto spread = [&](std::uint64_t m) -> std::uint64_t { return ballot((m & gm) != 0); };
    auto group_bits = [&](std::uint64_t m) -> std::uint64_t { return CG == 64 ? m : ((m >> shift) & full); };

    const probing_schemes::Geometry& geo = core.geometry();
    const W p = W(geo.windows);
    auto* __restrict__ slots = core.slot_ptr();
    const index_t len = end > first ? end - first : 0;

    if (!core.ready()) {
        for (index_t i = rank; i < len; i += CG) StatusHandler::handle(Status::not_initialized(), status_out, first + i);
        return;
    }

    K key[U];
    V val[U];
    W w[U], step[U];
    std::uint32_t pos[U], rnd[U];
    // When a slot reads as just its key (no packed value bits), a claiming
    // lane always saw EMPTY there, so the CAS's expected word is the EMPTY
    // probe and the window read need not be kept.
    constexpr bool keep_seen = !std::is_same<Probe, K>::value;
    using Bits = typename std::conditional<(CG <= 32), std::uint32_t, std::uint64_t>::type;
    Bits emp[U];            // group bits: EMPTY lanes of the current window not tried yet
    Probe seen[keep_seen ? U : 1];   // this lane's slot as read in the current window
    std::uint64_t live[U], claim[U];
    constexpr bool multi = is_multi_value_core<Core>::value;
    index_t count[multi ? U : 1];   // values of the key met so far (the same in every lane of the group)
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if constexpr (multi) count[u] = 0;
        key[u] = K{};
        val[u] = V{};
        w[u] = 0;
        step[u] = 1;
        pos[u] = 0;
        rnd[u] = 0;
        emp[u] = 0;
        if constexpr (keep_seen) seen[u] = Probe{};
        live[u] = 0;
        claim[u] = 0;
    }
    // requests of the slots in the current round
    Probe fresh[U];
    typename Core::cas_word old[U];
    bool lead[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        fresh[u] = Probe{};
        old[u] = typename Core::cas_word{};
        lead[u] = false;
    }
    std::uint64_t dups = 0, exhausted = 0;   // groups that saw a duplicate (multi-value: reached the
                                             // maximum) / ran out of windows
    std::uint64_t exhausted_dup = 0;         // multi-value: ran out of windows after values of the key
    index_t next = 0;
    std::uint64_t avail = 0;
    std::uint32_t stage_pos = 0;
    // keys and values of the next stage, loaded one stage ahead
    K ahead = K{};
    V ahead_value = V{};
    if (len > 0) {
        const index_t i0 = first + (rank < len ? rank : 0);
        ahead = keys_in[i0];
        if constexpr (HasValues) ahead_value = values_in[i0];
    }

    while (true) {
        if (avail == 0 && next < len) {
            const bool in = next + rank < len;
            const index_t i = first + next + (in ? rank : 0);
            const K k = ahead;
            const V kv = ahead_value;
            const bool valid = in && core.is_valid_key(k);
            if (valid) {
                Cursor c;
                c.begin(k, core.hash_seed(), geo);
                Entry e;
                e.key = k;
                e.w = W(c.w);
                e.step = W(c.step);
                e.value = kv;
                stage[rank] = e;
            }
            if (in && !valid) {
                join(Status::invalid_key());
                StatusHandler::handle(Status::invalid_key(), status_out, i);
            }
            avail = group_bits(ballot(valid));
            {   // the next stage's pairs, loaded in place once this stage's are
                // used, unconditionally (see retrieve_refill_range)
                const index_t nn = next + CG + rank;
                const index_t i2 = first + (nn < len ? nn : len - 1);
                ahead = keys_in[i2];
                if constexpr (HasValues) ahead_value = values_in[i2];
            }
            stage_pos = std::uint32_t(next);
            next += CG;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        auto take_slot = [&](int u) {
            const bool take = !mine(live[u]) && avail != 0;
            if (take) {
                const unsigned j = unsigned(ffs64(avail) - 1);
                avail &= avail - 1;
                const Entry e = stage[j];
                key[u] = e.key;
                w[u] = e.w;
                step[u] = e.step;
                val[u] = e.value;
                pos[u] = stage_pos + j;
                rnd[u] = 0;
                if constexpr (multi) count[u] = 0;
            }
            const std::uint64_t t = ballot(take);
            live[u] |= t;
            claim[u] &= ~t;
        };
        // issue: the window load of a loading slot (a slot not loading reads
        // window 0: one line it shares), the CAS of a claiming slot's leader
        auto issue_slot = [&](int u) {
            const W at = mine(live[u] & ~claim[u]) ? w[u] : W(0);
            fresh[u] = Core::load_insert_probe(slots + index_t(at) * CG + rank);
            lead[u] = mine(claim[u]) && rank == unsigned(ffs64(emp[u]) - 1);
            old[u] = typename Core::cas_word{};
            if (lead[u]) {
                Probe expect = Core::empty_probe();
                if constexpr (keep_seen) expect = seen[u];
                old[u] = core.cas_issue(slots + index_t(w[u]) * CG + rank, expect, key[u], val[u]);
            }
        };
        auto process_slot = [&](int u) {
            const std::uint64_t loading = live[u] & ~claim[u];
            const std::uint64_t claiming = live[u] & claim[u];
            // loading slots: a hit is a duplicate (multi-value: counted; the
            // maximum reached ends the insert), EMPTY lanes start the claim
            const K tk = Core::probe_key(fresh[u]);
            const std::uint64_t hitl = ballot(tk == key[u]) & loading;
            std::uint64_t ghit;
            if constexpr (multi) {
                count[u] += index_t(__builtin_popcountll(group_bits(hitl)));
                ghit = spread(ballot(count[u] >= core.max_values) & loading);
            } else {
                ghit = spread(hitl);
            }
            const std::uint64_t em = ballot(Core::is_empty_key(tk));
            const std::uint64_t gemp = spread(em & loading) & ~ghit;
            if constexpr (keep_seen)
                if (mine(loading)) seen[u] = fresh[u];
            // claiming slots: the leader's CAS outcome, which every lane of
            // the group reads from the leader (one LDS-crossbar permute)
            const unsigned leader = unsigned(ffs64(emp[u]) - 1);
            int outcome = -1;
            if (lead[u]) {
                if constexpr (keep_seen) {
                    outcome = core.cas_outcome(old[u], seen[u], key[u]);
                } else {
                    Probe expect = Core::empty_probe();
                    outcome = core.cas_outcome(old[u], expect, key[u]);
                }
            }
            const int code = __builtin_amdgcn_ds_bpermute(int((shift + leader) << 2), outcome);
            const std::uint64_t won = ballot(code == 1) & claiming;
            std::uint64_t dupc = ballot(code == 2) & claiming;
            if constexpr (multi) {   // the key won this slot first: one more value, then the next EMPTY lane
                if (mine(dupc)) count[u] += 1;
                dupc &= ballot(count[u] >= core.max_values);
            }
            std::uint64_t retry = 0;   // only slots whose window read carries value bits
            if constexpr (keep_seen) retry = ballot(code == 3) & claiming;
            if (mine(gemp)) emp[u] = Bits(group_bits(em));
            const std::uint64_t lost = claiming & ~won & ~dupc & ~retry;
            if (mine(lost)) emp[u] &= Bits(emp[u] - 1);
            const std::uint64_t drained = lost & ballot(emp[u] == 0);
            // finished keys: duplicate (found in the window or returned by
            // the CAS) or claimed
            const std::uint64_t dup = ghit | dupc;
            const std::uint64_t done = dup | won;
            dups |= dup;
            if (rank == 0 && mine(done)) {
                if constexpr (multi) {
                    Status st = mine(dup) ? Status::duplicate_key() + Status::max_values_for_key_reached()
                                          : (count[u] > 0 ? Status::duplicate_key() : Status::none());
                    if (mine(won)) {
                        core.finish_insert(index_t(w[u]) * CG + leader, val[u], Core::publishes_value, st, join);
                        if (count[u] == 0 && new_keys != nullptr) *new_keys += 1;
                    }
                    StatusHandler::handle(st, status_out, first + pos[u]);
                } else {
                    const unsigned at_lane = mine(ghit) ? unsigned(ffs64(group_bits(hitl)) - 1) : leader;
                    Status st = mine(dup) ? Status::duplicate_key() : Status::none();
                    core.finish_insert(index_t(w[u]) * CG + at_lane, val[u], mine(won) && Core::publishes_value, st,
                                       join);
                    StatusHandler::handle(st, status_out, first + pos[u]);
                }
            }
            // slots moving to the next window, or out of windows
            const std::uint64_t adv = (loading & ~ghit & ~gemp) | drained;
            const std::uint64_t last = adv & ballot(rnd[u] + 1 >= budget);
            exhausted |= last;
            Status ple = Status::probing_length_exceeded();
            if constexpr (multi) {
                const std::uint64_t met = last & ballot(count[u] > 0);
                exhausted_dup |= met;
                if (mine(met)) ple = ple + Status::duplicate_key();
            }
            if (rank == 0 && mine(last)) StatusHandler::handle(ple, status_out, first + pos[u]);
            const std::uint64_t move = adv & ~last;
            const W nw = w[u] + step[u];
            if (rank == 0 && mine(move)) mark_overflow(core, w[u], step[u]);
            if (mine(move)) {
                w[u] = nw >= p ? W(nw - p) : nw;
                rnd[u] += 1;
            }
            claim[u] = (claim[u] | gemp) & ~(done | adv);
            live[u] &= ~(done | last);
        };
        // one round: refill, every slot's load or CAS, then every result
        std::uint64_t any_live = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            take_slot(u);
            any_live |= live[u];
        }
        if (any_live == 0 && ballot(avail != 0 || next < len) == 0) break;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < U; ++u) issue_slot(u);
#pragma unroll
        for (int u = 0; u < U; ++u) process_slot(u);
    }
    if constexpr (multi) {
        if (mine(dups)) join(Status::duplicate_key() + Status::max_values_for_key_reached());
        if (mine(exhausted_dup)) join(Status::duplicate_key());
    } else {
        if (mine(dups)) join(Status::duplicate_key());
    }
    if (mine(exhausted)) join(Status::probing_length_exceeded());
}

}  // namespace detail

namespace kernels {

// Wavefronts per SIMD the refill lookup kernel is compiled for (its
// register budget: 512 / waves VGPRs); 0 leaves the choice to the compiler.
// At 4 (128 VGPRs) the u64 kernel runs 4 wavefronts per SIMD instead of 3
// (141 VGPRs), which is worth 13 % on a 4.8 GB table.
#ifndef WARPCORE_REFILL_WAVES
#define WARPCORE_REFILL_WAVES 4
#endif
#if WARPCORE_REFILL_WAVES > 0
#define WC_REFILL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WARPCORE_REFILL_WAVES)))
#else
#define WC_REFILL_OCCUPANCY
#endif
// The same for the refill insert kernel.
#ifndef WARPCORE_REFILL_INSERT_WAVES
#define WARPCORE_REFILL_INSERT_WAVES 0
#endif
#if WARPCORE_REFILL_INSERT_WAVES > 0
#define WC_REFILL_INSERT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WARPCORE_REFILL_INSERT_WAVES)))
#else
#define WC_REFILL_INSERT_OCCUPANCY
#endif

// Bulk lookups: group g of the grid takes keys [g * per, (g + 1) * per).
template<class Core, class StatusHandler, int U, class W, bool Pipe>
WC_D void retrieve_refill_body(const typename Core::key_type* __restrict__ keys_in, index_t num_in,
                               typename Core::out_type* __restrict__ values_out, const Core& core,
                               index_t probing_length, typename StatusHandler::base_type* __restrict__ status_out,
                               detail::RefillEntry<typename Core::key_type, W>* stage) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    detail::DeferredJoin join;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    const index_t per = SDIV(num_in, groups);
    const index_t first = gid * per < num_in ? gid * per : num_in;
    const index_t end = first + per < num_in ? first + per : num_in;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, core.geometry().windows);
    index_t tally = 0;
    detail::retrieve_refill_range<Core, StatusHandler, U, W, Pipe>(keys_in, first, end, values_out, core, budget,
                                                                   status_out, stage + (threadIdx.x & ~(CG - 1)),
                                                                   join, &tally);
    join.flush(core.status_word());
    if constexpr (detail::is_multi_value_core<Core>::value) core.add_tally(tally);
}

template<class Core, class StatusHandler, int U, class W>
__global__ __launch_bounds__(256) WC_REFILL_OCCUPANCY void retrieve_refill(const typename Core::key_type* __restrict__ keys_in,
                                                       index_t num_in, typename Core::out_type* __restrict__ values_out,
                                                       Core core, index_t probing_length,
                                                       typename StatusHandler::base_type* __restrict__ status_out) {
    __shared__ detail::RefillEntry<typename Core::key_type, W> stage[256];
    retrieve_refill_body<Core, StatusHandler, U, W, false>(keys_in, num_in, values_out, core, probing_length,
                                                           status_out, stage);
}

// Wavefronts per SIMD the pipelined refill lookup kernel is compiled for
// (0: the compiler's choice); its loads in flight hold their registers
// across a whole pass.
#ifndef WARPCORE_REFILL_PIPE_WAVES
#define WARPCORE_REFILL_PIPE_WAVES 0
#endif
#if WARPCORE_REFILL_PIPE_WAVES > 0
#define WC_REFILL_PIPE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WARPCORE_REFILL_PIPE_WAVES)))
#else
#define WC_REFILL_PIPE_OCCUPANCY
#endif
template<class Core, class StatusHandler, int U, class W>
__global__ __launch_bounds__(256) WC_REFILL_PIPE_OCCUPANCY void retrieve_refill_pipe(
    const typename Core::key_type* __restrict__ keys_in, index_t num_in,
    typename Core::out_type* __restrict__ values_out, Core core, index_t probing_length,
    typename StatusHandler::base_type* __restrict__ status_out) {
    __shared__ detail::RefillEntry<typename Core::key_type, W> stage[256];
    retrieve_refill_body<Core, StatusHandler, U, W, true>(keys_in, num_in, values_out, core, probing_length,
                                                          status_out, stage);
}

// Bulk inserts. num_dev (nullable): the key count is min(num_in, *num_dev),
// read on the device, so a caller whose count is produced by an earlier
// kernel on the stream needs no host round trip (num_in bounds it).
template<class Core, class StatusHandler, int U, class W, bool HasValues>
__global__ __launch_bounds__(256) WC_REFILL_INSERT_OCCUPANCY void insert_refill(const typename Core::key_type* __restrict__ keys_in,
                                                     const typename Core::value_type* __restrict__ values_in,
                                                     index_t num_in, Core core, index_t probing_length,
                                                     typename StatusHandler::base_type* __restrict__ status_out,
                                                     const index_t* __restrict__ num_dev) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    __shared__ detail::RefillInsertEntry<typename Core::key_type, W, typename Core::value_type> stage[256];
    detail::DeferredJoin join;
    if (num_dev != nullptr && *num_dev < num_in) num_in = *num_dev;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    const index_t per = SDIV(num_in, groups);
    const index_t first = gid * per < num_in ? gid * per : num_in;
    const index_t end = first + per < num_in ? first + per : num_in;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, core.geometry().windows);
    index_t new_keys = 0;
    detail::insert_refill_range<Core, StatusHandler, U, W, HasValues>(
        keys_in, values_in, first, end, core, budget, status_out, stage + (threadIdx.x & ~(CG - 1)), join, &new_keys);
    join.flush(core.status_word());
    if constexpr (detail::is_multi_value_core<Core>::value) core.add_keys(new_keys);
}

}  // namespace kernels

namespace detail {

// Grid of a refill launch: every block resident at once (the occupancy the
// device reports for the kernel, at most 8 blocks of 256 per CU), since
// each group owns an equal contiguous share of the keys and a second round
// of blocks would run at lower occupancy. Queried once per kernel.
template<auto Kernel>
WC_H unsigned refill_grid() noexcept {
    static const unsigned grid = [] {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, Kernel, int(block_size()), 0) != hipSuccess ||
            blocks < 1)
            blocks = 1;
        return unsigned(num_compute_units()) * unsigned(blocks < 8 ? blocks : 8);
    }();
    return grid;
}

// Grid of a lane-owner launch over num_in keys: the resident grid, or
// fewer blocks when that would leave lanes without a key (each wavefront
// takes an equal share, so a share under 64 keys idles lanes for the whole
// launch).
template<auto Kernel>
WC_H unsigned owner_grid(index_t num_in) noexcept {
    const index_t full = refill_grid<Kernel>();
    const index_t needed = SDIV(num_in, index_t(block_size()));   // one key per lane
    return unsigned(needed < 1 ? 1 : needed < full ? needed : full);
}

template<class Core, class StatusHandler, int U, bool HasValues>
WC_H void launch_insert_refill_as(const typename Core::key_type* keys_in, const typename Core::value_type* values_in,
                                  index_t num_in, const Core& core, index_t probing_length,
                                  typename StatusHandler::base_type* status_out, hipStream_t stream,
                                  const index_t* num_dev) noexcept {
    if (core.geometry().windows < (index_t(1) << 31)) {
        constexpr auto k = kernels::insert_refill<Core, StatusHandler, U, std::uint32_t, HasValues>;
        k<<<refill_grid<k>(), block_size(), 0, stream>>>(keys_in, values_in, num_in, core, probing_length, status_out,
                                                         num_dev);
    } else {
        constexpr auto k = kernels::insert_refill<Core, StatusHandler, U, index_t, HasValues>;
        k<<<refill_grid<k>(), block_size(), 0, stream>>>(keys_in, values_in, num_in, core, probing_length, status_out,
                                                         num_dev);
    }
}

// Slots per lane of the lane-owner kernels for a core: 2 (a window read by
// half as many lanes, one 16-byte load each) for packed 8-byte slots in
// windows of 4..16 slots when owner_pairs() is on, else 1.
template<class C, class = void>
struct has_packed_slots : std::false_type {};
template<class C>
struct has_packed_slots<C, std::void_t<decltype(C::packed)>> : std::integral_constant<bool, C::packed> {};

template<class Core>
constexpr bool owner_pairs_capable() noexcept {
    return has_packed_slots<Core>::value && !is_multi_value_core<Core>::value && Core::cg_size() >= 4 &&
           Core::cg_size() <= 16;
}

template<class Core, class StatusHandler, bool HasValues, unsigned SPL>
WC_H void launch_insert_owner_as(const typename Core::key_type* keys_in, const typename Core::value_type* values_in,
                                 index_t num_in, const Core& core, index_t probing_length,
                                 typename StatusHandler::base_type* status_out, hipStream_t stream,
                                 const index_t* num_dev) noexcept {
    if (core.geometry().windows < (index_t(1) << 31)) {
        constexpr auto k = kernels::insert_owner<Core, StatusHandler, std::uint32_t, HasValues, SPL>;
        k<<<owner_grid<k>(num_in), block_size(), 0, stream>>>(keys_in, values_in, num_in, core, probing_length,
                                                              status_out, num_dev);
    } else {
        constexpr auto k = kernels::insert_owner<Core, StatusHandler, index_t, HasValues, SPL>;
        k<<<owner_grid<k>(num_in), block_size(), 0, stream>>>(keys_in, values_in, num_in, core, probing_length,
                                                              status_out, num_dev);
    }
}

template<class Core, class StatusHandler, bool HasValues>
WC_H void launch_insert_owner(const typename Core::key_type* keys_in, const typename Core::value_type* values_in,
                              index_t num_in, const Core& core, index_t probing_length,
                              typename StatusHandler::base_type* status_out, hipStream_t stream,
                              const index_t* num_dev) noexcept {
    if constexpr (owner_pairs_capable<Core>()) {
        if (owner_pairs()) {
            launch_insert_owner_as<Core, StatusHandler, HasValues, 2>(keys_in, values_in, num_in, core,
                                                                      probing_length, status_out, stream, num_dev);
            return;
        }
    }
    launch_insert_owner_as<Core, StatusHandler, HasValues, 1>(keys_in, values_in, num_in, core, probing_length,
                                                              status_out, stream, num_dev);
}

template<class Core, class StatusHandler, unsigned SPL, bool Erase>
WC_H void launch_lookup_owner_as(const typename Core::key_type* keys_in, index_t num_in,
                                 typename Core::out_type* values_out, const Core& core, index_t probing_length,
                                 typename StatusHandler::base_type* status_out, hipStream_t stream) noexcept {
    if (core.geometry().windows < (index_t(1) << 31)) {
        if constexpr (Erase) {
            constexpr auto k = kernels::erase_owner<Core, StatusHandler, std::uint32_t, SPL>;
            k<<<owner_grid<k>(num_in), block_size(), 0, stream>>>(keys_in, num_in, core, probing_length, status_out);
        } else {
            constexpr auto k = kernels::retrieve_owner<Core, StatusHandler, std::uint32_t, SPL>;
            k<<<owner_grid<k>(num_in), block_size(), 0, stream>>>(keys_in, num_in, values_out, core, probing_length,
                                                                  status_out);
        }
    } else {
        if constexpr (Erase) {
            constexpr auto k = kernels::erase_owner<Core, StatusHandler, index_t, SPL>;
            k<<<owner_grid<k>(num_in), block_size(), 0, stream>>>(keys_in, num_in, core, probing_length, status_out);
        } else {
            constexpr auto k = kernels::retrieve_owner<Core, StatusHandler, index_t, SPL>;
            k<<<owner_grid<k>(num_in), block_size(), 0, stream>>>(keys_in, num_in, values_out, core, probing_length,
                                                                  status_out);
        }
    }
}

template<class Core, class StatusHandler, bool Erase = false>
WC_H void launch_lookup_owner(const typename Core::key_type* keys_in, index_t num_in,
                              typename Core::out_type* values_out, const Core& core, index_t probing_length,
                              typename StatusHandler::base_type* status_out, hipStream_t stream) noexcept {
    if constexpr (owner_pairs_capable<Core>()) {
        if (owner_pairs()) {
            launch_lookup_owner_as<Core, StatusHandler, 2, Erase>(keys_in, num_in, values_out, core, probing_length,
                                                                  status_out, stream);
            return;
        }
    }
    launch_lookup_owner_as<Core, StatusHandler, 1, Erase>(keys_in, num_in, values_out, core, probing_length,
                                                          status_out, stream);
}

template<class Core, class StatusHandler>
WC_H void launch_retrieve_owner(const typename Core::key_type* keys_in, index_t num_in,
                                typename Core::out_type* values_out, const Core& core, index_t probing_length,
                                typename StatusHandler::base_type* status_out, hipStream_t stream) noexcept {
    launch_lookup_owner<Core, StatusHandler, false>(keys_in, num_in, values_out, core, probing_length, status_out,
                                                    stream);
}

template<class Core, class StatusHandler>
WC_H void launch_erase_owner(const typename Core::key_type* keys_in, index_t num_in, const Core& core,
                             index_t probing_length, typename StatusHandler::base_type* status_out,
                             hipStream_t stream) noexcept {
    launch_lookup_owner<Core, StatusHandler, true>(keys_in, num_in, nullptr, core, probing_length, status_out,
                                                   stream);
}

// Lane-owner kernels apply to single-value cores whose windows are 2..16
// slots (a 128-byte line of 8- or 16-byte slots at most: a group reads one
// window per owner, so wider windows would hold 32 or 64 windows per lane).
template<class Core>
constexpr bool owner_capable() noexcept {
    return !is_multi_value_core<Core>::value && Core::cg_size() >= 2 && Core::cg_size() <= 16;
}
// Inserts also for multi-value cores (a claim stores one more value).
template<class Core>
constexpr bool owner_insert_capable() noexcept {
    return Core::cg_size() >= 2 && Core::cg_size() <= 16;
}

template<class Core, class StatusHandler, int U, bool HasValues>
WC_H void launch_insert_refill(const typename Core::key_type* keys_in, const typename Core::value_type* values_in,
                               index_t num_in, const Core& core, index_t probing_length,
                               typename StatusHandler::base_type* status_out, hipStream_t stream,
                               const index_t* num_dev = nullptr) noexcept {
    if constexpr (owner_insert_capable<Core>()) {
        if (owner_inserts()) {
            launch_insert_owner<Core, StatusHandler, HasValues>(keys_in, values_in, num_in, core, probing_length,
                                                                status_out, stream, num_dev);
            return;
        }
    }
    launch_insert_refill_as<Core, StatusHandler, U, HasValues>(keys_in, values_in, num_in, core, probing_length,
                                                               status_out, stream, num_dev);
}

template<class Core, class StatusHandler, int U, bool Pipe>
WC_H void launch_retrieve_refill_as(const typename Core::key_type* keys_in, index_t num_in,
                                    typename Core::out_type* values_out, const Core& core, index_t probing_length,
                                    typename StatusHandler::base_type* status_out, hipStream_t stream) noexcept {
    if (core.geometry().windows < (index_t(1) << 31)) {
        constexpr auto k = Pipe ? kernels::retrieve_refill_pipe<Core, StatusHandler, U, std::uint32_t>
                                : kernels::retrieve_refill<Core, StatusHandler, U, std::uint32_t>;
        k<<<refill_grid<k>(), block_size(), 0, stream>>>(keys_in, num_in, values_out, core, probing_length, status_out);
    } else {
        constexpr auto k = Pipe ? kernels::retrieve_refill_pipe<Core, StatusHandler, U, index_t>
                                : kernels::retrieve_refill<Core, StatusHandler, U, index_t>;
        k<<<refill_grid<k>(), block_size(), 0, stream>>>(keys_in, num_in, values_out, core, probing_length, status_out);
    }
}

template<class Core, class StatusHandler, int U>
WC_H void launch_retrieve_refill(const typename Core::key_type* keys_in, index_t num_in,
                                 typename Core::out_type* values_out, const Core& core, index_t probing_length,
                                 typename StatusHandler::base_type* status_out, hipStream_t stream) noexcept {
    if constexpr (owner_capable<Core>()) {
        if (owner_lookups()) {
            launch_retrieve_owner<Core, StatusHandler>(keys_in, num_in, values_out, core, probing_length, status_out,
                                                       stream);
            return;
        }
    }
    constexpr int UP = U == WARPCORE_REFILL_KEYS ? WARPCORE_REFILL_PIPE_KEYS : U;
    if (pipelined_lookups())
        launch_retrieve_refill_as<Core, StatusHandler, UP, true>(keys_in, num_in, values_out, core, probing_length,
                                                                 status_out, stream);
    else
        launch_retrieve_refill_as<Core, StatusHandler, U, false>(keys_in, num_in, values_out, core, probing_length,
                                                                 status_out, stream);
}

}  // namespace detail
}  // namespace warpcore
