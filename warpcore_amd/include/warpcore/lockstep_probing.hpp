// Lockstep cooperative probing: a group of CG lanes takes U keys at a time
// and walks their windows together, one window per key per round.
//
// The streaming kernels (batched_probing.hpp) refill a group's key slots
// as soon as a key finishes; the price is a per-slot state machine whose
// divergent branches cost more issue slots than the probing itself. Here
// the group's U keys start together and every round issues U window loads
// (a finished key reads the table's first window: every finished key of
// the CU reads that one line, so it stays in L1), then
// updates every key with selects: no per-key branches, one group-uniform
// loop. Because all U keys start at round 0, each live key has visited
// r + 1 windows in round r, so the probing budget is one shared counter.
//
// The keys of a batch are hashed once: lane j < U of the group loads key j,
// computes its first window and step, and stages them in LDS; every lane
// then reads the U entries (same address across the group, an LDS
// broadcast). A hit is written by the lane that holds the slot (no
// shuffle); statuses are written by lane u for key u.
//
// Per key the windows are read in the same order as the one-key protocol
// (single_value_hash_table.cuh:267-307), so results and statuses are those
// of the reference.
#pragma once

#include "batched_probing.hpp"
#include "engine.hpp"
#include "tile.hpp"

namespace warpcore {
namespace detail {

// Schemes whose window step is fixed per key (double hashing, linear
// probing): the lockstep kernel advances every key by its own step.
template<class Scheme> struct fixed_step_scheme : std::false_type {};
template<class H1, class H2, index_t CG>
struct fixed_step_scheme<probing_schemes::DoubleHashing<H1, H2, CG>> : std::true_type {};
template<class H, index_t CG>
struct fixed_step_scheme<probing_schemes::LinearProbing<H, CG>> : std::true_type {};

// Keys per group and batch of the lockstep kernels.
#ifndef WARPCORE_LOCKSTEP_KEYS
#define WARPCORE_LOCKSTEP_KEYS 8
#endif
#ifndef WARPCORE_LOCKSTEP_INSERT_KEYS
#define WARPCORE_LOCKSTEP_INSERT_KEYS 4
#endif
template<unsigned CG> constexpr int lockstep_keys() noexcept {
    return CG < WARPCORE_LOCKSTEP_KEYS ? int(CG) : WARPCORE_LOCKSTEP_KEYS;
}
template<unsigned CG> constexpr int lockstep_insert_keys() noexcept {
    return CG < WARPCORE_LOCKSTEP_INSERT_KEYS ? int(CG) : WARPCORE_LOCKSTEP_INSERT_KEYS;
}

// One staged key: the key and its window cursor.
template<class K, class W>
struct alignas(16) StagedKey {
    K key;
    W w, step;
    std::uint32_t state;   // kKeyOk / kKeyInvalid / kNotReady / kPastEnd
};

}  // namespace detail

namespace detail {

// Lookups of keys_in[first], keys_in[first + stride], ... below `end`, U
// at a time per group (lanes of one group call this together). W: window
// index type, std::uint32_t when the table has fewer than 2^31 windows,
// else index_t. `stage` holds U entries of this group.
template<class Core, class StatusHandler, int U, class W>
WC_D void retrieve_lockstep_range(const typename Core::key_type* __restrict__ keys_in, index_t first, index_t end,
                                  index_t stride, typename Core::out_type* __restrict__ values_out, const Core& core,
                                  index_t budget, typename StatusHandler::base_type* __restrict__ status_out,
                                  StagedKey<typename Core::key_type, W>* my, DeferredJoin& join) {
    using K = typename Core::key_type;
    using O = typename Core::out_type;
    using Slot = typename Core::probe_slot_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    using Entry = StagedKey<K, W>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    static_assert(U >= 1 && unsigned(U) <= CG, "a batch holds at most one key per lane");
    const unsigned lane = lane_id();
    const unsigned rank = lane & (CG - 1);
    const unsigned shift = lane & ~(CG - 1);
    constexpr std::uint64_t full = tile<CG>::full_mask();
    auto gballot = [&](bool p) -> std::uint64_t {
        const std::uint64_t m = __builtin_amdgcn_ballot_w64(p);
        if constexpr (CG == 64) return m;
        else return (m >> shift) & full;
    };
    const probing_schemes::Geometry& geo = core.geometry();
    const W p = W(geo.windows);
    const Slot* __restrict__ slots = core.slot_ptr();

    for (index_t base = first; base < end; base += stride) {
        // stage the batch: lane j < U hashes key base + j
        if (rank < unsigned(U)) {
            const index_t i = base + rank;
            const K key = keys_in[i < end ? i : end - 1];
            Entry e;
            e.key = key;
            e.w = 0;
            e.step = 1;
            if (i >= end) e.state = kPastEnd;
            else if (!core.ready()) e.state = kNotReady;
            else if (!core.is_valid_key(key)) e.state = kKeyInvalid;
            else {
                Cursor c;
                c.begin(key, core.hash_seed(), geo);
                e.w = W(c.w);
                e.step = W(c.step);
                e.state = kKeyOk;
            }
            my[rank] = e;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        K key[U];
        W w[U], step[U];
        Status st[U];
        std::uint32_t live = 0;   // bit u: key u still probing
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Entry e = my[u];
            key[u] = e.key;
            w[u] = e.w;
            step[u] = e.step;
            st[u] = Status::none();
            if (e.state == kKeyOk) live |= 1u << u;
            else if (e.state == kNotReady) st[u] = Status::not_initialized();
            else if (e.state == kKeyInvalid) {
                st[u] = Status::invalid_key();
                join(st[u]);
            }
        }
        __builtin_amdgcn_wave_barrier();   // entries read before the next batch overwrites them

        for (index_t round = 0; live != 0; ++round) {
            Slot s[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {   // a finished key reads the first window: one line every
                const W at = (live >> u) & 1u ? w[u] : W(0);   // finished key shares, an L1 hit
                s[u] = load_slot(slots + index_t(at) * CG + rank);
            }
            const bool last = round + 1 >= budget;
            std::uint32_t next = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool on = (live >> u) & 1u;
                const K tk = Core::key_of(s[u]);
                const std::uint64_t hit = gballot(tk == key[u]);
                const std::uint64_t empty = gballot(Core::is_empty_key(tk));
                if (on && hit != 0 && rank == unsigned(ffs64(hit) - 1))
                    values_out[base + u] = core.found(Core::value_of(s[u]));
                const bool miss = on && hit == 0 && empty != 0;
                const bool out_of_windows = on && hit == 0 && empty == 0 && last;
                if (miss) {
                    O o{};
                    st[u] = core.missing(o);
                    join(st[u]);
                    if (!st[u].has_any() && rank == 0) values_out[base + u] = o;
                }
                if (out_of_windows) {
                    st[u] = Status::probing_length_exceeded();
                    join(st[u]);
                }
                const bool cont = on && hit == 0 && empty == 0 && !last;
                if (cont) {
                    W nw = w[u] + step[u];
                    w[u] = nw >= p ? nw - p : nw;
                    next |= 1u << u;
                }
            }
            live = next;
        }

#pragma unroll
        for (int u = 0; u < U; ++u)
            if (rank == unsigned(u) && base + u < end) StatusHandler::handle(st[u], status_out, base + u);
    }
}

// Erase of keys_in[first], keys_in[first + stride], ... below `end`, U at a
// time per group (single_value_hash_table.cuh:425-468 per key): the lane
// holding a hit writes the tombstone over the key; a window with an EMPTY
// lane and no hit ends in key_not_found; misses and exhausted budgets are
// returned, only invalid keys are joined. `Core` provides erase_slot().
template<class Core, class StatusHandler, int U, class W>
WC_D void erase_lockstep_range(const typename Core::key_type* __restrict__ keys_in, index_t first, index_t end,
                               index_t stride, const Core& core, index_t budget,
                               typename StatusHandler::base_type* __restrict__ status_out,
                               StagedKey<typename Core::key_type, W>* my, DeferredJoin& join) {
    using K = typename Core::key_type;
    using Slot = typename Core::probe_slot_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    using Entry = StagedKey<K, W>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    static_assert(U >= 1 && unsigned(U) <= CG, "a batch holds at most one key per lane");
    const unsigned lane = lane_id();
    const unsigned rank = lane & (CG - 1);
    const unsigned shift = lane & ~(CG - 1);
    constexpr std::uint64_t full = tile<CG>::full_mask();
    auto gballot = [&](bool p) -> std::uint64_t {
        const std::uint64_t m = __builtin_amdgcn_ballot_w64(p);
        if constexpr (CG == 64) return m;
        else return (m >> shift) & full;
    };
    const probing_schemes::Geometry& geo = core.geometry();
    const W p = W(geo.windows);
    const Slot* __restrict__ slots = core.slot_ptr();

    for (index_t base = first; base < end; base += stride) {
        if (rank < unsigned(U)) {
            const index_t i = base + rank;
            const K key = keys_in[i < end ? i : end - 1];
            Entry e;
            e.key = key;
            e.w = 0;
            e.step = 1;
            if (i >= end) e.state = kPastEnd;
            else if (!core.ready()) e.state = kNotReady;
            else if (!core.is_valid_key(key)) e.state = kKeyInvalid;
            else {
                Cursor c;
                c.begin(key, core.hash_seed(), geo);
                e.w = W(c.w);
                e.step = W(c.step);
                e.state = kKeyOk;
            }
            my[rank] = e;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        K key[U];
        W w[U], step[U];
        Status st[U];
        std::uint32_t live = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Entry e = my[u];
            key[u] = e.key;
            w[u] = e.w;
            step[u] = e.step;
            st[u] = Status::none();
            if (e.state == kKeyOk) live |= 1u << u;
            else if (e.state == kNotReady) st[u] = Status::not_initialized();
            else if (e.state == kKeyInvalid) {
                st[u] = Status::invalid_key();
                join(st[u]);
            }
        }
        __builtin_amdgcn_wave_barrier();

        for (index_t round = 0; live != 0; ++round) {
            Slot s[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const W at = (live >> u) & 1u ? w[u] : W(0);
                s[u] = load_slot(slots + index_t(at) * CG + rank);
            }
            const bool last = round + 1 >= budget;
            std::uint32_t next = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool on = (live >> u) & 1u;
                const K tk = Core::key_of(s[u]);
                const std::uint64_t hit = gballot(tk == key[u]);
                const std::uint64_t empty = gballot(Core::is_empty_key(tk));
                if (on && hit != 0 && rank == unsigned(ffs64(hit) - 1)) core.erase_slot(index_t(w[u]) * CG + rank);
                if (on && hit == 0 && empty != 0) st[u] = Status::key_not_found();
                if (on && hit == 0 && empty == 0 && last) st[u] = Status::probing_length_exceeded();
                if (on && hit == 0 && empty == 0 && !last) {
                    W nw = w[u] + step[u];
                    w[u] = nw >= p ? nw - p : nw;
                    next |= 1u << u;
                }
            }
            live = next;
        }

#pragma unroll
        for (int u = 0; u < U; ++u)
            if (rank == unsigned(u) && base + u < end) StatusHandler::handle(st[u], status_out, base + u);
    }
}

// Insert with U keys per group in lockstep (single_value_hash_table.cuh:742-823
// per key). A key is in one of two phases: `load` (read its current
// window) or `cas` (its window had EMPTY lanes; the lowest untried one
// claims). Every round issues all window loads (keys not loading read the
// first window) and all claiming CASes before any result is used.
// Keys finish at different rounds; a finished key's loads read the first
// window until the group's batch is done. `Core` provides the hooks of insert_batched.
template<class Core, class StatusHandler, int U, class W, bool HasValues>
WC_D void insert_lockstep_range(const typename Core::key_type* __restrict__ keys_in,
                                const typename Core::value_type* __restrict__ values_in, index_t first, index_t end,
                                index_t stride, const Core& core, index_t budget,
                                typename StatusHandler::base_type* __restrict__ status_out,
                                StagedKey<typename Core::key_type, W>* my, DeferredJoin& join) {
    using K = typename Core::key_type;
    using V = typename Core::value_type;
    using Probe = typename Core::insert_probe_type;
    using Cursor = WindowCursor<typename Core::probing_scheme_type>;
    using Entry = StagedKey<K, W>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    static_assert(U >= 1 && unsigned(U) <= CG, "a batch holds at most one key per lane");

    const unsigned lane = lane_id();
    const unsigned rank = lane & (CG - 1);
    const unsigned shift = lane & ~(CG - 1);
    constexpr std::uint64_t full = tile<CG>::full_mask();
    auto gballot = [&](bool p) -> std::uint64_t {
        const std::uint64_t m = __builtin_amdgcn_ballot_w64(p);
        if constexpr (CG == 64) return m;
        else return (m >> shift) & full;
    };
    const probing_schemes::Geometry& geo = core.geometry();
    const W p = W(geo.windows);
    auto* __restrict__ slots = core.slot_ptr();

    for (index_t base = first; base < end; base += stride) {
        if (rank < unsigned(U)) {
            const index_t i = base + rank;
            const K key = keys_in[i < end ? i : end - 1];
            Entry e;
            e.key = key;
            e.w = 0;
            e.step = 1;
            if (i >= end) e.state = kPastEnd;
            else if (!core.ready()) e.state = kNotReady;
            else if (!core.is_valid_key(key)) e.state = kKeyInvalid;
            else {
                Cursor c;
                c.begin(key, core.hash_seed(), geo);
                e.w = W(c.w);
                e.step = W(c.step);
                e.state = kKeyOk;
            }
            my[rank] = e;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        K key[U];
        V val[U];
        W w[U], step[U];
        index_t left[U];
        std::uint64_t empty[U];   // EMPTY lanes of the current window not tried yet
        Probe seen[U];            // this lane's slot as read in the current window
        Status st[U];
        std::uint32_t live = 0, claiming = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Entry e = my[u];
            key[u] = e.key;
            w[u] = e.w;
            step[u] = e.step;
            left[u] = budget;
            empty[u] = 0;
            seen[u] = Probe{};
            st[u] = Status::none();
            if constexpr (HasValues) {
                const index_t i = base + u;
                val[u] = values_in[i < end ? i : end - 1];
            } else {
                val[u] = V{};
            }
            if (e.state == kKeyOk) live |= 1u << u;
            else if (e.state == kNotReady) st[u] = Status::not_initialized();
            else if (e.state == kKeyInvalid) {
                st[u] = Status::invalid_key();
                join(st[u]);
            }
        }
        __builtin_amdgcn_wave_barrier();

        while (live != 0) {
            Probe fresh[U];
            typename Core::cas_word old[U];
            bool lead[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {   // keys not loading read the first window (an L1 hit)
                const bool loading = ((live & ~claiming) >> u) & 1u;
                fresh[u] = Core::load_insert_probe(slots + index_t(loading ? w[u] : W(0)) * CG + rank);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                lead[u] = ((claiming >> u) & 1u) && rank == unsigned(ffs64(empty[u]) - 1);
                old[u] = typename Core::cas_word{};
                if (lead[u]) old[u] = core.cas_issue(slots + index_t(w[u]) * CG + rank, seen[u], key[u], val[u]);
            }
            std::uint32_t next_live = live, next_claiming = claiming;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!((live >> u) & 1u)) continue;
                index_t where = ~index_t(0);
                bool published = false;
                bool finished = false;
                bool advance = false;
                if (!((claiming >> u) & 1u)) {
                    seen[u] = fresh[u];
                    const K tk = Core::probe_key(seen[u]);
                    const std::uint64_t hit = gballot(tk == key[u]);
                    if (hit) {
                        st[u] = Status::duplicate_key();
                        where = index_t(w[u]) * CG + unsigned(ffs64(hit) - 1);
                        finished = true;
                    } else {
                        empty[u] = gballot(Core::is_empty_key(tk));
                        if (empty[u]) next_claiming |= 1u << u;
                        else advance = true;
                    }
                } else {
                    const unsigned leader = unsigned(ffs64(empty[u]) - 1);
                    const int outcome = lead[u] ? core.cas_outcome(old[u], seen[u], key[u]) : -1;
                    if (gballot(outcome == 2)) {
                        st[u] = Status::duplicate_key();
                        where = index_t(w[u]) * CG + leader;
                        finished = true;
                    } else if (gballot(outcome == 1)) {
                        where = index_t(w[u]) * CG + leader;
                        published = Core::publishes_value;
                        finished = true;
                    } else if (!gballot(outcome == 3)) {
                        empty[u] &= empty[u] - 1;
                        if (!empty[u]) {
                            next_claiming &= ~(1u << u);
                            advance = true;
                        }
                    }
                }
                if (advance) {
                    if (--left[u] == 0) {
                        st[u] = Status::probing_length_exceeded();
                        finished = true;
                    } else {
                        const W nw = w[u] + step[u];
                        w[u] = nw >= p ? nw - p : nw;
                    }
                }
                if (finished) {
                    if (st[u].has_any()) join(st[u]);
                    if (rank == 0 && where != ~index_t(0)) core.finish_insert(where, val[u], published, st[u], join);
                    // finish_insert may add a flag (counter overflow) on lane 0
                    st[u] = Status::from_base(tile<CG>::shfl(st[u].base(), 0));
                    next_live &= ~(1u << u);
                    next_claiming &= ~(1u << u);
                }
            }
            live = next_live;
            claiming = next_claiming;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (rank == unsigned(u) && base + u < end) StatusHandler::handle(st[u], status_out, base + u);
    }
}

}  // namespace detail

namespace kernels {

template<class Core, class StatusHandler, int U, class W>
__global__ __launch_bounds__(256) void retrieve_lockstep(const typename Core::key_type* __restrict__ keys_in,
                                                         index_t num_in,
                                                         typename Core::out_type* __restrict__ values_out,
                                                         Core core, index_t probing_length,
                                                         typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    __shared__ detail::StagedKey<typename Core::key_type, W> stage[256 / CG * U];
    detail::DeferredJoin join;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, core.geometry().windows);
    detail::retrieve_lockstep_range<Core, StatusHandler, U, W>(keys_in, gid * U, num_in, groups * U, values_out, core,
                                                               budget, status_out, stage + (threadIdx.x / CG) * U,
                                                               join);
    join.flush(core.status_word());
}

template<class Core, class StatusHandler, int U, class W>
__global__ __launch_bounds__(256) void erase_lockstep(const typename Core::key_type* __restrict__ keys_in,
                                                      index_t num_in, Core core, index_t probing_length,
                                                      typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    __shared__ detail::StagedKey<typename Core::key_type, W> stage[256 / CG * U];
    detail::DeferredJoin join;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, core.geometry().windows);
    detail::erase_lockstep_range<Core, StatusHandler, U, W>(keys_in, gid * U, num_in, groups * U, core, budget,
                                                            status_out, stage + (threadIdx.x / CG) * U, join);
    join.flush(core.status_word());
}

template<class Core, class StatusHandler, int U, class W, bool HasValues>
__global__ __launch_bounds__(256) void insert_lockstep(const typename Core::key_type* __restrict__ keys_in,
                                                       const typename Core::value_type* __restrict__ values_in,
                                                       index_t num_in, Core core, index_t probing_length,
                                                       typename StatusHandler::base_type* __restrict__ status_out) {
    constexpr unsigned CG = unsigned(Core::cg_size());
    __shared__ detail::StagedKey<typename Core::key_type, W> stage[256 / CG * U];
    detail::DeferredJoin join;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    const index_t groups = (index_t(gridDim.x) * blockDim.x) / CG;
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, core.geometry().windows);
    detail::insert_lockstep_range<Core, StatusHandler, U, W, HasValues>(
        keys_in, values_in, gid * U, num_in, groups * U, core, budget, status_out, stage + (threadIdx.x / CG) * U,
        join);
    join.flush(core.status_word());
}

}  // namespace kernels

namespace detail {

// Grid of a lockstep launch: one group per U keys, at most
// WARPCORE_LOCKSTEP_BLOCKS_PER_CU blocks of 256 per CU (the loop is
// grid-stride).
#ifndef WARPCORE_LOCKSTEP_BLOCKS_PER_CU
#define WARPCORE_LOCKSTEP_BLOCKS_PER_CU 8
#endif
template<unsigned CG, int U>
WC_H unsigned grid_for_lockstep(index_t items) noexcept {
    const index_t groups = SDIV(items, index_t(U));
    const index_t blocks = SDIV(groups * CG, block_size());
    const index_t cap = index_t(num_compute_units()) * WARPCORE_LOCKSTEP_BLOCKS_PER_CU;
    const index_t g = std::min(blocks, cap);
    return unsigned(g == 0 ? 1 : g);
}

// Kernels of bulk calls on tables probed by cooperative groups
// (process-wide setting): bit 0 set = lockstep inserts, bit 1 set =
// lockstep lookups and erases, bit 2 set = refill lookups (refill_probing.hpp,
// before bit 1), bit 3 set = refill inserts (before bit 0); a clear bit
// selects the streaming state machine. Bit 4 pipelines the refill lookups
// (each slot's request consumed as soon as it returns). Bits 6 / 7 select
// the lane-owner lookups / inserts for single-value tables with windows of
// 2..16 slots (owner_probing.hpp), ahead of the refill kernels. The
// default, lane-owner inserts and lookups, else refill inserts and
// pipelined refill lookups, is the fastest measured on MI355X (Gops/s,
// 2^28 keys at LF 0.9, lookups at 50 % hits; insert / lookup):
//   SVHT<u64,u64> cg 8:  refill 12.0 / 16.1, pipelined lookups 16.6-18.8,
//                        lane owners 15.9 / 17.5
//   SVHT<u32,u32> cg 16: refill 11.0 / 18.9, pipelined lookups 20.5-23.2,
//                        lane owners 18.6 / 23.5
// (lockstep batches wait for their longest probe sequence, which at load
// factor 0.9 is many windows; refill slots do not wait; a pipelined refill
// insert measured 11.3 and 11.7 and was dropped).
inline std::atomic<int>& group_kernel_setting() noexcept {
    static std::atomic<int> k{476};
    return k;
}
WC_H int group_kernel() noexcept { return group_kernel_setting().load(std::memory_order_relaxed); }
WC_H void set_group_kernel(int k) noexcept { group_kernel_setting().store(k, std::memory_order_relaxed); }
WC_H bool lockstep_inserts() noexcept { return (group_kernel() & 1) != 0; }
WC_H bool lockstep_lookups() noexcept { return (group_kernel() & 2) != 0; }
WC_H bool refill_lookups() noexcept { return (group_kernel() & 4) != 0; }
WC_H bool refill_inserts() noexcept { return (group_kernel() & 8) != 0; }
// bit 4: pipelined refill lookups
WC_H bool pipelined_lookups() noexcept { return (group_kernel() & 16) != 0; }
// bit 6: lane-owner lookups, bit 7: lane-owner inserts (owner_probing.hpp;
// single-value tables; before the refill bits)
WC_H bool owner_lookups() noexcept { return (group_kernel() & 64) != 0; }
WC_H bool owner_inserts() noexcept { return (group_kernel() & 128) != 0; }
// bit 8: lane-owner kernels read packed 8-byte slots two per lane (a window
// of W slots by W/2 lanes, one 16-byte load each)
WC_H bool owner_pairs() noexcept { return (group_kernel() & 256) != 0; }

template<class Core, class StatusHandler, int U>
WC_H void launch_erase_lockstep(const typename Core::key_type* keys_in, index_t num_in, const Core& core,
                                index_t probing_length, typename StatusHandler::base_type* status_out,
                                hipStream_t stream) noexcept {
    constexpr unsigned CG = unsigned(Core::cg_size());
    const unsigned grid = grid_for_lockstep<CG, U>(num_in);
    if (core.geometry().windows < (index_t(1) << 31))
        kernels::erase_lockstep<Core, StatusHandler, U, std::uint32_t>
            <<<grid, block_size(), 0, stream>>>(keys_in, num_in, core, probing_length, status_out);
    else
        kernels::erase_lockstep<Core, StatusHandler, U, index_t>
            <<<grid, block_size(), 0, stream>>>(keys_in, num_in, core, probing_length, status_out);
}

template<class Core, class StatusHandler, int U>
WC_H void launch_retrieve_lockstep(const typename Core::key_type* keys_in, index_t num_in,
                                   typename Core::out_type* values_out, const Core& core, index_t probing_length,
                                   typename StatusHandler::base_type* status_out, hipStream_t stream) noexcept {
    constexpr unsigned CG = unsigned(Core::cg_size());
    const unsigned grid = grid_for_lockstep<CG, U>(num_in);
    if (core.geometry().windows < (index_t(1) << 31))
        kernels::retrieve_lockstep<Core, StatusHandler, U, std::uint32_t>
            <<<grid, block_size(), 0, stream>>>(keys_in, num_in, values_out, core, probing_length, status_out);
    else
        kernels::retrieve_lockstep<Core, StatusHandler, U, index_t>
            <<<grid, block_size(), 0, stream>>>(keys_in, num_in, values_out, core, probing_length, status_out);
}

template<class Core, class StatusHandler, int U, bool HasValues>
WC_H void launch_insert_lockstep(const typename Core::key_type* keys_in, const typename Core::value_type* values_in,
                                 index_t num_in, const Core& core, index_t probing_length,
                                 typename StatusHandler::base_type* status_out, hipStream_t stream) noexcept {
    constexpr unsigned CG = unsigned(Core::cg_size());
    const unsigned grid = grid_for_lockstep<CG, U>(num_in);
    if (core.geometry().windows < (index_t(1) << 31))
        kernels::insert_lockstep<Core, StatusHandler, U, std::uint32_t, HasValues>
            <<<grid, block_size(), 0, stream>>>(keys_in, values_in, num_in, core, probing_length, status_out);
    else
        kernels::insert_lockstep<Core, StatusHandler, U, index_t, HasValues>
            <<<grid, block_size(), 0, stream>>>(keys_in, values_in, num_in, core, probing_length, status_out);
}

}  // namespace detail
}  // namespace warpcore
