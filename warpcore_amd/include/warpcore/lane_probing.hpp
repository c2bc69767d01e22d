// Lane-per-key probing for the bulk operations.
//
// The table layout, windows, probing sequence and per-key protocol are the
// ones of the cooperative kernels (one window = CG slots, a hit anywhere in
// the window ends the probe, the lowest EMPTY slot is claimed first, then
// the next, a CAS returning the key itself is a duplicate); what changes is
// who executes it. Here one lane runs the whole protocol of one key: it
// reads the key's window as 16-byte chunks (8 x dwordx4 for a 128-byte
// window), scans the CG slots in registers and issues its own CAS. Control
// flow is then per key instead of per group, so every wavefront instruction
// serves 64 keys rather than 64/CG, and the 64 lanes' hash and window
// arithmetic runs fully parallel.
//
// Each lane keeps U keys in flight as a streaming state machine: every
// round all U windows and the next input keys are read (no load under a
// branch, which would make the compiler wait for all outstanding loads
// before issuing the next), then each key advances, and a finished key
// writes its result and is replaced by the lane's next input key.
//
// Used when a window is a whole number of 16-byte chunks of at most 256
// bytes; other shapes use the cooperative kernels of batched_probing.hpp.
#pragma once

#include "batched_probing.hpp"

namespace warpcore {
namespace detail {

template<class Slot, unsigned CG>
struct LaneWindow {
    static constexpr unsigned bytes = unsigned(sizeof(Slot)) * CG;
    static constexpr unsigned chunks = bytes / 16;
    static constexpr unsigned per_chunk = 16 / unsigned(sizeof(Slot));
    static constexpr bool usable = sizeof(Slot) <= 16 && (16 % sizeof(Slot)) == 0 && bytes % 16 == 0 &&
                                   bytes >= 16 && bytes <= 256;
};

// Makes the compiler hold the whole chunk in registers at this point; used
// after all loads of a round are issued, so that it neither splits a
// 16-byte load nor defers part of it until after a compare.
WC_D void pin(uint4& c) noexcept { asm volatile("" : "+v"(c.x), "+v"(c.y), "+v"(c.z), "+v"(c.w)); }

template<class Slot>
WC_D Slot chunk_slot(const uint4& c, unsigned i) noexcept {
    Slot s;
    __builtin_memcpy(&s, reinterpret_cast<const char*>(&c) + i * sizeof(Slot), sizeof(Slot));
    return s;
}

}  // namespace detail

namespace kernels {

template<class Core, class StatusHandler, int U>
__global__ __launch_bounds__(256) void retrieve_lanes(const typename Core::key_type* __restrict__ keys_in,
                                                      index_t num_in, typename Core::out_type* __restrict__ values_out,
                                                      Core core, index_t probing_length,
                                                      typename StatusHandler::base_type* __restrict__ status_out) {
    using K = typename Core::key_type;
    using O = typename Core::out_type;
    using Slot = typename Core::probe_slot_type;
    using Cursor = detail::WindowCursor<typename Core::probing_scheme_type>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    using W = detail::LaneWindow<Slot, CG>;
    static_assert(W::usable, "window shape not handled lane-per-key");
    detail::DeferredJoin join;
    const index_t tid = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const index_t lanes = index_t(gridDim.x) * blockDim.x;
    const index_t stride = lanes * U;
    const probing_schemes::Geometry& geo = core.geometry();
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, geo.windows);
    const uint4* __restrict__ table = reinterpret_cast<const uint4*>(core.slot_ptr());
    const index_t last = num_in - 1;

    index_t idx[U], nidx[U];
    K key[U], nkey[U];
    Cursor cur[U];
    std::uint32_t phase[U];
    Status st[U];
    O out[U];

    auto start = [&](int u) {
        out[u] = O{};
        st[u] = Status::none();
        cur[u].left = budget;
        if (idx[u] >= num_in) { phase[u] = detail::kDone; return; }
        if (!core.ready()) { st[u] = Status::not_initialized(); phase[u] = detail::kEmit; return; }
        if (!core.is_valid_key(key[u])) { st[u] = Status::invalid_key(); join(st[u]); phase[u] = detail::kEmit; return; }
        cur[u].begin(key[u], core.hash_seed(), geo);
        cur[u].left = budget;
        phase[u] = detail::kLoad;
    };
#pragma unroll
    for (int u = 0; u < U; ++u) {
        idx[u] = tid + index_t(u) * lanes;
        nidx[u] = idx[u] + stride;
        key[u] = keys_in[idx[u] < num_in ? idx[u] : last];
        cur[u].w = 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) start(u);

    while (true) {
        bool busy = false;
#pragma unroll
        for (int u = 0; u < U; ++u) busy |= phase[u] != detail::kDone;
        if (!busy) break;
        uint4 win[U][W::chunks];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint4* p = table + cur[u].w * W::chunks;
#pragma unroll
            for (unsigned c = 0; c < W::chunks; ++c) win[u][c] = p[c];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) nkey[u] = keys_in[nidx[u] < num_in ? nidx[u] : last];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (unsigned c = 0; c < W::chunks; ++c) detail::pin(win[u][c]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (phase[u] == detail::kLoad) {
                bool hit = false, empty = false;
                O found{};
                // the reference takes the lowest hitting lane; scan top down so
                // the lowest hit is the value kept
#pragma unroll
                for (int c = int(W::chunks) - 1; c >= 0; --c) {
#pragma unroll
                    for (int i = int(W::per_chunk) - 1; i >= 0; --i) {
                        const Slot s = detail::chunk_slot<Slot>(win[u][c], unsigned(i));
                        const bool h = Core::key_of(s) == key[u];
                        found = h ? core.found(Core::value_of(s)) : found;
                        hit |= h;
                        empty |= Core::is_empty_key(Core::key_of(s));
                    }
                }
                if (hit) {
                    out[u] = found;
                    phase[u] = detail::kEmit;
                } else if (empty) {
                    st[u] = core.missing(out[u]);
                    join(st[u]);
                    phase[u] = detail::kEmit;
                } else if (!cur[u].advance(geo)) {
                    st[u] = Status::probing_length_exceeded();
                    join(st[u]);
                    phase[u] = detail::kEmit;
                }
            }
            if (phase[u] == detail::kEmit) {
                if (!st[u].has_any()) values_out[idx[u]] = out[u];
                StatusHandler::handle(st[u], status_out, idx[u]);
                idx[u] = nidx[u];
                key[u] = nkey[u];
                nidx[u] += stride;
                start(u);
            }
        }
    }
    join.flush(core.status_word());
}

template<class Core, class StatusHandler, int U, bool HasValues>
__global__ __launch_bounds__(256) void insert_lanes(const typename Core::key_type* __restrict__ keys_in,
                                                    const typename Core::value_type* __restrict__ values_in,
                                                    index_t num_in, Core core, index_t probing_length,
                                                    typename StatusHandler::base_type* __restrict__ status_out) {
    using K = typename Core::key_type;
    using V = typename Core::value_type;
    using Slot = typename Core::probe_slot_type;
    using Cursor = detail::WindowCursor<typename Core::probing_scheme_type>;
    constexpr unsigned CG = unsigned(Core::cg_size());
    using W = detail::LaneWindow<Slot, CG>;
    static_assert(W::usable, "window shape not handled lane-per-key");
    static_assert(CG <= 64, "window mask is 64 bits");
    detail::DeferredJoin join;
    const index_t tid = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const index_t lanes = index_t(gridDim.x) * blockDim.x;
    const index_t stride = lanes * U;
    const probing_schemes::Geometry& geo = core.geometry();
    const index_t budget = probing_schemes::detail::window_budget(probing_length, CG, geo.windows);
    auto* __restrict__ slots = core.slot_ptr();
    const uint4* __restrict__ table = reinterpret_cast<const uint4*>(slots);
    const index_t last = num_in - 1;

    index_t idx[U], nidx[U];
    K key[U], nkey[U];
    V val[U], nval[U];
    Cursor cur[U];
    std::uint32_t phase[U];
    std::uint64_t empty[U];                 // untried EMPTY slots of the current window
    typename Core::insert_probe_type seen[U];  // expected word of the next CAS (EMPTY;
                                               // refreshed if a packed slot's value bits differ)
    index_t where[U];
    Status st[U];
    bool published[U];

    auto start = [&](int u) {
        st[u] = Status::none();
        empty[u] = 0;
        where[u] = ~index_t(0);
        published[u] = false;
        cur[u].left = budget;
        if (idx[u] >= num_in) { phase[u] = detail::kDone; return; }
        if (!core.ready()) { st[u] = Status::not_initialized(); phase[u] = detail::kEmit; return; }
        if (!core.is_valid_key(key[u])) { st[u] = Status::invalid_key(); join(st[u]); phase[u] = detail::kEmit; return; }
        cur[u].begin(key[u], core.hash_seed(), geo);
        cur[u].left = budget;
        phase[u] = detail::kLoad;
    };
#pragma unroll
    for (int u = 0; u < U; ++u) {
        idx[u] = tid + index_t(u) * lanes;
        nidx[u] = idx[u] + stride;
        const index_t a = idx[u] < num_in ? idx[u] : last;
        key[u] = keys_in[a];
        if constexpr (HasValues) val[u] = values_in[a];
        else val[u] = V{};
        nval[u] = V{};
        cur[u].w = 0;
        seen[u] = typename Core::insert_probe_type{};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) start(u);

    while (true) {
        bool busy = false;
#pragma unroll
        for (int u = 0; u < U; ++u) busy |= phase[u] != detail::kDone;
        if (!busy) break;
        uint4 win[U][W::chunks];
        typename Core::cas_word old[U];
        bool issue[U];
        // all window reads, next-input reads and CASes before any result is used
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint4* p = table + cur[u].w * W::chunks;
#pragma unroll
            for (unsigned c = 0; c < W::chunks; ++c) win[u][c] = p[c];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const index_t b = nidx[u] < num_in ? nidx[u] : last;
            nkey[u] = keys_in[b];
            if constexpr (HasValues) nval[u] = values_in[b];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            issue[u] = phase[u] == detail::kCas;
            old[u] = typename Core::cas_word{};
            if (issue[u]) {
                const unsigned c = unsigned(__builtin_ctzll(empty[u]));
                old[u] = core.cas_issue(slots + cur[u].w * CG + c, seen[u], key[u], val[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (phase[u] == detail::kLoad) {
                std::uint64_t hits = 0, empties = 0;
#pragma unroll
                for (unsigned c = 0; c < W::chunks; ++c) {
#pragma unroll
                    for (unsigned i = 0; i < W::per_chunk; ++i) {
                        const Slot s = detail::chunk_slot<Slot>(win[u][c], i);
                        const unsigned bit = c * W::per_chunk + i;
                        hits |= std::uint64_t(Core::key_of(s) == key[u]) << bit;
                        empties |= std::uint64_t(Core::is_empty_key(Core::key_of(s))) << bit;
                    }
                }
                if (hits) {
                    st[u] = Status::duplicate_key();
                    join(st[u]);
                    where[u] = cur[u].w * CG + unsigned(__builtin_ctzll(hits));
                    phase[u] = detail::kEmit;
                } else if (empties) {
                    empty[u] = empties;
                    seen[u] = Core::empty_probe();
                    phase[u] = detail::kCas;      // its CAS goes out next round
                } else if (!cur[u].advance(geo)) {
                    st[u] = Status::probing_length_exceeded();
                    join(st[u]);
                    phase[u] = detail::kEmit;
                }
            } else if (issue[u]) {
                const unsigned c = unsigned(__builtin_ctzll(empty[u]));
                const int outcome = core.cas_outcome(old[u], seen[u], key[u]);
                if (outcome == 2) {
                    st[u] = Status::duplicate_key();
                    join(st[u]);
                    where[u] = cur[u].w * CG + c;
                    phase[u] = detail::kEmit;
                } else if (outcome == 1) {
                    where[u] = cur[u].w * CG + c;
                    published[u] = Core::publishes_value;
                    phase[u] = detail::kEmit;
                } else if (outcome == 0) {
                    empty[u] &= empty[u] - 1;
                    if (empty[u]) {
                        seen[u] = Core::empty_probe();
                    } else if (cur[u].advance(geo)) {
                        phase[u] = detail::kLoad;
                    } else {
                        st[u] = Status::probing_length_exceeded();
                        join(st[u]);
                        phase[u] = detail::kEmit;
                    }
                }
                // outcome 3: the EMPTY slot's value bits moved, retry with seen refreshed
            }
            if (phase[u] == detail::kEmit) {
                if (where[u] != ~index_t(0)) core.finish_insert(where[u], val[u], published[u], st[u], join);
                StatusHandler::handle(st[u], status_out, idx[u]);
                idx[u] = nidx[u];
                key[u] = nkey[u];
                val[u] = nval[u];
                nidx[u] += stride;
                start(u);
            }
        }
    }
    join.flush(core.status_word());
}

}  // namespace kernels
}  // namespace warpcore
