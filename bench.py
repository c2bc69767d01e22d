#!/usr/bin/env python3
"""Hash-table hot-path benchmark on one MI355X.

Workload (the large single-value configuration): SingleValueHashTable
<u64, u64>, 2^28 unique keys at load factor 0.9 (min capacity 2^28 / 0.9 ->
298,876,376 slots, 4.78 GB of 16-byte slots), cooperative groups of 8 lanes
(one 128-byte line per probe window), then 2^28 lookups with a 50 % hit
rate. One step = init (EMPTY fill) + insert of the batch + retrieve of the
queries; every input is resident in HBM before timing starts.

value = (2^28 inserts + 2^28 lookups) per step / wall time per step, Gops/s.
Alongside: per-kernel times from HIP events recorded on the stream the
kernels run on, the roofline of the dominant kernel (algorithmic bytes: a
retrieve reads one 128-byte line plus key and value, an insert reads and
writes back one line plus key and value), the same workload for
SingleValueHashTable<u32, u32> (cg 16) and CountingHashTable<u64> over 2^28
Zipf(1.1) occurrences of 2^24 keys, the plain 2^20-key u32 table (MALL
resident, with the random-line rate at its own footprint), HashSet<u32>
(2^27 keys), the pattern-blocked BloomFilter (2^30 bits, k 8, 2^27 keys,
measured FPR), MultiValueHashTable<u32,u32> (2^26 pairs, two-pass
retrieve), the batched small tables, the serving loop, and the OpenMP CPU
table on a bounded sample as the reported CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import re
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "Gops/s insert+retrieve @LF0.9 u32/u64 on 1 GPU; served p99 vs launch-per-request"
LINE = 128          # bytes per probe window / cache line
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(op: str, kb: int, vb: int) -> int:
    if op == "retrieve":
        return LINE + kb // 8 + vb // 8
    if op == "insert":
        return 2 * LINE + kb // 8 + vb // 8
    if op == "count":
        return 2 * LINE + kb // 8
    raise ValueError(op)


# Revision of the probing kernels' code; a traffic figure measured on
# another revision is not used.
KERNEL_REVISION = "owner-4"


def read_traffic(family: str, log2n: int = 28, table: str = "u64"):
    """HBM bytes per launch of kernel `family` (insert_owner, retrieve_owner,
    ...) on this workload (2^28 keys; `table` "u64" for SVHT<u64,u64> cg 8,
    "u32" for SVHT<u32,u32> cg 16), measured by rocprofv3 --pmc (FETCH_SIZE
    + WRITE_SIZE, each scaled by a copy kernel of known size), if
    profiles/svht_u64_lf09_traffic.md holds a figure for that kernel at this
    KERNEL_REVISION; else None (a figure of another kernel is never used)."""
    f = REPO / "profiles" / "svht_u64_lf09_traffic.md"
    if log2n != 28 or not f.exists():
        return None
    key = family if table == "u64" else f"{table}\\.{family}"
    m = re.search(rf"^traffic_bytes_per_launch\.{key}@{KERNEL_REVISION}\s*=\s*([0-9.eE+]+)", f.read_text(),
                  re.M)
    return float(m.group(1)) if m else None


def kernel_family(lib, op: str) -> str:
    """The bulk kernel a SingleValueHashTable<u64,u64> (cg 8) call of `op`
    launches under the current kernel selection (wc_get_group_kernel)."""
    k = lib.wc_get_group_kernel()
    if op == "insert":
        return "insert_owner" if k & 128 else ("insert_refill" if k & 8 else
                                                ("insert_lockstep" if k & 1 else "insert_batched"))
    if k & 64:
        return "retrieve_owner"
    if k & 4:
        return "retrieve_refill_pipe" if k & 16 else "retrieve_refill"
    return "retrieve_lockstep" if k & 2 else "retrieve_batched"


def probes_lib():
    """ctypes handle of warpcore_amd/probes/libwarpcore_probes.so, prototypes
    from its header."""
    from warpcore_amd import _capi
    d = REPO / "warpcore_amd" / "probes"
    so = d / "libwarpcore_probes.so"
    if not so.exists():
        raise RuntimeError(f"{so} is missing: run the build first")
    h = ctypes.CDLL(str(so))
    for name, (restype, argtypes) in _capi.parse_header((d / "warpcore_probes.h").read_text()).items():
        fn = getattr(h, name)
        fn.restype, fn.argtypes = restype, argtypes
    return h


class Gpu:
    def __init__(self):
        import torch
        from warpcore_amd import _capi
        self.torch = torch
        self.lib = _capi.lib()
        self.probes = probes_lib()
        self.stream = torch.cuda.Stream()
        self.sp = ctypes.c_void_p(self.stream.cuda_stream)

    def empty(self, n, bits):
        t = {32: self.torch.int32, 64: self.torch.int64}[bits]
        return self.torch.empty(n, dtype=t, device="cuda")

    @staticmethod
    def p(t):
        return ctypes.c_void_p(t.data_ptr())

    def event(self):
        return self.torch.cuda.Event(enable_timing=True)

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} returned {rc}")


def run_svht(g: Gpu, kb: int, vb: int, cg: int, log2n: int, steps: int, warmup: int):
    torch = g.torch
    from warpcore_amd.tables import SingleValueHashTable
    n = 1 << log2n
    half = n // 2
    min_cap = math.ceil(n / 0.9)
    sfx = "u64" if kb == 64 else "u32"
    gen_unique = getattr(g.lib, f"wc_gen_unique_{sfx}")
    gen_uniform = getattr(g.lib, f"wc_gen_uniform_{'u64' if vb == 64 else 'u32'}")
    interleave = getattr(g.lib, f"wc_gen_interleave_{sfx}")
    keys, values = g.empty(n, kb), g.empty(n, vb)
    absent, queries, out = g.empty(half, kb), g.empty(n, kb), g.empty(n, vb)
    with torch.cuda.stream(g.stream):
        g.check(gen_unique(g.p(keys), n, 0x1000, g.sp), "gen keys")
        g.check(gen_uniform(g.p(values), n, 0xC1 + 1, g.sp), "gen values")
        g.check(gen_unique(g.p(absent), half, 0x1000 + n + 2, g.sp), "gen absent")
        g.check(interleave(g.p(queries), g.p(keys), g.p(absent), half, 0xC1 + 2, g.sp), "mix")
    table = SingleValueHashTable(min_cap, kb, vb, cg_size=cg)
    g.stream.synchronize()

    def step(ev=None):
        table.init(stream=g.stream)
        if ev:
            ev[0].record(g.stream)
        table.insert(keys, values, stream=g.stream)
        if ev:
            ev[1].record(g.stream)
        table.retrieve(queries, out, stream=g.stream)
        if ev:
            ev[2].record(g.stream)

    for _ in range(warmup):
        step()
    g.stream.synchronize()
    evs = [[g.event() for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        step(evs[i])
    g.stream.synchronize()
    wall = time.perf_counter() - t0
    t_ins = [e[0].elapsed_time(e[1]) for e in evs]
    t_ret = [e[1].elapsed_time(e[2]) for e in evs]

    # result check of the last step: every inserted key stored, half the lookups hit
    st = g.empty(n, 32)
    out.fill_(-1)
    table.retrieve(queries, out, stream=g.stream, status_out=st)
    g.stream.synchronize()
    hits = int((st == 0).sum().item())
    size = table.size(stream=g.stream)
    status = int(table.peek_status(stream=g.stream))
    ok = hits == half and size == n and status == 0x20  # key_not_found joined by the misses
    if not ok:
        raise RuntimeError(f"result check failed: hits={hits} size={size} status={status:#x}")
    res = dict(n=n, capacity=table.capacity(), cg=cg, wall_ms_per_step=1e3 * wall / steps,
               insert_ms=sum(t_ins) / steps, retrieve_ms=sum(t_ret) / steps,
               insert_ms_min=min(t_ins), retrieve_ms_min=min(t_ret), table_bytes=table.bytes_total())
    del table, keys, values, absent, queries, out, st
    torch.cuda.empty_cache()
    return res


def run_counting(g: Gpu, log2n: int, log2_distinct: int, s: float, steps: int, warmup: int):
    torch = g.torch
    from warpcore_amd.tables import CountingHashTable
    n, d = 1 << log2n, 1 << log2_distinct
    keys = g.empty(n, 64)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_zipf_u64(g.p(keys), n, d, s, 0xC3, 0x2000, g.sp), "gen zipf")
    table = CountingHashTable(math.ceil(d / 0.9), 64)
    g.stream.synchronize()
    for _ in range(warmup):
        table.init(stream=g.stream)
        table.insert(keys, stream=g.stream)
    g.stream.synchronize()
    evs = [[g.event(), g.event()] for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        table.init(stream=g.stream)
        evs[i][0].record(g.stream)
        table.insert(keys, stream=g.stream)
        evs[i][1].record(g.stream)
    g.stream.synchronize()
    wall = time.perf_counter() - t0
    t = [e[0].elapsed_time(e[1]) for e in evs]
    distinct = table.size(stream=g.stream)
    res = dict(n=n, distinct=distinct, capacity=table.capacity(), wall_ms_per_step=1e3 * wall / steps,
               insert_ms=sum(t) / steps, insert_ms_min=min(t))
    del table, keys
    torch.cuda.empty_cache()
    return res


def run_batched(g: Gpu, num_tables: int = 65536, steps: int = 3, warmup: int = 1, tune: bool = True,
                sample_tables: int = 256, seed: int = 0xC7):
    """65,536 SingleValueHashTable<u32,u32> of capacity log-uniform in
    2^6..2^16 slots, each filled to a load factor uniform in [0.5, 0.9]:
    insert of every table's keys, then retrieve of them, one persistent
    launch each, after the autotuner has chosen (lanes, window) per
    (capacity, fill) class and the launch. Beside it, plain per-table
    launches on a sample."""
    import numpy as np
    torch = g.torch
    from warpcore_amd.tables import SingleValueHashTable, TableBatch
    rng = np.random.default_rng(seed)
    caps = np.exp2(rng.uniform(6, 16, num_tables)).astype(np.uint64)
    n_t = np.maximum((caps * rng.uniform(0.5, 0.9, num_tables)).astype(np.uint64), 1)
    off = np.concatenate([[0], np.cumsum(n_t)]).astype(np.uint64)
    n = int(off[-1])
    keys, values, out = g.empty(n, 32), g.empty(n, 32), g.empty(n, 32)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_unique_u32(g.p(keys), n, 0x7000, g.sp), "gen keys")
        g.check(g.lib.wc_gen_uniform_u32(g.p(values), n, seed, g.sp), "gen values")
    offs = torch.from_numpy(off.view(np.int64)).to("cuda")
    batch = TableBatch(caps, 32, 32, cg_size=16)
    g.stream.synchronize()
    launch, tuned_ms, classes, totals = None, None, None, None
    if tune:
        t0 = time.perf_counter()
        launch, tuned_ms = batch.tune(keys, values, offs, keys, offs, out, stream=g.stream)
        tune_s = time.perf_counter() - t0
        rows, totals = batch.tune_report()
        classes = [{"capacity": f"2^{2 * r['capacity_band']}..2^{2 * r['capacity_band'] + 2}",
                    "fill": ["<0.6", "0.6-0.7", "0.7-0.8", ">=0.8"][r["fill_band"]],
                    "tables": r["tables"], "keys": r["keys"], "chosen": list(r["chosen_shape"]),
                    "fixed_ms": round(r["fixed_ns"] / 1e6, 4), "chosen_ms": round(r["chosen_ns"] / 1e6, 4),
                    "gain": round(r["fixed_ns"] / max(r["chosen_ns"], 1), 3),
                    "slots_ratio": round(r["chosen_slots"] / max(r["fixed_slots"], 1), 3)} for r in rows]
    for _ in range(warmup):
        batch.init(stream=g.stream)
        batch.insert(keys, values, offs, stream=g.stream)
        batch.retrieve(keys, offs, out, stream=g.stream)
    g.stream.synchronize()
    ev = [[g.event() for _ in range(3)] for _ in range(steps)]
    for i in range(steps):
        batch.init(stream=g.stream)
        ev[i][0].record(g.stream)
        batch.insert(keys, values, offs, stream=g.stream)
        ev[i][1].record(g.stream)
        batch.retrieve(keys, offs, out, stream=g.stream)
        ev[i][2].record(g.stream)
    g.stream.synchronize()
    ins = min(e[0].elapsed_time(e[1]) for e in ev)
    ret = min(e[1].elapsed_time(e[2]) for e in ev)
    ok = bool(torch.equal(out, values))
    # plain launches: one SingleValueHashTable per table for a sample of tables
    pick = np.linspace(0, num_tables - 1, sample_tables).astype(np.int64)
    tables = [SingleValueHashTable(int(caps[t]), 32, 32, cg_size=16) for t in pick]
    # views cut before timing, so the timed loop is the launches themselves
    parts = [(keys[int(off[t]):int(off[t + 1])], values[int(off[t]):int(off[t + 1])], out[int(off[t]):int(off[t + 1])])
             for t in pick]
    sample_keys = sum(int(off[t + 1] - off[t]) for t in pick)
    g.stream.synchronize()
    e0, e1 = g.event(), g.event()
    e0.record(g.stream)
    for tab, (k, v, o) in zip(tables, parts):
        tab.insert(k, v, stream=g.stream)
        tab.retrieve(k, o, stream=g.stream)
    e1.record(g.stream)
    g.stream.synchronize()
    plain_ms = e0.elapsed_time(e1)
    res = {"tables": num_tables, "keys": n, "total_slots": batch.total_slots(),
           "gops": round(2 * n / (ins + ret) / 1e6, 3), "insert_gops": round(n / ins / 1e6, 3),
           "retrieve_gops": round(n / ret / 1e6, 3), "results_ok": ok, "plain_sample_tables": sample_tables,
           "plain_per_table_gops": round(2 * sample_keys / plain_ms / 1e6, 3)}
    if tune:
        res["tuner"] = {"launch": {"item_keys": launch[0], "blocks_per_cu": launch[1], "block_size": launch[2]},
                        "fixed": "(16,16) every table, launch 2048 keys/item, 8 blocks/CU of 256",
                        "fixed_gops": round(2 * n / totals[0] / 1e6, 3),
                        "tuned_gops_in_tune": round(2 * n / totals[1] / 1e6, 3),
                        "tune_seconds": round(tune_s, 2), "classes": classes}
    del batch, tables, keys, values, out, offs
    torch.cuda.empty_cache()
    return res


def run_serving(g: Gpu, num_tables: int = 64, total_bytes: float = 100e9, num_requests: int = 2000,
                load: float = 0.5, seed: int = 0xC8, server_kw=None):
    """Open-loop request stream against 64 resident SingleValueHashTable
    <u64,u64> (about 100 GB of slots): sizes log-uniform 2^8..2^22 keys,
    30 % insert / 60 % retrieve / 10 % erase, uniform u64 keys with 20 %
    Zipf(1.0) hot keys. Phase 1: every request at t=0 (throughput). Phase 2:
    Poisson arrivals at `load` x the one-launch-per-request throughput
    (p50/p99 latency). Both through the server and one launch per request."""
    import numpy as np
    torch = g.torch
    from warpcore_amd.tables import Server
    rng = np.random.default_rng(seed)
    caps = np.full(num_tables, int(total_bytes / 16 / num_tables), np.uint64)
    sizes = np.exp2(rng.uniform(8, 22, num_requests)).astype(np.uint64)
    ops = rng.choice([0, 1, 2], size=num_requests, p=[0.3, 0.6, 0.1]).astype(np.uint32)
    tabs = rng.integers(0, num_tables, size=num_requests).astype(np.uint32)
    n = int(sizes.sum())
    keys, values, out, st = g.empty(n, 64), g.empty(n, 64), g.empty(n, 64), g.empty(n, 32)
    hot = g.empty(n, 64)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_uniform_u64(g.p(keys), n, seed, g.sp), "gen keys")
        g.check(g.lib.wc_gen_zipf_u64(g.p(hot), n, 1 << 20, 1.0, seed + 1, 0x9000, g.sp), "gen hot")
        g.check(g.lib.wc_gen_uniform_u64(g.p(values), n, seed + 2, g.sp), "gen values")
        take = torch.rand(n, device="cuda") < 0.2
        keys[take] = hot[take]
        # 0 and ~0 are not keys
        keys[(keys == 0) | (keys == -1)] = 0x1234567
    g.stream.synchronize()
    del hot, take
    res = {"tables": num_tables, "table_slots": int(caps[0]), "requests": num_requests, "keys": n}

    def one(mode, arrival, count=num_requests, pick=None):
        srv = Server(caps, 64, 64, **(server_kw or {}))
        sel = slice(0, count) if pick is None else pick
        lat, mk = srv.drive(mode, ops[sel], tabs[sel], sizes[sel], arrival[sel], keys, values, out, st)
        stats = srv.stats()
        steps = (stats["steps"], stats["direct_steps"]) if mode == 0 else (count, 0)
        del srv
        torch.cuda.empty_cache()
        return lat, mk, steps

    zero = np.zeros(num_requests)
    # warm-up of both paths (table allocation, clocks) before anything is timed
    one(1, zero, 300)
    one(0, zero, 300)
    lat_b, mk_b, _ = one(1, zero)
    lat_s, mk_s, steps_s = one(0, zero)
    res["throughput"] = {"server_gkeys_s": round(n / mk_s / 1e3, 3), "launch_per_request_gkeys_s":
                         round(n / mk_b / 1e3, 3), "speedup": round(mk_b / mk_s, 3), "server_steps": steps_s[0],
                         "server_direct_steps": steps_s[1]}
    # throughput per request-size class: the class's requests alone, all
    # at t=0, through the server and one launch per request
    per_class = []
    for lo, hi in [(8, 12), (12, 16), (16, 19), (19, 22.01)]:
        pick = np.flatnonzero((sizes >= 2 ** lo) & (sizes < 2 ** hi))
        if len(pick) == 0:
            continue
        kc = int(sizes[pick].sum())
        _, mk_cb, _ = one(1, zero, pick=pick)
        _, mk_cs, _ = one(0, zero, pick=pick)
        per_class.append({"sizes": f"2^{lo}..2^{int(hi)}", "requests": int(len(pick)), "keys": kc,
                          "key_share": round(kc / n, 4), "server_gkeys_s": round(kc / mk_cs / 1e3, 3),
                          "launch_per_request_gkeys_s": round(kc / mk_cb / 1e3, 3),
                          "speedup": round(mk_cb / mk_cs, 2)})
    res["throughput"]["per_size_class"] = per_class
    rate = load * num_requests / mk_b          # requests per microsecond
    arrival = np.cumsum(rng.exponential(1.0 / rate, size=num_requests))
    lat_b, mk_b2, _ = one(1, arrival)
    lat_s, mk_s2, steps_s2 = one(0, arrival)
    pct = lambda a, q: round(float(np.percentile(a, q)), 1)
    res["latency"] = {"offered_load": load, "offered_gkeys_s": round(n / arrival[-1] / 1e3, 3),
                      "server_p50_us": pct(lat_s, 50), "server_p99_us": pct(lat_s, 99),
                      "launch_per_request_p50_us": pct(lat_b, 50), "launch_per_request_p99_us": pct(lat_b, 99),
                      "server_steps": steps_s2[0], "server_direct_steps": steps_s2[1]}
    del keys, values, out, st
    torch.cuda.empty_cache()
    return res


def run_line_roofline(g: Gpu, table_bytes: int, lines: int = 1 << 28, reps: int = 3, claim: bool = False):
    """Rate at which random 128-byte lines of a table_bytes footprint are
    served (wc_util_gather_lines: 8 lanes per line, 8 lines in flight per
    group), the practical ceiling of the probing kernels at that size; with
    `claim`, also the rate of line reads each followed by a dependent
    compare-and-swap on one of its words (wc_util_claim_lines), the
    ceiling of one-window inserts."""
    torch = g.torch
    buf = torch.empty(table_bytes // 16 * 2, dtype=torch.int64, device="cuda")
    buf.fill_(7)
    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    g.stream.wait_stream(torch.cuda.current_stream())
    def best_of(fn, what):
        best = float("inf")
        for r in range(reps + 1):
            e0, e1 = g.event(), g.event()
            e0.record(g.stream)
            g.check(fn(g.p(buf), table_bytes, lines, 0x51 + r, g.p(sink), g.sp), what)
            e1.record(g.stream)
            g.stream.synchronize()
            if r:
                best = min(best, e0.elapsed_time(e1))
        return best

    best = best_of(g.probes.wc_util_gather_lines, "gather")
    res = {"glines_s": round(lines / (best * 1e-3) / 1e9, 2), "gbs": round(lines / (best * 1e-3) / 1e9 * LINE, 1),
           "footprint_bytes": table_bytes, "ms": round(best, 3)}
    if claim:   # each line read followed by a dependent CAS on one of its words: the insert's pattern
        cbest = best_of(g.probes.wc_util_claim_lines, "claim")
        res["claim_glines_s"] = round(lines / (cbest * 1e-3) / 1e9, 2)
    del buf, sink
    torch.cuda.empty_cache()
    return res


def _timed(g: Gpu, fn, reps: int, before=None):
    """Min and mean over reps of fn() between HIP events on g.stream, after one
    untimed call; before() (untimed, e.g. a table init) runs ahead of each."""
    times = []
    for r in range(reps + 1):
        if before:
            before()
        e0, e1 = g.event(), g.event()
        e0.record(g.stream)
        fn()
        e1.record(g.stream)
        g.stream.synchronize()
        if r:
            times.append(e0.elapsed_time(e1))
    return min(times), sum(times) / len(times)


def _rate(n, ms, bytes_per_op):
    """Gops/s and algorithmic GB/s of n ops in ms milliseconds."""
    return {"gops": round(n / ms / 1e6, 3), "ms": round(ms, 4), "bytes_per_op": bytes_per_op,
            "gbs": round(n * bytes_per_op / ms / 1e6, 1)}


def run_small_svht(g: Gpu, load: float, reps: int = 20):
    """The plain configuration: SingleValueHashTable<u32,u32>, 2^20 unique keys,
    min capacity ceil(2^20 / load), lookups of all 2^20 keys plus 2^20 absent
    keys in hashed order. The table (16.8 MB at load 0.8) stays in the MALL,
    so the line rate at its own footprint is reported beside it."""
    torch = g.torch
    from warpcore_amd.tables import SingleValueHashTable
    n = 1 << 20
    keys, vals, absent = g.empty(n, 32), g.empty(n, 32), g.empty(n, 32)
    queries, out = g.empty(2 * n, 32), g.empty(2 * n, 32)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_unique_u32(g.p(keys), n, 0xC1, g.sp), "gen keys")
        g.check(g.lib.wc_gen_uniform_u32(g.p(vals), n, 0xC1 + 1, g.sp), "gen values")
        g.check(g.lib.wc_gen_unique_u32(g.p(absent), n, 0xC1 + n + 3, g.sp), "gen absent")
        g.check(g.lib.wc_gen_interleave_u32(g.p(queries), g.p(keys), g.p(absent), n, 0xC1 + 2, g.sp), "mix")
    t = SingleValueHashTable(math.ceil(n / load), 32, 32)
    g.stream.synchronize()
    ins, _ = _timed(g, lambda: t.insert(keys, vals, stream=g.stream), reps, before=lambda: t.init(stream=g.stream))
    ret, _ = _timed(g, lambda: t.retrieve(queries, out, stream=g.stream), reps)
    st = g.empty(2 * n, 32)
    t.retrieve(queries, out, stream=g.stream, status_out=st)
    g.stream.synchronize()
    if int((st == 0).sum()) != n or t.size(stream=g.stream) != n:
        raise RuntimeError("small-table result check failed")
    lr = run_line_roofline(g, t.bytes_total(), lines=1 << 26, claim=True)
    res = {"capacity": t.capacity(), "table_bytes": t.bytes_total(), "cg": t.cg_size(),
           "insert": _rate(n, ins, algorithmic_bytes("insert", 32, 32)),
           "retrieve": _rate(2 * n, ret, algorithmic_bytes("retrieve", 32, 32)),
           "line_rate_at_footprint_glines_s": lr["glines_s"],
           "claim_rate_at_footprint_glines_s": lr["claim_glines_s"]}
    del t, keys, vals, absent, queries, out, st
    torch.cuda.empty_cache()
    return res


def run_hashset(g: Gpu, log2n: int = 27, reps: int = 3):
    """HashSet<u32>: 2^27 keys at load factor 0.9 (151,011,248 slots, 604 MB),
    insert, then contains of all keys and of 2^27 absent keys."""
    torch = g.torch
    from warpcore_amd.tables import HashSet
    n = 1 << log2n
    keys, absent = g.empty(n, 32), g.empty(n, 32)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_unique_u32(g.p(keys), n, 0x3000, g.sp), "gen keys")
        g.check(g.lib.wc_gen_unique_u32(g.p(absent), n, 0x3000 + n + 5, g.sp), "gen absent")
    s = HashSet(math.ceil(n / 0.9), 32)
    flags = torch.empty(n, dtype=torch.uint8, device="cuda")
    g.stream.synchronize()
    ins, _ = _timed(g, lambda: s.insert(keys, stream=g.stream), reps, before=lambda: s.init(stream=g.stream))
    hit, _ = _timed(g, lambda: s.contains(keys, flags, stream=g.stream), reps)
    ok = int(flags.sum()) == n
    miss, _ = _timed(g, lambda: s.contains(absent, flags, stream=g.stream), reps)
    ok = ok and int(flags.sum()) == 0 and s.size(stream=g.stream) == n
    if not ok:
        raise RuntimeError("hash set result check failed")
    res = {"capacity": s.capacity(), "cg": s.cg_size(), "insert": _rate(n, ins, 2 * LINE + 4),
           "contains_hits": _rate(n, hit, LINE + 4 + 1), "contains_misses": _rate(n, miss, LINE + 4 + 1)}
    del s, keys, absent, flags
    torch.cuda.empty_cache()
    return res


def run_bloom(g: Gpu, log2_bits: int = 30, k: int = 8, log2n: int = 27, reps: int = 3):
    """Pattern-blocked BloomFilter: 2^30 bits (128 MB), k = 8, 2^27 u64 inserts,
    2^27 queries of which half are absent; measured false-positive rate."""
    torch = g.torch
    from warpcore_amd.tables import BloomFilter
    n, half = 1 << log2n, 1 << (log2n - 1)
    keys, absent, queries = g.empty(n, 64), g.empty(half, 64), g.empty(n, 64)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_unique_u64(g.p(keys), n, 0x4000, g.sp), "gen keys")
        g.check(g.lib.wc_gen_unique_u64(g.p(absent), half, 0x4000 + n + 9, g.sp), "gen absent")
        g.check(g.lib.wc_gen_interleave_u64(g.p(queries), g.p(keys), g.p(absent), half, 0x4001, g.sp), "mix")
    bf = BloomFilter(1 << log2_bits, k, 64)
    flags = torch.empty(n, dtype=torch.uint8, device="cuda")
    g.stream.synchronize()
    ins, _ = _timed(g, lambda: bf.insert(keys, stream=g.stream), reps, before=lambda: bf.init(stream=g.stream))
    qry, _ = _timed(g, lambda: bf.retrieve(queries, flags, stream=g.stream), reps)
    bf.retrieve(keys, flags, stream=g.stream)
    g.stream.synchronize()
    if int(flags.sum()) != n:
        raise RuntimeError("bloom filter returned a false negative")
    bf.retrieve(absent, flags[:half], stream=g.stream)
    g.stream.synchronize()
    fpr = float(flags[:half].float().mean())
    res = {"bytes": bf.bytes_total(), "cg": bf.cg_size(), "insert": _rate(n, ins, 2 * LINE + 8),
           "query": _rate(n, qry, LINE + 8 + 1), "fpr_measured": round(fpr, 5), "fpr_model": round(bf.fpr(n), 5)}
    del bf, keys, absent, queries, flags
    torch.cuda.empty_cache()
    return res


def run_bucketlist(g: Gpu, log2n: int = 27, mult: int = 8, reps: int = 3):
    """BucketListHashTable<u32,u32> of the reference's bucket-list benchmark
    shape: 2^27 values on 2^24 keys (each key 8 times, shuffled), key
    capacity 2^24 / 0.9, value capacity 2^27 / 0.5, buckets growing by 1.1
    from 1 value, cg 8; insert, then the two-pass retrieve of every key."""
    torch = g.torch
    from warpcore_amd.tables import BucketListHashTable
    n = 1 << log2n
    d = n // mult
    ukeys = g.empty(d, 32)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_unique_u32(g.p(ukeys), d, 0x6000, g.sp), "gen keys")
    g.stream.synchronize()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0xC6)
    pk = ukeys.repeat(mult)[torch.randperm(n, device="cuda", generator=gen)].contiguous()
    pv = g.empty(n, 32)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_uniform_u32(g.p(pv), n, 0xC7, g.sp), "gen values")
    torch.cuda.synchronize()
    t = BucketListHashTable(math.ceil(d / 0.9), math.ceil(n / 0.5), 32, 32, grow_factor=1.1, min_bucket_size=1,
                            cg_size=8)
    g.stream.synchronize()
    ins, _ = _timed(g, lambda: t.insert(pk, pv, stream=g.stream), reps, before=lambda: t.init(stream=g.stream))
    beg = torch.empty(d, dtype=torch.int64, device="cuda")
    end = torch.empty(d, dtype=torch.int64, device="cuda")
    vout = g.empty(n, 32)
    got = []
    ret, _ = _timed(g, lambda: got.append(t.retrieve(ukeys, beg, end, vout, stream=g.stream)), reps)
    if got[-1] != n:
        raise RuntimeError(f"bucket-list result check failed: {got[-1]} values of {n}")
    res = {"keys": d, "values": n, "insert_gops": round(n / ins / 1e6, 3), "insert_ms": round(ins, 4),
           "retrieve_values_gops": round(n / ret / 1e6, 3), "retrieve_ms": round(ret, 4)}
    del t, ukeys, pk, pv, beg, end, vout
    torch.cuda.empty_cache()
    return res


def run_mvht(g: Gpu, log2n: int = 26, mean: int = 8, reps: int = 3):
    """MultiValueHashTable<u32,u32>: 2^26 pairs, per-key multiplicity geometric
    with mean 8 (about 2^23 keys), load factor 0.9 (75,505,624 slots); insert,
    then the two-pass retrieve of every key (count, exclusive scan, gather)."""
    import numpy as np
    torch = g.torch
    from warpcore_amd.tables import MultiValueHashTable
    n = 1 << log2n
    rng = np.random.default_rng(0xC4)
    mult = rng.geometric(1 / mean, size=n // mean + 4096)
    mult = mult[: np.searchsorted(np.cumsum(mult), n) + 1]
    d = len(mult)
    ukeys = g.empty(d, 32)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_unique_u32(g.p(ukeys), d, 0x5000, g.sp), "gen keys")
    g.stream.synchronize()
    m = torch.from_numpy(mult).to("cuda")
    pk = torch.repeat_interleave(ukeys, m)[:n]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0xC4)
    pk = pk[torch.randperm(n, device="cuda", generator=gen)].contiguous()
    pv = g.empty(n, 32)
    with torch.cuda.stream(g.stream):
        g.check(g.lib.wc_gen_uniform_u32(g.p(pv), n, 0xC5, g.sp), "gen values")
    torch.cuda.synchronize()
    t = MultiValueHashTable(math.ceil(n / 0.9), 32, 32)
    g.stream.synchronize()
    ins, _ = _timed(g, lambda: t.insert(pk, pv, stream=g.stream), reps, before=lambda: t.init(stream=g.stream))
    nk = t.num_keys(stream=g.stream)
    beg = torch.empty(nk, dtype=torch.int64, device="cuda")
    end = torch.empty(nk, dtype=torch.int64, device="cuda")
    vout = g.empty(n, 32)
    got = []
    ret, _ = _timed(g, lambda: got.append(t.retrieve(ukeys[:nk], beg, end, vout, stream=g.stream)), reps)
    if nk != d or got[-1] != n:
        raise RuntimeError(f"multi-value result check failed: keys {nk} of {d}, values {got[-1]} of {n}")
    # two probe passes per key, offsets and scan (16 B per key), one value per pair out
    per_key = 2 * LINE + 16 + 4
    res = {"capacity": t.capacity(), "cg": t.cg_size(), "keys": nk, "pairs": n,
           "insert": _rate(n, ins, 2 * LINE + 8),
           "retrieve_pairs_gops": round(n / ret / 1e6, 3), "retrieve_ms": round(ret, 4),
           "retrieve_gbs": round((nk * per_key + n * 4) / ret / 1e6, 1)}
    del t, ukeys, m, pk, pv, beg, end, vout
    torch.cuda.empty_cache()
    return res


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_pass(oracle, kb: int, n: int, min_cap: int, cg: int, hit_frac: float, seed: int):
    """One init + insert + retrieve pass of the OpenMP table: n distinct keys
    (fmix of a counter, sentinels removed), lookups of n keys of which
    hit_frac are stored. Returns (insert s, retrieve s, capacity)."""
    import numpy as np
    lib = oracle.lib()
    dt = np.uint64 if kb == 64 else np.uint32
    absent_n = int(n * (1 - hit_frac)) if hit_frac < 1 else n
    idx = np.arange(0x1000, 0x1000 + n + absent_n + n // 64 + 1024, dtype=np.uint64)
    allk = np.empty_like(idx)
    lib.wco_fmix64_array(oracle._p(idx), oracle._p(allk), len(idx))
    if kb == 32:
        allk = np.unique(allk.astype(np.uint32))   # distinct after the cut to 32 bits
        np.random.default_rng(seed).shuffle(allk)
    allk = allk[(allk != 0) & (allk != dt(np.iinfo(dt).max))].astype(dt)
    keys, absent = allk[:n], allk[n:n + absent_n]
    vals = np.random.default_rng(seed).integers(0, np.iinfo(dt).max, n, dtype=dt)
    if hit_frac < 1:
        hits = keys[: n - absent_n]
        queries = np.empty(len(hits) + len(absent), dt)
        queries[0::2], queries[1::2] = hits, absent
    else:
        queries = np.concatenate([keys, absent])
    out = np.zeros(len(queries), dt)
    t = oracle.OmpTable(min_cap, kb, kb, cg=cg)
    t.init()
    ti, _ = t.insert(keys, vals, status=False)
    tr, _ = t.retrieve(queries, out, status=False)
    return ti, tr, len(queries), t.capacity()


def cpu_baseline():
    """The OpenMP table (same hashing, probing, CAS protocol and statuses) on
    the host cores at the configurations' full sizes: SVHT<u64,u64> with 2^28
    keys at LF 0.9 and 2^28 lookups at 50 % hits (the metric's workload), and
    SVHT<u32,u32> with 2^20 keys at capacity 2^20/0.8 and lookups of the 2^20
    keys plus 2^20 absent ones. One pass each (about 10-20 s of CPU work)."""
    import oracle
    n2 = 1 << 28
    ti, tr, q2, cap2 = _omp_pass(oracle, 64, n2, math.ceil(n2 / 0.9), 8, 0.5, 0xC2)
    n1 = 1 << 20
    ti1, tr1, q1, cap1 = _omp_pass(oracle, 32, n1, math.ceil(n1 / 0.8), 8, 1.0, 0xC1)
    threads = oracle.omp_threads()
    return dict(value=round((n2 + q2) / (ti + tr) / 1e9, 4), unit="Gops/s", cores=threads, kind="port",
                cpu=_cpu_model(), nproc=os.cpu_count(),
                insert_gops=round(n2 / ti / 1e9, 4), retrieve_gops=round(q2 / tr / 1e9, 4),
                c1_u32u32_2p20=dict(gops=round((n1 + q1) / (ti1 + tr1) / 1e9, 4), capacity=cap1,
                                    insert_gops=round(n1 / ti1 / 1e9, 4), retrieve_gops=round(q1 / tr1 / 1e9, 4)),
                sample=f"OpenMP u64/u64 table ({threads} threads; fmix64 double hashing over cg-8 line windows, "
                       f"std::atomic CAS): the full workload, 2^28 unique keys at LF 0.9 ({cap2} slots) then "
                       f"2^28 lookups at 50% hits, one pass; c1: u32/u32, 2^20 keys, capacity 2^20/0.8, "
                       f"lookups of 2^20 stored + 2^20 absent keys")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2-keys", type=int, default=28)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip every workload but the headline u64 table")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched small tables")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving loop")
    ap.add_argument("--no-structures", action="store_true",
                    help="skip the 2^20-key u32 table, hash set, bloom filter and multi-value table")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (PMC calibration)")
    only = {"hashset": run_hashset, "bloom": run_bloom, "mvht": run_mvht, "bucketlist": run_bucketlist,
            "batched": run_batched, "serving": run_serving,
            "counting": lambda g: run_counting(g, 28, 24, 1.1, 5, 1)}
    ap.add_argument("--only", choices=sorted(only),
                    help="run this one structure's workload alone and print its result as JSON "
                         "(for profiling runs; no benchmark line)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: the hash-table hot path has no multi-GPU "
              "path (tables would shard by key hash, one process per GPU).", file=sys.stderr)
        sys.exit(2)

    g = Gpu()
    if args.only:
        print(json.dumps({args.only: only[args.only](g)}), flush=True)
        return
    log(f"[bench] svht u64/u64 2^{args.log2_keys} keys, steps={args.steps} warmup={args.warmup}")
    r64 = run_svht(g, 64, 64, 8, args.log2_keys, args.steps, args.warmup)
    log(f"[bench] u64: {r64}")
    extras = {}
    if not args.no_extras:
        r32 = run_svht(g, 32, 32, 16, args.log2_keys, args.steps, args.warmup)
        log(f"[bench] u32: {r32}")
        rc = run_counting(g, args.log2_keys, 24, 1.1, args.steps, args.warmup)
        log(f"[bench] counting: {rc}")
        extras = {
            "u32u32_cg16": {"gops": round(2 * r32["n"] / r32["wall_ms_per_step"] / 1e6, 3),
                            "insert_gops": round(r32["n"] / r32["insert_ms"] / 1e6, 3),
                            "retrieve_gops": round(r32["n"] / r32["retrieve_ms"] / 1e6, 3),
                            "capacity": r32["capacity"]},
            "counting_u64_zipf1.1": {"gops": round(rc["n"] / rc["wall_ms_per_step"] / 1e6, 3),
                                     "insert_gops": round(rc["n"] / rc["insert_ms"] / 1e6, 3),
                                     "occurrences": rc["n"], "distinct": rc["distinct"],
                                     "capacity": rc["capacity"]},
        }
        if not args.no_batched:
            log("[bench] batched tables ...")
            extras["batched_65536_u32u32"] = run_batched(g)
            log(f"[bench] batched: {extras['batched_65536_u32u32']}")
        if not args.no_structures:
            log("[bench] plain u32 table, hash set, bloom filter, multi-value table ...")
            extras["svht_u32u32_2p20"] = {f"lf{lf}": run_small_svht(g, lf) for lf in (0.8, 0.9)}
            extras["hashset_u32_2p27"] = run_hashset(g)
            extras["bloom_2p30bits_k8"] = run_bloom(g)
            extras["mvht_u32u32_2p26"] = run_mvht(g)
            extras["bucketlist_u32u32_2p27"] = run_bucketlist(g)
            for k in ("svht_u32u32_2p20", "hashset_u32_2p27", "bloom_2p30bits_k8", "mvht_u32u32_2p26",
                      "bucketlist_u32u32_2p27"):
                log(f"[bench] {k}: {extras[k]}")
        if not args.no_serving:
            log("[bench] serving ...")
            extras["serving_64x1.56GB_u64u64"] = run_serving(g)
            log(f"[bench] serving: {extras['serving_64x1.56GB_u64u64']}")
    if args.calibrate_copy_mib:
        torch = g.torch
        # wc_util_copy: 16 bytes per lane and access, each byte read and written once
        a = torch.empty(args.calibrate_copy_mib << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        g.check(g.probes.wc_util_copy(g.p(b), g.p(a), a.numel(), None), "calibration copy")
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibrate_copy_mib} MiB done")

    n = r64["n"]
    per_op = {"insert": algorithmic_bytes("insert", 64, 64), "retrieve": algorithmic_bytes("retrieve", 64, 64)}
    kt = {"insert": r64["insert_ms"], "retrieve": r64["retrieve_ms"]}
    dom = max(kt, key=kt.get)
    achieved = per_op[dom] * n / (kt[dom] * 1e-3) / 1e9
    other = "retrieve" if dom == "insert" else "insert"
    achieved_other = per_op[other] * n / (kt[other] * 1e-3) / 1e9
    fam = {op: kernel_family(g.lib, op) for op in ("insert", "retrieve")}
    roofline = {"bound": "hbm", "kernel": f"{fam[dom]} [SVHT u64/u64 cg 8]", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(fam[dom], args.log2_keys),
                "traffic_source": f"profiles/svht_u64_lf09_traffic.md @{KERNEL_REVISION}",
                "bytes_per_op": per_op[dom], "ops_per_launch": n, "avg_kernel_ms": round(kt[dom], 4),
                "other_kernel": {"kernel": f"{fam[other]} [SVHT u64/u64 cg 8]", "achieved": round(achieved_other, 1),
                                 "frac": round(achieved_other / HBM_PEAK_GBS, 4), "bytes_per_op": per_op[other],
                                 "avg_kernel_ms": round(kt[other], 4),
                                 "traffic": read_traffic(fam[other], args.log2_keys)}}
    # practical ceiling for random whole lines at the headline table's size
    lr = run_line_roofline(g, r64["table_bytes"], claim=True)
    # inserts against the read-then-claim rate (one window and one CAS per key at best)
    lr["insert_vs_claim"] = round(n / (kt["insert"] * 1e-3) / 1e9 / lr["claim_glines_s"], 4)
    for op in ("retrieve", "insert"):
        tr = read_traffic(fam[op], args.log2_keys)
        if tr:   # measured HBM bytes per op in 128-byte line equivalents (reads and write-backs)
            lines_per_op = tr / LINE / n
            lr[f"{op}_lines_per_op"] = round(lines_per_op, 3)
            lr[f"{op}_frac"] = round(n / (kt[op] * 1e-3) * lines_per_op / 1e9 / lr["glines_s"], 4)
    log(f"[bench] line roofline: {lr}")
    if "u32u32_cg16" in extras:   # the same for the u32 table (2.4 GB of 8-byte slots)
        lr32 = run_line_roofline(g, r32["table_bytes"])
        for op in ("retrieve", "insert"):
            tr = read_traffic(fam[op], args.log2_keys, "u32")
            if tr:
                lines_per_op = tr / LINE / r32["n"]
                lr32[f"{op}_lines_per_op"] = round(lines_per_op, 3)
                lr32[f"{op}_frac"] = round(r32["n"] / (r32[f"{op}_ms"] * 1e-3) * lines_per_op / 1e9 /
                                           lr32["glines_s"], 4)
        extras["u32u32_cg16"]["line_roofline"] = lr32
        log(f"[bench] u32 line roofline: {lr32}")
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline()
        log(f"[bench] cpu: {cpu}")

    out = {
        "metric": METRIC,
        "value": round(2 * n / r64["wall_ms_per_step"] / 1e6, 3),
        "unit": "Gops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["wall_ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: 2^28 distinct fmix64-scrambled u64 keys, splitmix64 u64 values, "
                "2^28 lookups (2^27 stored + 2^27 absent keys)",
        "config": {"workload": f"SingleValueHashTable<u64,u64> init+insert+retrieve, 2^{args.log2_keys} keys, "
                               f"LF 0.9 ({r64['capacity']} slots, {r64['table_bytes'] / 1e9:.2f} GB), cg 8, "
                               f"50% hit lookups",
                   "keys": n, "capacity": r64["capacity"], "cg_size": 8,
                   "insert_gops": round(n / r64["insert_ms"] / 1e6, 3),
                   "retrieve_gops": round(n / r64["retrieve_ms"] / 1e6, 3),
                   "kernel_only_gops": round(2 * n / (r64["insert_ms"] + r64["retrieve_ms"]) / 1e6, 3)},
        "roofline": roofline,
        "line_roofline": lr,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
