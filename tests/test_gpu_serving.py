"""Serving loop: a mixed request stream through the server gives, request by
request, the results of running the same requests one plain launch at a
time in arrival order (per-table linearised order); admission refusals;
concurrent clients."""
import threading

import numpy as np
import pytest
import torch

from datagen import dev_empty, from_dev, to_dev, unique_keys

pytestmark = pytest.mark.gpu

S_NONE, S_DUP, S_NOT_FOUND, S_OOM = 0, 16, 32, 128
INSERT, RETRIEVE, ERASE = 0, 1, 2


def _stream(num_requests, num_tables, seed, lo=8, hi=16, universe=1 << 18):
    """Requests over a shared key universe per table so lookups hit: 30 %
    insert, 60 % retrieve, 10 % erase; sizes log-uniform 2^lo..2^hi; 20 % of
    the keys of a request from a Zipf(1.0) hot set. Values are a function of
    (key, request) so duplicates inside one insert carry one value; erase
    keys are distinct within a request."""
    rng = np.random.default_rng(seed)
    uni = unique_keys(universe, 64, seed=seed)
    hot = uni[: 1 << 10]
    w = 1.0 / np.arange(1, len(hot) + 1)
    w /= w.sum()
    ops = rng.choice([INSERT, RETRIEVE, ERASE], size=num_requests, p=[0.3, 0.6, 0.1]).astype(np.uint32)
    ops[: num_tables] = INSERT   # every table starts with an insert
    tables = rng.integers(0, num_tables, size=num_requests).astype(np.uint32)
    tables[: num_tables] = np.arange(num_tables)
    sizes = np.exp2(rng.uniform(lo, hi, size=num_requests)).astype(np.uint64)
    keys, vals = [], []
    for r in range(num_requests):
        n = int(sizes[r])
        k = rng.choice(uni, size=n)
        nh = n // 5
        k[:nh] = rng.choice(hot, size=nh, p=w)
        if ops[r] == ERASE:
            k = np.unique(k)
            rng.shuffle(k)
            sizes[r] = len(k)
        keys.append(k.astype(np.uint64))
        vals.append((k * np.uint64(3) + np.uint64(r)).astype(np.uint64))
    arrival = np.cumsum(rng.exponential(200.0, size=num_requests))   # microseconds
    return ops, tables, sizes, arrival, keys, vals


def _run(mode, caps, ops, tables, sizes, arrival, keys, vals, **kw):
    from warpcore_amd.tables import Server
    srv = Server(caps, 64, 64, **kw)
    K, V = to_dev(np.concatenate(keys)), to_dev(np.concatenate(vals))
    out = dev_empty(K.numel(), 64, 0x5A5A)
    st = dev_empty(K.numel(), 32, 0x77)
    torch.cuda.synchronize()
    lat, mk = srv.drive(mode, ops, tables, sizes, arrival, K, V, out, st)
    torch.cuda.synchronize()
    sizes_t = [srv.table_size(t) for t in range(len(caps))]
    status_t = [int(srv.table_status(t)) for t in range(len(caps))]
    res = from_dev(out, np.uint64), from_dev(st, np.uint32), lat, mk, sizes_t, status_t, srv.stats()
    del srv
    return res


@pytest.mark.parametrize("use_graph,max_step_keys,direct_keys",
                         [(True, 0, None), (False, 0, None), (True, 5000, None), (True, 0, 2048)])
def test_server_matches_one_launch_per_request(cuda, use_graph, max_step_keys, direct_keys):
    """max_step_keys=5000: requests of up to 2^16 keys run in chunks over
    several steps. direct_keys=2048: with a short backlog, requests of 2048
    keys or more run alone on their table's bulk kernels, the rest in
    steps of the step kernel."""
    print(f"[serving] use_graph={use_graph} max_step_keys={max_step_keys} direct_keys={direct_keys}", flush=True)
    caps = np.full(6, 1 << 20, np.uint64)
    ops, tables, sizes, arrival, keys, vals = _stream(240, len(caps), seed=5)
    out_s, st_s, lat_s, mk_s, size_s, tst_s, stats = _run(0, caps, ops, tables, sizes, arrival, keys, vals,
                                                          use_graph=use_graph, max_step_keys=max_step_keys,
                                                          direct_keys=direct_keys)
    out_p, st_p, lat_p, mk_p, size_p, tst_p, _ = _run(1, caps, ops, tables, sizes, arrival, keys, vals)
    assert stats["requests"] == len(ops) and stats["keys"] == int(sizes.sum())
    if direct_keys is None:
        assert stats["direct_steps"] == 0
    else:   # large requests, and steps of one whole request
        assert int(np.sum(sizes >= direct_keys)) > 0 and stats["direct_steps"] > 0
    if max_step_keys:
        assert stats["steps"] >= int(sizes.sum()) // max_step_keys
    else:
        assert 1 <= stats["steps"] <= len(ops)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    for r in range(len(ops)):
        s = slice(off[r], off[r + 1])
        if ops[r] == RETRIEVE:
            np.testing.assert_array_equal(st_s[s], st_p[s], err_msg=f"request {r}")
            np.testing.assert_array_equal(out_s[s], out_p[s], err_msg=f"request {r}")
        elif ops[r] == INSERT:
            # which duplicate of a key reports `none` is unspecified
            np.testing.assert_array_equal(np.sort(st_s[s]), np.sort(st_p[s]), err_msg=f"request {r}")
        else:
            np.testing.assert_array_equal(st_s[s], st_p[s], err_msg=f"request {r}")
    assert size_s == size_p
    assert tst_s == tst_p
    assert np.all(lat_s > 0) and np.all(lat_p > 0)


def test_server_refuses_inserts_over_budget(cuda):
    from warpcore_amd.tables import Server
    srv = Server([1000], 64, 64, max_load=0.5)
    cap = srv.table_capacity(0)
    budget = int(cap * 0.5)
    a = to_dev(unique_keys(budget - 10, 64, seed=1))
    b = to_dev(unique_keys(20, 64, seed=2, exclude=from_dev(a, np.uint64)))
    sa, sb = dev_empty(a.numel(), 32, 0x77), dev_empty(b.numel(), 32, 0x77)
    torch.cuda.synchronize()
    ra = srv.submit(INSERT, 0, a, a, None, sa)
    rb = srv.submit(INSERT, 0, b, b, None, sb)
    assert srv.wait(ra) is True
    assert srv.wait(rb) is False                      # 20 more keys than the budget left
    assert np.all(from_dev(sa, np.uint32) == S_NONE)
    assert np.all(from_dev(sb, np.uint32) == S_OOM)
    assert srv.table_size(0) == budget - 10
    # a lookup after the refused insert sees only the first batch
    out, st = dev_empty(b.numel(), 64), dev_empty(b.numel(), 32)
    torch.cuda.synchronize()
    assert srv.wait(srv.submit(RETRIEVE, 0, b, None, out, st))
    assert np.all(from_dev(st, np.uint32) == S_NOT_FOUND)


def test_server_concurrent_clients(cuda):
    from warpcore_amd.tables import Server
    caps = np.full(8, 1 << 19, np.uint64)
    srv = Server(caps, 64, 64)
    rng = np.random.default_rng(3)
    universe = [unique_keys(1 << 14, 64, seed=100 + t) for t in range(8)]
    results = {}

    def client(c):
        # client c owns tables 2c and 2c+1: insert, then look up twice
        mine = []
        for t in (2 * c, 2 * c + 1):
            k = to_dev(universe[t])
            v = to_dev(universe[t] ^ np.uint64(0xABCD))
            s = dev_empty(k.numel(), 32)
            torch.cuda.synchronize()
            mine.append((srv.submit(INSERT, t, k, v, None, s), k, v, s))
            q = to_dev(np.concatenate([universe[t][::2], universe[(t + 1) % 8][:100]]))
            o, qs = dev_empty(q.numel(), 64, 0x11), dev_empty(q.numel(), 32)
            torch.cuda.synchronize()
            mine.append((srv.submit(RETRIEVE, t, q, None, o, qs), q, o, qs))
        results[c] = mine

    threads = [threading.Thread(target=client, args=(c,)) for c in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    srv.drain()
    for c in range(4):
        for j, t in enumerate((2 * c, 2 * c + 1)):
            _, _, _, s = results[c][2 * j]
            assert np.all(from_dev(s, np.uint32) == S_NONE)
            _, q, o, qs = results[c][2 * j + 1]
            qn = from_dev(q, np.uint64)
            hit = np.isin(qn, universe[t])
            got_o, got_s = from_dev(o, np.uint64), from_dev(qs, np.uint32)
            assert np.all(got_s[hit] == S_NONE) and np.all(got_s[~hit] == S_NOT_FOUND)
            np.testing.assert_array_equal(got_o[hit], qn[hit] ^ np.uint64(0xABCD))
    assert srv.stats()["requests"] == 16


def _pair_multiset(keys, st):
    order = np.lexsort((st, keys))
    return keys[order], st[order]


def test_server_matches_oracle_per_table(cuda, oracle):
    """A served stream equals the oracle fed each table's requests in
    arrival order (the server's per-table linearised order): every lookup's
    values and statuses, every erase's statuses, every insert's multiset of
    (key, status) pairs (which duplicate of a key reports `none` is free),
    the final sizes and status words. 8 tables of 2^20 slots, requests of
    2^8..2^18 keys arriving as a Poisson stream."""
    print("[serving] oracle replay: 8 tables, 160 requests of 2^8..2^18 keys", flush=True)
    caps = np.full(8, 1 << 20, np.uint64)
    ops, tables, sizes, arrival, keys, vals = _stream(160, len(caps), seed=17, lo=8, hi=18, universe=1 << 18)
    out_s, st_s, lat_s, mk_s, size_s, tst_s, stats = _run(0, caps, ops, tables, sizes, arrival, keys, vals)
    assert stats["requests"] == len(ops) and stats["failed"] == 0
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    models = [oracle.SVHT(int(c), 64, 64, cg=8) for c in caps]
    for r in range(len(ops)):
        s = slice(off[r], off[r + 1])
        o = models[tables[r]]
        if ops[r] == INSERT:
            exp = o.insert(keys[r], vals[r])
            gk, gs = _pair_multiset(keys[r], st_s[s])
            ek, es = _pair_multiset(keys[r], exp)
            np.testing.assert_array_equal(gs, es, err_msg=f"insert request {r}")
        elif ops[r] == RETRIEVE:
            ev, es = o.retrieve(keys[r])
            np.testing.assert_array_equal(st_s[s], es, err_msg=f"retrieve request {r}")
            hit = es == S_NONE
            np.testing.assert_array_equal(out_s[s][hit], ev[hit], err_msg=f"retrieve request {r}")
            assert np.all(out_s[s][~hit] == np.uint64(0x5A5A)), f"request {r}: a miss wrote its output"
        else:
            es = o.erase(keys[r])
            np.testing.assert_array_equal(st_s[s], es, err_msg=f"erase request {r}")
    assert size_s == [m.size() for m in models]
    assert tst_s == [m.status() for m in models]


def test_server_config_shape_matches_launch_per_request(cuda):
    """The serving configuration's shape: 64 resident SingleValueHashTable
    <u64,u64> of 97.66 M slots (100 GB), 2,000 requests of 2^8..2^22 keys
    (log-uniform) arriving as a Poisson stream, 30 % insert / 60 % retrieve /
    10 % erase, keys from a 2^28-key universe with 20 % Zipf(1.0) hot keys.
    The server and one plain launch per request (each on its own 100 GB of
    tables) must agree on every lookup's value and status, every erase's
    status, every insert's count of `none` statuses and the sum of the keys
    reporting it, and every table's size and status word."""
    from warpcore_amd import _capi
    from warpcore_amd.tables import Server
    lib = _capi.lib()
    num_tables, num_requests = 64, 2000
    rng = np.random.default_rng(0xC5)
    caps = np.full(num_tables, int(100e9 / 16 / num_tables), np.uint64)
    sizes = np.exp2(rng.uniform(8, 22, num_requests)).astype(np.uint64)
    ops = rng.choice([INSERT, RETRIEVE, ERASE], size=num_requests, p=[0.3, 0.6, 0.1]).astype(np.uint32)
    tabs = rng.integers(0, num_tables, size=num_requests).astype(np.uint32)
    n = int(sizes.sum())
    print(f"[serving] config shape: {num_tables} tables x {int(caps[0])} slots, {num_requests} requests, "
          f"{n} keys", flush=True)
    U = 1 << 28
    universe = torch.empty(U, dtype=torch.int64, device="cuda")
    assert lib.wc_gen_unique_u64(universe.data_ptr(), U, 1, None) == 0
    g = torch.Generator(device="cuda")
    g.manual_seed(0xC5)
    idx = torch.randint(0, U, (n,), device="cuda", generator=g)
    # hot keys: Zipf(1.0) ranks k over the first 2^20 universe keys (universe[k - 1])
    hot = torch.empty(n, dtype=torch.int64, device="cuda")
    assert lib.wc_gen_zipf_u64(hot.data_ptr(), n, 1 << 20, 1.0, 0xC6, 1, None) == 0
    take = torch.rand(n, device="cuda", generator=g) < 0.2
    keys = torch.where(take, hot, universe[idx])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    for r in np.flatnonzero(ops == ERASE):   # distinct keys within an erase: a contiguous universe range
        m = int(sizes[r])
        start = int(rng.integers(0, U - m))
        keys[off[r]:off[r + 1]] = universe[start:start + m]
    del universe, hot, take, idx
    req = torch.repeat_interleave(torch.arange(num_requests, device="cuda"),
                                  torch.from_numpy(sizes.astype(np.int64)).to("cuda"))
    values = keys * 3 + req          # one value per (key, request): duplicates in an insert agree
    rate = 0.5 * num_requests / 150e3   # about half the launch-per-request throughput
    arrival = np.cumsum(rng.exponential(1.0 / rate, size=num_requests))
    torch.cuda.synchronize()

    def run(mode):
        srv = Server(caps, 64, 64)
        out = torch.full((n,), 0x5A5A, dtype=torch.int64, device="cuda")
        st = torch.full((n,), 0x77, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        srv.drive(mode, ops, tabs, sizes, arrival, keys, values, out, st)
        torch.cuda.synchronize()
        sizes_t = [srv.table_size(t) for t in range(num_tables)]
        status_t = [int(srv.table_status(t)) for t in range(num_tables)]
        stats = srv.stats() if mode == 0 else None
        del srv
        torch.cuda.empty_cache()
        return out, st, sizes_t, status_t, stats

    out_s, st_s, size_s, tst_s, stats = run(0)
    print(f"[serving] server: {stats}", flush=True)
    out_p, st_p, size_p, tst_p, _ = run(1)
    assert stats["requests"] == num_requests and stats["failed"] == 0
    op_of = torch.from_numpy(ops.astype(np.int64)).to("cuda")[req]
    look, ins, era = op_of == RETRIEVE, op_of == INSERT, op_of == ERASE
    assert torch.equal(st_s[look], st_p[look]) and torch.equal(out_s[look], out_p[look])
    assert torch.equal(st_s[era], st_p[era])
    hits = int((st_s[look] == S_NONE).sum())
    assert hits > 0.05 * int(look.sum())         # the stream really looks up stored keys
    for st in (st_s, st_p):
        assert bool(((st[ins] == S_NONE) | (st[ins] == S_DUP)).all())

    def none_by_request(st):
        first = (st == S_NONE) & ins
        cnt = torch.zeros(num_requests, dtype=torch.int64, device="cuda").index_add_(0, req, first.long())
        ksum = torch.zeros(num_requests, dtype=torch.int64, device="cuda").index_add_(0, req,
                                                                                      torch.where(first, keys, 0))
        return cnt, ksum

    cs, ks = none_by_request(st_s)
    cp, kp = none_by_request(st_p)
    assert torch.equal(cs, cp) and torch.equal(ks, kp)
    assert size_s == size_p and tst_s == tst_p


def test_server_insert_erase_cycles_stay_admitted(cuda):
    """Erase leaves tombstones that keep their slots; an insert that no
    longer fits the budget has its table compacted (live pairs re-inserted)
    and is admitted. Repeated inserts of stored keys claim nothing: the
    recount of the table's claimed slots admits them. Each round: lookups
    see exactly the live keys."""
    from warpcore_amd.tables import Server
    srv = Server([1 << 17], 64, 64)
    budget = srv.table_budget(0)
    n = int(budget * 0.6)
    live = None
    for rnd in range(4):
        print(f"[serving] cycle {rnd}", flush=True)
        k = unique_keys(n, 64, seed=200 + rnd)
        kd, vd = to_dev(k), to_dev(k ^ np.uint64(rnd + 1))
        for rep in range(3):   # the first insert claims n slots, the repeats none
            st = dev_empty(n, 32, 0x77)
            torch.cuda.synchronize()
            assert srv.wait(srv.submit(INSERT, 0, kd, vd, None, st)), f"round {rnd} insert {rep} refused"
            assert np.all(from_dev(st, np.uint32) == (S_NONE if rep == 0 else S_DUP))
        assert srv.table_reserved(0) <= budget
        q = np.concatenate([k, live]) if live is not None else k
        qd = to_dev(q)
        out, qs = dev_empty(len(q), 64, 0x1), dev_empty(len(q), 32)
        torch.cuda.synchronize()
        assert srv.wait(srv.submit(RETRIEVE, 0, qd, None, out, qs))
        got_s = from_dev(qs, np.uint32)
        assert np.all(got_s[:n] == S_NONE) and np.all(got_s[n:] == S_NOT_FOUND)
        np.testing.assert_array_equal(from_dev(out, np.uint64)[:n], k ^ np.uint64(rnd + 1))
        es = dev_empty(n, 32, 0x77)
        torch.cuda.synchronize()
        assert srv.wait(srv.submit(ERASE, 0, kd, None, None, es))
        assert np.all(from_dev(es, np.uint32) == S_NONE)
        live = k
    assert srv.stats()["compactions"] >= 2
    assert srv.table_size(0) == 0


def test_server_failed_step_fails_its_requests(cuda):
    """Failure handling forced from the host: the scheduler's second step
    is reported as a launch failure without running. wait() returns the
    failure code for its request; the other requests are served, and the
    table holds only what the served inserts stored."""
    from warpcore_amd.tables import Server
    srv = Server([1 << 18], 64, 64)
    srv.fail_step(2)
    batches = [unique_keys(1000, 64, seed=300 + i) for i in range(3)]
    codes = []
    for b in batches:
        kd = to_dev(b)
        st = dev_empty(len(b), 32)
        torch.cuda.synchronize()
        rid = srv.submit(INSERT, 0, kd, kd, None, st)
        codes.append(srv.wait_code(rid))   # one request per step: waited before the next
    assert codes == [Server.SERVED, Server.FAILED, Server.SERVED]
    assert srv.stats()["failed"] == 1
    assert srv.table_size(0) == 2000
    assert srv.wait_code(10 ** 9) == Server.UNKNOWN_ID
