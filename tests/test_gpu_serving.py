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
