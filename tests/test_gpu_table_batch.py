"""TableBatch: thousands of small tables in one persistent launch, checked
per table against the oracle and against plain per-table launches."""
import numpy as np
import pytest
import torch

from datagen import dev_empty, dtype, from_dev, to_dev, unique_keys

pytestmark = pytest.mark.gpu

S_NONE, S_PLE, S_INVALID_KEY, S_DUP, S_NOT_FOUND = 0, 2, 8, 16, 32


def _workload(num_tables, kb, seed, lo=6, hi=14, lf=(0.5, 0.9), absent_frac=0.5):
    """Per table: capacity log-uniform in 2^lo..2^hi, fill uniform in lf,
    distinct keys (tables may share keys), queries = keys + absent keys."""
    rng = np.random.default_rng(seed)
    caps = np.exp2(rng.uniform(lo, hi, num_tables)).astype(np.uint64)
    fills = rng.uniform(*lf, num_tables)
    n = np.maximum((caps * fills).astype(np.int64), 1)
    keys, queries = [], []
    for t in range(num_tables):
        k = unique_keys(int(n[t]), kb, seed=seed * 100003 + t)
        a = unique_keys(int(n[t] * absent_frac) + 1, kb, seed=seed * 100003 + t + 7_000_000, exclude=k)
        keys.append(k)
        q = np.concatenate([k, a])
        rng.shuffle(q)
        queries.append(q)
    vals = [unique_keys(len(k), kb, seed=seed + 99 + t) for t, k in enumerate(keys)]
    off = np.concatenate([[0], np.cumsum([len(k) for k in keys])]).astype(np.uint64)
    qoff = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.uint64)
    return caps, keys, vals, queries, off, qoff


@pytest.mark.parametrize("kb,cg", [(32, 16), (32, 8), (64, 8), (32, 4)])
def test_table_batch_matches_oracle_per_table(cuda, oracle, kb, cg):
    from warpcore_amd.tables import TableBatch
    print(f"[batch] kb={kb} cg={cg}", flush=True)
    caps, keys, vals, queries, off, qoff = _workload(600, kb, seed=3 + cg)
    b = TableBatch(caps, kb, kb, cg_size=cg)
    assert b.num_tables() == len(caps)
    for t in (0, 17, 599):
        assert b.capacity(t) >= caps[t] and b.capacity(t) % cg == 0
    K, V, Q = (to_dev(np.concatenate(x)) for x in (keys, vals, queries))
    st = dev_empty(K.numel(), 32, 0x77)
    b.insert(K, V, to_dev(off), status_out=st)
    assert np.all(from_dev(st, np.uint32) == S_NONE)
    out = dev_empty(Q.numel(), kb, 0xAB)
    qst = dev_empty(Q.numel(), 32, 0x77)
    b.retrieve(Q, to_dev(qoff), out, status_out=qst)
    got, gst = from_dev(out, dtype(kb)), from_dev(qst, np.uint32)
    words = from_dev(b.status_words(), np.uint32)
    for t in range(0, len(caps), 7):
        o = oracle.SVHT(int(caps[t]), kb, kb, cg=cg)
        assert np.all(o.insert(keys[t], vals[t]) == S_NONE)
        ov, ost = o.retrieve(queries[t])
        sl = slice(int(qoff[t]), int(qoff[t + 1]))
        np.testing.assert_array_equal(gst[sl], ost)
        hit = ost == S_NONE
        np.testing.assert_array_equal(got[sl][hit], ov[hit])
        assert np.all(got[sl][~hit] == np.array(0xAB, dtype(kb)))   # misses leave the output alone
        assert words[t] == o.status()


def test_table_batch_matches_plain_launches(cuda):
    from warpcore_amd.tables import SingleValueHashTable, TableBatch
    caps, keys, vals, queries, off, qoff = _workload(64, 32, seed=11, lo=6, hi=16)
    # duplicates inside some tables: value overwrite, one `none` per key
    for t in range(0, 64, 5):
        keys[t] = np.concatenate([keys[t], keys[t][: len(keys[t]) // 3]])
        vals[t] = np.concatenate([vals[t], vals[t][: len(vals[t]) // 3]])
    off = np.concatenate([[0], np.cumsum([len(k) for k in keys])]).astype(np.uint64)
    b = TableBatch(caps, 32, 32, cg_size=16)
    K, V, Q = (to_dev(np.concatenate(x)) for x in (keys, vals, queries))
    st = dev_empty(K.numel(), 32)
    b.insert(K, V, to_dev(off), status_out=st)
    out, qst = dev_empty(Q.numel(), 32, 0xAB), dev_empty(Q.numel(), 32)
    b.retrieve(Q, to_dev(qoff), out, status_out=qst)
    gst, got, gq = from_dev(st, np.uint32), from_dev(out, np.uint32), from_dev(qst, np.uint32)
    for t in range(0, 64, 3):
        plain = SingleValueHashTable(int(caps[t]), 32, 32, cg_size=16)
        pst, pout, pq = dev_empty(len(keys[t]), 32), dev_empty(len(queries[t]), 32, 0xAB), dev_empty(len(queries[t]), 32)
        plain.insert(to_dev(keys[t]), to_dev(vals[t]), status_out=pst)
        plain.retrieve(to_dev(queries[t]), pout, status_out=pq)
        s = slice(int(off[t]), int(off[t + 1]))
        q = slice(int(qoff[t]), int(qoff[t + 1]))
        # duplicate pairs carry equal values here, so every result is determined
        np.testing.assert_array_equal(np.sort(gst[s]), np.sort(from_dev(pst, np.uint32)))
        np.testing.assert_array_equal(gq[q], from_dev(pq, np.uint32))
        np.testing.assert_array_equal(got[q], from_dev(pout, np.uint32))


def test_table_batch_full_tables_and_invalid_keys(cuda):
    from warpcore_amd.tables import TableBatch
    caps = np.array([64, 64, 200], np.uint64)
    b = TableBatch(caps, 32, 32, cg_size=16)
    c0 = b.capacity(0)
    k0 = unique_keys(c0 + 20, 32, seed=1)          # 20 more keys than slots
    k1 = np.array([5, 0, 6, 0xFFFFFFFF], np.uint32)  # EMPTY and TOMBSTONE are invalid
    k2 = unique_keys(100, 32, seed=2)
    keys = [k0, k1, k2]
    off = np.concatenate([[0], np.cumsum([len(k) for k in keys])]).astype(np.uint64)
    K = to_dev(np.concatenate(keys))
    st = dev_empty(K.numel(), 32)
    b.insert(K, K, to_dev(off), status_out=st)
    s = from_dev(st, np.uint32)
    s0, s1 = s[: len(k0)], s[len(k0): len(k0) + 4]
    assert np.sum(s0 == S_NONE) == c0 and np.sum(s0 == S_PLE) == 20
    assert list(s1) == [S_NONE, S_INVALID_KEY, S_NONE, S_INVALID_KEY]
    words = from_dev(b.status_words(), np.uint32)
    assert words[0] == S_PLE and words[1] == S_INVALID_KEY and words[2] == S_NONE
    b.init()
    assert np.all(from_dev(b.status_words(), np.uint32) == 0)


def test_table_batch_tuner_keeps_results(cuda):
    from warpcore_amd.tables import TableBatch
    caps, keys, vals, queries, off, qoff = _workload(300, 32, seed=21)
    b = TableBatch(caps, 32, 32, cg_size=16)
    K, V, Q = (to_dev(np.concatenate(x)) for x in (keys, vals, queries))
    O, QO = to_dev(off), to_dev(qoff)
    out = dev_empty(Q.numel(), 32)
    launch, ms = b.tune(K, V, O, Q, QO, out)
    assert launch[0] in (4, 8) and launch[1] in (32, 64, 128, 256) and launch[2] in (4, 8, 16)
    assert ms > 0 and b.get_launch() == launch
    for cfg in [launch, (4, 32, 4), (8, 256, 16)]:
        b.set_launch(*cfg)
        b.init()
        st = dev_empty(K.numel(), 32)
        b.insert(K, V, O, status_out=st)
        out, qst = dev_empty(Q.numel(), 32, 0xAB), dev_empty(Q.numel(), 32)
        b.retrieve(Q, QO, out, status_out=qst)
        assert np.all(from_dev(st, np.uint32) == S_NONE)
        gq = from_dev(qst, np.uint32)
        got = from_dev(out, np.uint32)
        for t in range(0, 300, 11):
            q = queries[t]
            sl = slice(int(qoff[t]), int(qoff[t + 1]))
            lookup = dict(zip(keys[t].tolist(), vals[t].tolist()))
            exp = np.array([lookup.get(x, 0xAB) for x in q.tolist()], np.uint32)
            np.testing.assert_array_equal(got[sl], exp)
            np.testing.assert_array_equal(gq[sl], np.where(np.isin(q, keys[t]), S_NONE, S_NOT_FOUND))


def test_table_batch_full_configuration(cuda, oracle):
    """The batched configuration at full size: 65,536 SingleValueHashTable
    <u32,u32> of capacity log-uniform in 2^6..2^16 slots, each filled to a
    load factor uniform in [0.5, 0.9] (about 433 M keys). Size-independent
    properties over every table: all inserts report none, every stored key
    returns its value, every absent key reports key_not_found and leaves its
    output untouched, and the per-table status words are exact; plus the
    oracle on 48 tables spread over the whole capacity range (the largest
    ones, of 2^15..2^16 slots, included)."""
    from warpcore_amd import _capi
    from warpcore_amd.tables import TableBatch
    lib = _capi.lib()
    num_tables = 65536
    rng = np.random.default_rng(0xC7)
    caps = np.exp2(rng.uniform(6, 16, num_tables)).astype(np.uint64)
    n_t = np.maximum((caps * rng.uniform(0.5, 0.9, num_tables)).astype(np.uint64), 1)
    off = np.concatenate([[0], np.cumsum(n_t)]).astype(np.uint64)
    n = int(off[-1])
    print(f"[batch] full configuration: {num_tables} tables, {n} keys", flush=True)
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    absent = torch.empty(n, dtype=torch.int32, device="cuda")
    values = torch.empty(n, dtype=torch.int32, device="cuda")
    # fmix32 is a bijection: offsets 0x7000 + [0, n) and 0x7000 + [n, 2n) give disjoint key sets
    assert lib.wc_gen_unique_u32(keys.data_ptr(), n, 0x7000, None) == 0
    assert lib.wc_gen_unique_u32(absent.data_ptr(), n, 0x7000 + n, None) == 0
    assert lib.wc_gen_uniform_u32(values.data_ptr(), n, 0xC7, None) == 0
    offs = to_dev(off)
    b = TableBatch(caps, 32, 32, cg_size=16)
    st = torch.full((n,), 0x77, dtype=torch.int32, device="cuda")
    b.insert(keys, values, offs, status_out=st)
    assert bool((st == S_NONE).all())
    assert bool((b.status_words() == S_NONE).all())
    out = torch.full((n,), 0xAB, dtype=torch.int32, device="cuda")
    qst = torch.full((n,), 0x77, dtype=torch.int32, device="cuda")
    b.retrieve(keys, offs, out, status_out=qst)
    assert bool((qst == S_NONE).all()) and torch.equal(out, values)
    miss = torch.full((n,), 0xAB, dtype=torch.int32, device="cuda")
    b.retrieve(absent, offs, miss, status_out=qst)
    assert bool((qst == S_NOT_FOUND).all()) and bool((miss == 0xAB).all())
    assert bool((b.status_words() == S_NOT_FOUND).all())
    torch.cuda.synchronize()
    del miss, qst, st
    # oracle per table on a spread of capacities, the largest included
    order = np.argsort(caps, kind="stable")
    sample = np.unique(np.concatenate([order[np.linspace(0, num_tables - 1, 40).astype(np.int64)], order[-8:]]))
    assert caps[sample].max() >= (1 << 15)
    b.init()
    st = torch.full((n,), 0x77, dtype=torch.int32, device="cuda")
    b.insert(keys, values, offs, status_out=st)
    # queries of table t: its keys, then its absent keys (two lookups over
    # the same offsets)
    out1 = torch.full((n,), 0xAB, dtype=torch.int32, device="cuda")
    out2 = torch.full((n,), 0xAB, dtype=torch.int32, device="cuda")
    s1 = torch.full((n,), 0x77, dtype=torch.int32, device="cuda")
    s2 = torch.full((n,), 0x77, dtype=torch.int32, device="cuda")
    b.retrieve(keys, offs, out1, status_out=s1)
    b.retrieve(absent, offs, out2, status_out=s2)
    torch.cuda.synchronize()
    words = from_dev(b.status_words(), np.uint32)
    for t in sample.tolist():
        lo, hi = int(off[t]), int(off[t + 1])
        k = from_dev(keys[lo:hi], np.uint32)
        a = from_dev(absent[lo:hi], np.uint32)
        v = from_dev(values[lo:hi], np.uint32)
        o = oracle.SVHT(int(caps[t]), 32, 32, cg=16)
        np.testing.assert_array_equal(from_dev(st[lo:hi], np.uint32), o.insert(k, v), err_msg=f"table {t}")
        q = np.concatenate([k, a])
        ov, ost = o.retrieve(q)
        gs = np.concatenate([from_dev(s1[lo:hi], np.uint32), from_dev(s2[lo:hi], np.uint32)])
        gv = np.concatenate([from_dev(out1[lo:hi], np.uint32), from_dev(out2[lo:hi], np.uint32)])
        np.testing.assert_array_equal(gs, ost, err_msg=f"table {t}")
        hit = ost == S_NONE
        np.testing.assert_array_equal(gv[hit], ov[hit], err_msg=f"table {t}")
        assert np.all(gv[~hit] == 0xAB)
        assert words[t] == o.status(), f"table {t}"
