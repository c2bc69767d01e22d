"""BucketListHashTable on the GPU vs the oracle: per-key counts and value
multisets, pool usage, statuses, out-of-memory and max_values paths, and
exact value order on sequential appends."""
import numpy as np
import pytest
import torch

from datagen import dev_empty, dtype, from_dev, to_dev, unique_keys

pytestmark = pytest.mark.gpu

S_NONE, S_PLE, S_INVALID_CONFIG, S_INVALID_KEY, S_DUP, S_NOT_FOUND = 0, 2, 4, 8, 16, 32
S_OOM, S_NOT_INIT, S_DRY_RUN, S_MAXV = 1 << 7, 1 << 8, 1 << 9, 1 << 11


def _pairs(num_keys, bits, mean, seed):
    rng = np.random.default_rng(seed)
    keys = unique_keys(num_keys, bits, seed)
    mult = rng.geometric(1.0 / mean, size=num_keys)
    k = np.repeat(keys, mult)
    v = unique_keys(len(k), bits, seed + 1)
    perm = rng.permutation(len(k))
    return keys, mult, k[perm], v[perm]


def _expected(k, v):
    out = {}
    for key, val in zip(k.tolist(), v.tolist()):
        out.setdefault(key, []).append(val)
    return {key: sorted(vals) for key, vals in out.items()}


def _retrieve(t, q, total_hint, bits):
    n = len(q)
    begin, end = dev_empty(n, 64), dev_empty(n, 64)
    vals = dev_empty(max(total_hint, 1), bits, 0xAB)
    st = dev_empty(n, 32, 0x77)
    total = t.retrieve(to_dev(q), begin, end, vals, status_out=st)
    return (total, from_dev(begin, np.uint64), from_dev(end, np.uint64), from_dev(vals, dtype(bits)),
            from_dev(st, np.uint32))


@pytest.mark.parametrize("kb,cg,mean,grow,minb,maxb", [
    (32, 8, 8.0, 1.1, 1, 1023),
    (64, 8, 6.0, 2.0, 4, 64),
    (32, 4, 20.0, 1.5, 2, 16),
    (64, 16, 3.0, 1.0, 3, 3),
    (32, 1, 5.0, 1.1, 1, 1023),
])
def test_bucketlist_matches_oracle(cuda, oracle, kb, cg, mean, grow, minb, maxb):
    from warpcore_amd.tables import BucketListHashTable
    print(f"[bucketlist] kb={kb} cg={cg} mean={mean} grow={grow} min={minb} max={maxb}", flush=True)
    keys, mult, k, v = _pairs(4000, kb, mean, seed=7 + cg)
    key_cap, value_cap = int(len(keys) / 0.8), 4 * len(k) + 1000
    t = BucketListHashTable(key_cap, value_cap, kb, kb, grow_factor=grow, min_bucket_size=minb,
                            max_bucket_size=maxb, cg_size=cg)
    o = oracle.BucketList(key_cap, value_cap, kb, kb, cg=cg, grow=grow, min_bucket=minb, max_bucket=maxb)
    assert t.key_capacity() == o.key_capacity()
    st = dev_empty(len(k), 32, 0x77)
    t.insert(to_dev(k), to_dev(v), status_out=st)
    s_gpu = from_dev(st, np.uint32)
    s_ora = o.insert(k, v)
    assert set(np.unique(s_gpu)) <= {S_NONE, S_DUP}
    assert np.sum(s_gpu == S_NONE) == np.sum(s_ora == S_NONE) == len(keys)
    assert t.num_keys() == o.num_keys() == len(keys)
    assert t.num_values() == len(k)
    # bucket sizes depend only on each key's count, so pool usage is exact
    assert t.value_cells_used() == o.next_free()
    assert t.peek_status() == o.status()
    absent = unique_keys(300, kb, seed=77, exclude=keys)
    q = np.concatenate([keys, absent])
    q = q[np.random.default_rng(2).permutation(len(q))]
    total, b, e, vals, qst = _retrieve(t, q, len(k) + 8, kb)
    ob, oe, ov, ost = o.retrieve(q)
    assert total == len(k)
    np.testing.assert_array_equal(b, ob)
    np.testing.assert_array_equal(e, oe)
    np.testing.assert_array_equal(qst, ost)
    assert np.all(vals[len(k):] == np.array(0xAB, dtype(kb)))
    exp = _expected(k, v)
    for key, lo, hi in zip(q.tolist(), b, e):
        assert sorted(vals[int(lo):int(hi)].tolist()) == exp.get(key, [])


@pytest.mark.parametrize("kb,minb,grow", [(32, 1, 1.1), (64, 3, 2.0)])
def test_bucketlist_value_order_on_sequential_appends(cuda, oracle, kb, minb, grow):
    """One pair per launch fixes the append order: values come back tail
    bucket first exactly as the oracle walks them."""
    from warpcore_amd.tables import BucketListHashTable
    keys, mult, k, v = _pairs(12, kb, 12.0, seed=5)
    t = BucketListHashTable(100, 4000, kb, kb, grow_factor=grow, min_bucket_size=minb, max_bucket_size=5)
    o = oracle.BucketList(100, 4000, kb, kb, grow=grow, min_bucket=minb, max_bucket=5)
    for i in range(len(k)):
        st = dev_empty(1, 32, 0x77)
        t.insert(to_dev(k[i:i + 1]), to_dev(v[i:i + 1]), status_out=st)
        assert from_dev(st, np.uint32)[0] == o.insert(k[i:i + 1], v[i:i + 1])[0]
    total, b, e, vals, _ = _retrieve(t, keys, len(k), kb)
    ob, oe, ov, _ = o.retrieve(keys)
    assert total == len(ov)
    np.testing.assert_array_equal(vals[:total], ov)
    assert t.value_cells_used() == o.next_free()


def test_bucketlist_out_of_memory(cuda, oracle):
    from warpcore_amd.tables import BucketListHashTable
    keys, mult, k, v = _pairs(500, 32, 10.0, seed=9)
    value_cap = len(k) // 2
    t = BucketListHashTable(2000, value_cap, 32, 32, grow_factor=1.5, min_bucket_size=2, max_bucket_size=32)
    st = dev_empty(len(k), 32)
    t.insert(to_dev(k), to_dev(v), status_out=st)
    s = from_dev(st, np.uint32)
    oom = (s & S_OOM) != 0
    assert oom.any() and (~oom).any()
    assert set(np.unique(s & ~np.uint32(S_OOM))) <= {S_NONE, S_DUP}
    assert t.peek_status() & S_OOM
    stored = t.num_values()
    assert stored == np.sum(~oom)
    assert t.value_cells_used() >= value_cap       # the failed allocation is counted too
    total, b, e, vals, _ = _retrieve(t, keys, stored + 8, 32)
    assert total == stored
    exp = _expected(k, v)
    for key, lo, hi in zip(keys.tolist(), b, e):
        got = vals[int(lo):int(hi)].tolist()
        assert len(set(got)) == len(got) and set(got) <= set(exp[key])
    # sequential run against the oracle: the same pairs hit out-of-memory
    t2 = BucketListHashTable(2000, 60, 32, 32, grow_factor=1.5, min_bucket_size=2, max_bucket_size=32)
    o2 = oracle.BucketList(2000, 60, 32, 32, grow=1.5, min_bucket=2, max_bucket=32)
    for i in range(120):
        st = dev_empty(1, 32)
        t2.insert(to_dev(k[i:i + 1]), to_dev(v[i:i + 1]), status_out=st)
        assert from_dev(st, np.uint32)[0] == o2.insert(k[i:i + 1], v[i:i + 1])[0], i
    total, b, e, vals, _ = _retrieve(t2, keys, 200, 32)
    ob, oe, ov, _ = o2.retrieve(keys)
    np.testing.assert_array_equal(b, ob)
    np.testing.assert_array_equal(vals[:total], ov)


def _arenas(value_cap):
    """Arenas of a value pool: the largest power of two up to 256 leaving
    every arena at least 2^16 cells."""
    a = 1
    while a < 256 and value_cap // (2 * a) >= 1 << 16:
        a *= 2
    return a


def test_bucketlist_arenas_match_oracle(cuda, oracle):
    """A pool of 64 arenas: bucket sizes still depend only on each key's
    count, so pool usage, offsets and value multisets equal the oracle's."""
    from warpcore_amd.tables import BucketListHashTable
    keys, mult, k, v = _pairs(1 << 17, 32, 8.0, seed=41)
    key_cap, value_cap = int(len(keys) / 0.8), 4 * len(k)
    assert _arenas(value_cap) == 64
    t = BucketListHashTable(key_cap, value_cap, 32, 32, grow_factor=1.1, min_bucket_size=1, cg_size=8)
    empty = BucketListHashTable(key_cap, 1000, 32, 32, grow_factor=1.1, min_bucket_size=1, cg_size=8)
    # the counters of 64 arenas take one 128-byte line each; one arena, 8 bytes
    assert t.bytes_total() - empty.bytes_total() == (value_cap - 1000) * 8 + 64 * 128 - 8
    o = oracle.BucketList(key_cap, value_cap, 32, 32, cg=8, grow=1.1, min_bucket=1)
    st = dev_empty(len(k), 32, 0x77)
    t.insert(to_dev(k), to_dev(v), status_out=st)
    s_gpu = from_dev(st, np.uint32)
    s_ora = o.insert(k, v)
    assert np.sum(s_gpu == S_NONE) == np.sum(s_ora == S_NONE) == len(keys)
    assert set(np.unique(s_gpu)) <= {S_NONE, S_DUP}
    assert t.num_values() == len(k)
    assert t.value_cells_used() == o.next_free()
    total, b, e, vals, qst = _retrieve(t, keys, len(k), 32)
    ob, oe, ov, ost = o.retrieve(keys)
    assert total == len(k)
    np.testing.assert_array_equal(b, ob)
    np.testing.assert_array_equal(e, oe)
    order = np.argsort(k, kind="stable")
    ks, vs = k[order], v[order]
    starts = np.searchsorted(ks, keys)
    for i in np.random.default_rng(3).choice(len(keys), 2000, replace=False):
        want = np.sort(vs[starts[i]:starts[i] + mult[i]])
        np.testing.assert_array_equal(np.sort(vals[int(b[i]):int(e[i])]), want)
    np.testing.assert_array_equal(np.sort(vals[:total]), np.sort(v))


def test_bucketlist_arenas_out_of_memory(cuda):
    """An allocation fails only when no arena holds it: afterwards every
    arena counts as used to its end, the pairs not reported out_of_memory
    are all stored, and each key returns distinct values of its own."""
    from warpcore_amd.tables import BucketListHashTable
    keys, mult, k, v = _pairs(1 << 16, 32, 8.0, seed=43)
    value_cap = len(k) // 2
    assert _arenas(value_cap) >= 4
    t = BucketListHashTable(int(len(keys) / 0.8), value_cap, 32, 32, grow_factor=1.5, min_bucket_size=2,
                            max_bucket_size=32, cg_size=8)
    st = dev_empty(len(k), 32)
    t.insert(to_dev(k), to_dev(v), status_out=st)
    s = from_dev(st, np.uint32)
    oom = (s & S_OOM) != 0
    assert oom.any() and (~oom).any()
    assert set(np.unique(s & ~np.uint32(S_OOM))) <= {S_NONE, S_DUP}
    assert t.num_values() == np.sum(~oom)
    assert t.value_cells_used() == value_cap
    total, b, e, vals, _ = _retrieve(t, keys, len(k), 32)
    assert total == np.sum(~oom)
    stored = set(v[~oom].tolist())
    assert set(vals[:total].tolist()) == stored
    exp = _expected(k, v)
    for i in np.random.default_rng(4).choice(len(keys), 1000, replace=False):
        got = vals[int(b[i]):int(e[i])].tolist()
        assert len(set(got)) == len(got) and set(got) <= set(exp[int(keys[i])])


def test_bucketlist_out_of_memory_on_first_buckets(cuda):
    """A pool too small for every key's first bucket: a key whose first
    bucket does not fit reports out_of_memory on all its pairs (its handle
    is left full), a key that got one keeps its stored values, and no key
    holds more than its own values."""
    from warpcore_amd.tables import BucketListHashTable
    keys, mult, k, v = _pairs(500, 32, 6.0, seed=17)
    t = BucketListHashTable(2000, 300, 32, 32, grow_factor=1.5, min_bucket_size=2, max_bucket_size=32)
    st = dev_empty(len(k), 32)
    t.insert(to_dev(k), to_dev(v), status_out=st)
    s = from_dev(st, np.uint32)
    oom = (s & S_OOM) != 0
    assert set(np.unique(s & ~np.uint32(S_OOM))) <= {S_NONE, S_DUP}
    assert t.num_keys() == len(keys)
    stored = np.sum(~oom)
    assert t.num_values() == stored and t.value_cells_used() >= 300
    total, b, e, vals, _ = _retrieve(t, keys, stored + 8, 32)
    assert total == stored
    exp = _expected(k, v)
    empty_keys = 0
    for key, lo, hi in zip(keys.tolist(), b, e):
        got = vals[int(lo):int(hi)].tolist()
        assert len(set(got)) == len(got) and set(got) <= set(exp[key])
        n_ok = int(np.sum(~oom[k == key]))
        assert len(got) == n_ok
        empty_keys += n_ok == 0
    assert 0 < empty_keys < len(keys)   # some keys got no first bucket, others did


def test_bucketlist_max_values_per_key(cuda, oracle):
    from warpcore_amd.tables import BucketListHashTable
    keys, mult, k, v = _pairs(2000, 64, 6.0, seed=13)
    t = BucketListHashTable(4000, 4 * len(k), 64, 64, max_values_per_key=4)
    assert t.max_values_per_key() == 4
    st = dev_empty(len(k), 32)
    t.insert(to_dev(k), to_dev(v), status_out=st)
    s = from_dev(st, np.uint32)
    stored = np.minimum(mult, 4)
    assert t.num_values() == int(stored.sum())
    assert np.sum((s & S_MAXV) != 0) == int((mult - stored).sum())
    per = dev_empty(len(keys), 64)
    assert t.num_values(to_dev(keys), per) == int(stored.sum())
    np.testing.assert_array_equal(from_dev(per, np.uint64), stored.astype(np.uint64))
    big = BucketListHashTable(10, 10, 64, 64, max_values_per_key=1 << 30)
    assert big.max_values_per_key() == (1 << 20) - 1


def test_bucketlist_invalid_configuration(cuda):
    from warpcore_amd.tables import BucketListHashTable
    for kw in ({"grow_factor": 0.9}, {"min_bucket_size": 0}, {"min_bucket_size": 8, "max_bucket_size": 4},
               {"max_bucket_size": 1024}):
        t = BucketListHashTable(100, 1000, 32, 32, **kw)
        assert t.peek_status() & S_INVALID_CONFIG
        k = to_dev(np.arange(1, 11, dtype=np.uint32))
        t.insert(k, k)
        assert t.num_values() == 0


def test_bucketlist_dry_run_retrieve_all_and_densities(cuda):
    from warpcore_amd.tables import BucketListHashTable
    keys, mult, k, v = _pairs(3000, 32, 7.0, seed=17)
    t = BucketListHashTable(4000, 4 * len(k), 32, 32)
    t.insert(to_dev(k), to_dev(v))
    end = dev_empty(len(keys), 64)
    st = dev_empty(len(keys), 32)
    assert t.retrieve(to_dev(keys), None, end, None, status_out=st) == len(k)
    np.testing.assert_array_equal(from_dev(end, np.uint64), mult.astype(np.uint64))
    assert np.all(from_dev(st, np.uint32) == S_DRY_RUN)
    assert t.peek_status() & S_DRY_RUN
    ko = dev_empty(len(keys), 32)
    begin, end, vals = dev_empty(len(keys), 64), dev_empty(len(keys), 64), dev_empty(len(k), 32)
    nk, nv = t.retrieve_all(ko, begin, end, vals)
    assert (nk, nv) == (len(keys), len(k))
    got_keys = from_dev(ko, np.uint32)
    b, e, vv = from_dev(begin, np.uint64), from_dev(end, np.uint64), from_dev(vals, np.uint32)
    seg = {int(key): sorted(vv[int(lo):int(hi)].tolist()) for key, lo, hi in zip(got_keys, b, e)}
    assert seg == _expected(k, v)
    payload = len(keys) * 4 + len(k) * 4
    assert t.storage_density() == pytest.approx(payload / t.bytes_total(), rel=1e-5)
    rel = payload / (t.value_cells_used() * 8 + t.key_capacity() * (4 + 8))
    assert t.relative_storage_density() == pytest.approx(rel, rel=1e-5)
    assert t.value_load_factor() == pytest.approx(t.value_cells_used() / t.value_capacity(), rel=1e-6)
    t.init()
    assert t.num_keys() == 0 and t.num_values() == 0 and t.value_cells_used() == 0
