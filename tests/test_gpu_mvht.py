"""MultiValueHashTable on the GPU vs the oracle: per-key counts and value
multisets, statuses, the two-pass retrieve offsets, dry runs, retrieve_all,
max_values_per_key and probe-order parity on sequential inserts."""
import numpy as np
import pytest
import torch

from datagen import dev_empty, dtype, from_dev, kernel_shape, to_dev, unique_keys

pytestmark = pytest.mark.gpu

# bulk inserts and lookups: one key per group at a time ("groups") or
# several keys in flight per group ("refill", the default)
BULK_KERNELS = pytest.mark.parametrize("kernels", ["groups", "refill", "pipe", "owner"])

S_NONE, S_PLE, S_INVALID_KEY, S_DUP, S_NOT_FOUND = 0, 2, 8, 16, 32
S_DRY_RUN, S_PHASE, S_MAXV = 1 << 9, 1 << 10, 1 << 11


def _pairs(num_keys, bits, mean, seed):
    """Distinct keys with geometric multiplicities and distinct values."""
    rng = np.random.default_rng(seed)
    keys = unique_keys(num_keys, bits, seed)
    mult = rng.geometric(1.0 / mean, size=num_keys)
    k = np.repeat(keys, mult)
    v = unique_keys(len(k), bits, seed + 1)
    perm = rng.permutation(len(k))
    return keys, mult, k[perm], v[perm]


def _segments(begin, end, vals, keys):
    return {int(key): sorted(vals[int(b):int(e)].tolist()) for key, b, e in zip(keys, begin, end)}


def _expected(k, v):
    out = {}
    for key, val in zip(k.tolist(), v.tolist()):
        out.setdefault(key, []).append(val)
    return {key: sorted(vals) for key, vals in out.items()}


@BULK_KERNELS
@pytest.mark.parametrize("kb,cg,mean", [(32, 8, 8.0), (64, 8, 4.0), (32, 4, 2.0), (64, 16, 8.0), (32, 1, 3.0),
                                        (32, 64, 8.0)])
def test_mvht_insert_retrieve_matches_oracle(cuda, oracle, kb, cg, mean, kernels):
    from warpcore_amd.tables import MultiValueHashTable
    print(f"[mvht] kb={kb} cg={cg} mean={mean} {kernels}", flush=True)
    keys, mult, k, v = _pairs(6000, kb, mean, seed=11 + cg)
    cap = int(len(k) / 0.8)
    t = MultiValueHashTable(cap, kb, kb, cg_size=cg)
    o = oracle.MVHT(cap, kb, kb, cg=cg)
    assert t.capacity() == o.capacity()
    st = dev_empty(len(k), 32, 0x77)
    with kernel_shape(kernels):
        t.insert(to_dev(k), to_dev(v), status_out=st)
        s_gpu = from_dev(st, np.uint32)
        s_ora = o.insert(k, v)
        # one `none` per distinct key, duplicates for the rest (which pair is first is unspecified)
        assert set(np.unique(s_gpu)) <= {S_NONE, S_DUP}
        assert np.sum(s_gpu == S_NONE) == np.sum(s_ora == S_NONE) == len(keys)
        assert len(np.unique(k[s_gpu == S_NONE])) == len(keys)
        assert t.num_keys() == o.num_keys() == len(keys)
        assert t.size() == o.size() == len(k)
        assert t.peek_status() == o.status() == S_NONE
        absent = unique_keys(500, kb, seed=99, exclude=keys)
        q = np.concatenate([keys, absent])
        rng = np.random.default_rng(5)
        q = q[rng.permutation(len(q))]
        n = len(q)
        begin, end = dev_empty(n, 64), dev_empty(n, 64)
        vals = dev_empty(len(k) + 8, kb, 0xAB)
        qst = dev_empty(n, 32, 0x77)
        total = t.retrieve(to_dev(q), begin, end, vals, status_out=qst)
        ob, oe, ov, ost = o.retrieve(q)
        assert total == len(k) == len(ov)
        b, e = from_dev(begin, np.uint64), from_dev(end, np.uint64)
        np.testing.assert_array_equal(b, ob)        # offsets depend only on counts
        np.testing.assert_array_equal(e, oe)
        np.testing.assert_array_equal(from_dev(qst, np.uint32), ost)
        got = from_dev(vals, dtype(kb))
        assert np.all(got[len(k):] == np.array(0xAB, dtype(kb)))   # nothing written past the total
        seg = _segments(b, e, got, q)
        exp = _expected(k, v)
        for key in keys.tolist():
            assert seg[key] == exp[key]
        for key in absent.tolist():
            assert seg[key] == []
        assert t.peek_status() == S_NOT_FOUND


@pytest.mark.parametrize("kb", [32, 64])
def test_mvht_probe_order_parity_on_sequential_inserts(cuda, oracle, kb):
    """One pair per launch fixes the placement: values come back in the
    oracle's probe order, not just as the same multiset."""
    from warpcore_amd.tables import MultiValueHashTable
    keys, mult, k, v = _pairs(40, kb, 4.0, seed=3)
    t = MultiValueHashTable(400, kb, kb, cg_size=4)
    o = oracle.MVHT(400, kb, kb, cg=4)
    for i in range(len(k)):
        st = dev_empty(1, 32, 0x77)
        t.insert(to_dev(k[i:i + 1]), to_dev(v[i:i + 1]), status_out=st)
        assert from_dev(st, np.uint32)[0] == o.insert(k[i:i + 1], v[i:i + 1])[0]
    n = len(keys)
    begin, end, vals = dev_empty(n, 64), dev_empty(n, 64), dev_empty(len(k), kb)
    assert t.retrieve(to_dev(keys), begin, end, vals) == len(k)
    ob, oe, ov, _ = o.retrieve(keys)
    np.testing.assert_array_equal(from_dev(vals, dtype(kb)), ov)


@BULK_KERNELS
def test_mvht_max_values_per_key(cuda, oracle, kernels):
    from warpcore_amd.tables import MultiValueHashTable
    keys, mult, k, v = _pairs(3000, 32, 6.0, seed=8)
    t = MultiValueHashTable(len(k) * 2, 32, 32, max_values_per_key=3)
    o = oracle.MVHT(len(k) * 2, 32, 32, cg=8, max_values=3)
    st = dev_empty(len(k), 32)
    with kernel_shape(kernels):
        t.insert(to_dev(k), to_dev(v), status_out=st)
        s_gpu = from_dev(st, np.uint32)
        s_ora = o.insert(k, v)
        stored = np.minimum(mult, 3)
        assert t.size() == o.size() == int(stored.sum())
        assert t.num_keys() == len(keys)
        for code in (S_NONE, S_DUP, S_DUP | S_MAXV):
            assert np.sum(s_gpu == code) == np.sum(s_ora == code)
        assert t.peek_status() == o.status() == S_DUP | S_MAXV
        per = dev_empty(len(keys), 64)
        assert t.num_values(to_dev(keys), per) == int(stored.sum())
        np.testing.assert_array_equal(from_dev(per, np.uint64), stored.astype(np.uint64))
        # each kept value is one of the key's inserted values
        begin, end, vals = dev_empty(len(keys), 64), dev_empty(len(keys), 64), dev_empty(int(stored.sum()), 32)
        t.retrieve(to_dev(keys), begin, end, vals)
        seg = _segments(from_dev(begin, np.uint64), from_dev(end, np.uint64), from_dev(vals, np.uint32), keys)
        exp = _expected(k, v)
        for key in keys.tolist():
            assert set(seg[key]) <= set(exp[key]) and len(seg[key]) == min(len(exp[key]), 3)


def test_mvht_dry_run_and_num_values(cuda, oracle):
    from warpcore_amd.tables import MultiValueHashTable
    keys, mult, k, v = _pairs(2000, 64, 5.0, seed=21)
    t = MultiValueHashTable(len(k) * 2, 64, 64)
    t.insert(to_dev(k), to_dev(v))
    absent = unique_keys(100, 64, seed=2, exclude=keys)
    q = np.concatenate([keys, absent])
    end = dev_empty(len(q), 64, 0x33)
    st = dev_empty(len(q), 32)
    total = t.retrieve(to_dev(q), None, end, None, status_out=st)
    assert total == len(k)
    counts = np.concatenate([mult, np.zeros(100, int)]).astype(np.uint64)
    np.testing.assert_array_equal(from_dev(end, np.uint64), counts)   # raw counts, not scanned
    assert np.all(from_dev(st, np.uint32) == S_DRY_RUN)
    assert t.peek_status() & S_DRY_RUN
    assert t.peek_status() & S_NOT_FOUND
    per = dev_empty(len(q), 64)
    nst = dev_empty(len(q), 32)
    assert t.num_values(to_dev(q), per, status_out=nst) == len(k)
    np.testing.assert_array_equal(from_dev(per, np.uint64), counts)
    ns = from_dev(nst, np.uint32)
    assert np.all(ns[: len(keys)] == S_NONE) and np.all(ns[len(keys):] == S_NOT_FOUND)
    t.pop_status()
    assert t.peek_status() == S_NONE


@pytest.mark.parametrize("kb", [32, 64])
def test_mvht_retrieve_all(cuda, kb):
    from warpcore_amd.tables import MultiValueHashTable
    keys, mult, k, v = _pairs(5000, kb, 8.0, seed=31)
    t = MultiValueHashTable(int(len(k) / 0.9), kb, kb)
    t.insert(to_dev(k), to_dev(v))
    ko = dev_empty(len(keys), kb)
    assert t.retrieve_all_keys(ko) == len(keys)
    assert np.array_equal(np.sort(from_dev(ko, dtype(kb))), np.sort(keys))
    ko2 = dev_empty(len(keys), kb)
    begin, end, vals = dev_empty(len(keys), 64), dev_empty(len(keys), 64), dev_empty(len(k), kb)
    nk, nv = t.retrieve_all(ko2, begin, end, vals)
    assert (nk, nv) == (len(keys), len(k))
    got_keys = from_dev(ko2, dtype(kb))
    seg = _segments(from_dev(begin, np.uint64), from_dev(end, np.uint64), from_dev(vals, dtype(kb)), got_keys)
    assert seg == _expected(k, v)
    assert t.storage_density() == pytest.approx((len(keys) * kb / 8 + len(k) * kb / 8) /
                                                (t.capacity() * 2 * kb / 8), rel=1e-5)


@BULK_KERNELS
def test_mvht_full_table_and_invalid_keys(cuda, oracle, kernels):
    from warpcore_amd.tables import MultiValueHashTable
    t = MultiValueHashTable(256, 32, 32, cg_size=8)
    cap = t.capacity()
    k = np.repeat(unique_keys(cap // 4 + 10, 32, seed=4), 4)
    v = unique_keys(len(k), 32, seed=5)
    st = dev_empty(len(k), 32)
    with kernel_shape(kernels):
        t.insert(to_dev(k), to_dev(v), status_out=st)
    s = from_dev(st, np.uint32)
    assert t.size() == cap                           # every slot taken
    assert np.sum((s & S_PLE) != 0) == len(k) - cap  # the rest ran out of windows
    # a pair out of windows after values of its key also says duplicate_key
    assert set(np.unique(s)) <= {S_NONE, S_DUP, S_PLE, S_PLE | S_DUP}
    assert t.peek_status() & S_PLE
    bad = np.array([0, 0xFFFFFFFF], np.uint32)
    bst = dev_empty(2, 32)
    t2 = MultiValueHashTable(1000, 32, 32)
    with kernel_shape(kernels):
        t2.insert(to_dev(bad), to_dev(np.array([1, 2], np.uint32)), status_out=bst)
    assert np.all(from_dev(bst, np.uint32) == S_INVALID_KEY)
    assert t2.size() == 0 and t2.num_keys() == 0
    assert t2.peek_status() == S_INVALID_KEY


def test_mvht_large_query_offsets(cuda):
    """More than 1024 scan chunks of 2048 counts: offsets stay exact."""
    from warpcore_amd.tables import MultiValueHashTable
    n = (1 << 21) + 12345
    lib_keys = dev_empty(n, 32)
    from warpcore_amd import _capi
    import ctypes
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert _capi.lib().wc_gen_unique_u32(ctypes.c_void_p(lib_keys.data_ptr()), n, 77, stream) == 0
    t = MultiValueHashTable(int(n / 0.85), 32, 32)
    t.insert(lib_keys, lib_keys)
    torch.cuda.synchronize()
    assert t.num_keys() == n and t.size() == n
    begin, end, vals = dev_empty(n, 64), dev_empty(n, 64), dev_empty(n, 32)
    assert t.retrieve(lib_keys, begin, end, vals) == n
    idx = torch.arange(n, device="cuda", dtype=torch.int64)
    assert torch.equal(begin, idx)
    assert torch.equal(end, idx + 1)
    assert torch.equal(vals, lib_keys)              # value = key for every key


def test_mvht_empty_batches(cuda):
    from warpcore_amd.tables import MultiValueHashTable
    t = MultiValueHashTable(1000, 64, 64)
    e64 = torch.empty(0, dtype=torch.int64, device="cuda")
    t.insert(e64, e64)
    assert t.retrieve(e64, e64, e64, e64) == 0
    assert t.num_values(e64) == 0
    assert t.size() == 0 and t.num_keys() == 0
    ko = dev_empty(1, 64)
    assert t.retrieve_all_keys(ko) == 0
