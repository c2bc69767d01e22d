"""CountingHashTable, HashSet and the input generators on the GPU vs the oracle."""
import ctypes

import numpy as np
import pytest
import torch

from datagen import dev_empty, dtype, from_dev, to_dev, tomb, unique_keys

pytestmark = pytest.mark.gpu

S_NONE, S_PLE, S_INVALID_KEY, S_DUP, S_NOT_FOUND = 0, 2, 8, 16, 32


def _lib():
    from warpcore_amd import _capi
    return _capi.lib()


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _zipf_keys(n, distinct, s, seed, bits=64):
    out = dev_empty(n, bits)
    fn = _lib().wc_gen_zipf_u64 if bits == 64 else _lib().wc_gen_zipf_u32
    assert fn(ctypes.c_void_p(out.data_ptr()), n, distinct, s, seed, 0x2000, _stream()) == 0
    torch.cuda.synchronize()
    return from_dev(out, dtype(bits))


@pytest.fixture(params=[1, 0], ids=["combine", "each"])
def combine(request):
    lib = _lib()
    saved = lib.wc_get_counting_combine()
    lib.wc_set_counting_combine(request.param)
    yield request.param
    lib.wc_set_counting_combine(saved)


@pytest.mark.parametrize("kb", [64, 32])
def test_counting_status_word_without_status_output(cuda, combine, oracle, kb):
    """Inserts with no per-occurrence statuses still join duplicate_key
    into the table's status word exactly when the oracle does: not for a
    batch of distinct new keys, yes for repeats inside a batch (combined
    into one entry of several occurrences) and for keys already stored."""
    from warpcore_amd.tables import CountingHashTable
    cap = 1 << 15
    t = CountingHashTable(cap, kb)
    o = oracle.Counting(cap, kb)
    fresh = unique_keys(5000, kb, seed=81)
    t.insert(to_dev(fresh))
    o.insert(fresh)
    assert t.peek_status() == o.status() == S_NONE
    more = unique_keys(3000, kb, seed=82, exclude=fresh)
    rep = np.concatenate([more, more[:100]])          # 100 keys twice in one batch
    t.insert(to_dev(rep))
    o.insert(rep)
    assert t.peek_status() == o.status() == S_DUP
    t.pop_status()
    assert t.peek_status() == S_NONE
    again = fresh[:50]                                # stored keys, once each
    t.insert(to_dev(again))
    assert t.peek_status() == S_DUP
    out = dev_empty(len(again), 64)
    t.retrieve(to_dev(again), out)
    np.testing.assert_array_equal(from_dev(out, np.uint64), np.full(len(again), 2, np.uint64))


@pytest.mark.parametrize("kb,cg,s", [(64, 4, 1.1), (64, 8, 1.0), (32, 4, 1.1), (64, 16, 0.8), (64, 64, 1.1)])
def test_counting_zipf_exact_counts(cuda, shape, combine, oracle, kb, cg, s):
    from warpcore_amd.tables import CountingHashTable
    n, distinct = 1 << 20, 1 << 14
    print(f"[counting] kb={kb} cg={cg} s={s} combine={combine}", flush=True)
    keys = _zipf_keys(n, distinct, s, seed=0xC3 + cg, bits=kb)
    uniq, cnt = np.unique(keys, return_counts=True)
    t = CountingHashTable(-(-distinct * 10 // 9), kb, cg_size=cg)
    o = oracle.Counting(-(-distinct * 10 // 9), kb, cg=cg)
    assert t.capacity() == o.capacity()
    st = dev_empty(n, 32)
    t.insert(to_dev(keys), status_out=st)
    s_gpu = from_dev(st, np.uint32)
    s_ora = o.insert(keys)
    # exactly one `none` per distinct key, all others duplicates (which one is unspecified)
    assert set(np.unique(s_gpu)) <= {S_NONE, S_DUP}
    assert np.sum(s_gpu == S_NONE) == np.sum(s_ora == S_NONE) == len(uniq)
    assert len(np.unique(keys[s_gpu == S_NONE])) == len(uniq)
    assert t.size() == o.size() == len(uniq)
    assert t.peek_status() == o.status()
    absent = unique_keys(1000, kb, seed=5, exclude=uniq)
    q = np.concatenate([uniq, absent])
    out = dev_empty(len(q), 64, 0x1234)
    qst = dev_empty(len(q), 32)
    t.retrieve(to_dev(q), out, status_out=qst)
    got, gst = from_dev(out, np.uint64), from_dev(qst, np.uint32)
    exp, est = o.retrieve(q)
    np.testing.assert_array_equal(gst, est)
    np.testing.assert_array_equal(got[: len(uniq)], cnt.astype(np.uint64))
    np.testing.assert_array_equal(got[: len(uniq)], exp[: len(uniq)])
    assert np.all(got[len(uniq):] == 0x1234)   # misses leave the output untouched


def test_counting_max_count_and_retrieve_all(cuda, shape, combine, oracle):
    from warpcore_amd.tables import CountingHashTable
    keys = np.repeat(np.arange(1, 501, dtype=np.uint64), np.arange(1, 501) % 9 + 1)
    np.random.default_rng(3).shuffle(keys)
    t = CountingHashTable(4000, 64, max_count=5)
    o = oracle.Counting(4000, 64, cg=4, max_count=5)
    t.insert(to_dev(keys)); o.insert(keys)
    q = np.arange(1, 501, dtype=np.uint64)
    out = dev_empty(500, 64)
    t.retrieve(to_dev(q), out)
    exp, _ = o.retrieve(q)
    np.testing.assert_array_equal(from_dev(out, np.uint64), exp)
    assert np.all(exp == np.minimum(np.arange(1, 501) % 9 + 1, 5))
    ko, vo = dev_empty(500, 64), dev_empty(500, 64)
    assert t.retrieve_all(ko, vo) == 500
    k, v = from_dev(ko, np.uint64), from_dev(vo, np.uint64)
    order = np.argsort(k)
    np.testing.assert_array_equal(k[order], q)
    # retrieve_all reports the raw counters: exact below max_count; above it
    # concurrent increments may overshoot before the guard sees the cap
    raw = np.arange(1, 501) % 9 + 1
    below = raw <= 5
    np.testing.assert_array_equal(v[order][below], raw[below])
    assert np.all((v[order][~below] >= 5) & (v[order][~below] <= raw[~below]))
    # erase
    st = dev_empty(3, 32)
    t.erase(to_dev(np.array([1, 2, 10**6], np.uint64)), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), [S_NONE, S_NONE, S_NOT_FOUND])
    assert t.size() == 498


@pytest.mark.parametrize("kb,cg", [(32, 16), (64, 16), (32, 1), (32, 64), (64, 8)])
def test_hashset_parity(cuda, shape, oracle, kb, cg):
    from warpcore_amd.tables import HashSet
    n = 1 << 17
    keys = unique_keys(n, kb, seed=51)
    batch = np.concatenate([keys, keys[: n // 4], np.array([0, tomb(kb)], dtype(kb))])
    np.random.default_rng(52).shuffle(batch)
    t = HashSet(2 * n, kb, cg_size=cg)
    o = oracle.HashSet(2 * n, kb, cg=cg)
    assert t.capacity() == o.capacity()
    st = dev_empty(len(batch), 32)
    t.insert(to_dev(batch), status_out=st)
    s_gpu, s_ora = from_dev(st, np.uint32), o.insert(batch)
    assert np.sum(s_gpu == S_NONE) == np.sum(s_ora == S_NONE) == n
    assert np.sum(s_gpu == S_DUP) == np.sum(s_ora == S_DUP) == n // 4
    np.testing.assert_array_equal(s_gpu[s_ora == S_INVALID_KEY], S_INVALID_KEY)
    assert t.size() == o.size() == n
    assert t.peek_status() == o.status() == S_DUP | S_INVALID_KEY
    absent = unique_keys(n // 2, kb, seed=53, exclude=keys)
    q = np.concatenate([keys[::3], absent, np.array([0], dtype(kb))])
    f = dev_empty(len(q), 8, 7)
    qs = dev_empty(len(q), 32)
    t.contains(to_dev(q), f, status_out=qs)
    ef, es = o.contains(q)
    np.testing.assert_array_equal(from_dev(qs, np.uint32), es)
    gf = from_dev(f, np.uint8)
    np.testing.assert_array_equal(gf[es == S_NONE], ef[es == S_NONE])
    assert gf[-1] == 7   # invalid key: flag left untouched
    gone = keys[: n // 2]
    st = dev_empty(len(gone), 32)
    t.erase(to_dev(gone), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.erase(gone))
    assert t.size() == o.size() == n - n // 2
    ko = dev_empty(n, kb)
    m = t.retrieve_all(ko)
    assert m == n - n // 2
    np.testing.assert_array_equal(np.sort(from_dev(ko, dtype(kb))[:m]), np.sort(keys[n // 2:]))


@pytest.mark.parametrize("kb,cg", [(32, 16), (64, 8), (32, 4)])
def test_hashset_high_load(cuda, shape, oracle, kb, cg):
    """Membership at high load, where most absent keys walk several full
    windows: flags and statuses equal the oracle's at load 0.97, after
    erasing a third and re-inserting some of it (tombstones are never
    reused, so the re-inserts take EMPTY slots further along), and once the
    last EMPTY slot is taken (every absent key then walks the whole cycle and
    reports probing_length_exceeded)."""
    from warpcore_amd.tables import HashSet
    print(f"[hashset high load] kb={kb} cg={cg} shape={shape}", flush=True)
    cap = oracle.get_valid_capacity(1 << 16, cg)
    n = int(cap * 0.97)
    keys = unique_keys(n, kb, seed=71)
    absent = unique_keys(n, kb, seed=72, exclude=keys)
    t = HashSet(1 << 16, kb, cg_size=cg)
    o = oracle.HashSet(1 << 16, kb, cg=cg)
    assert t.capacity() == cap

    def check(q):
        f = dev_empty(len(q), 8, 7)
        qs = dev_empty(len(q), 32)
        t.contains(to_dev(q), f, status_out=qs)
        ef, es = o.contains(q)
        np.testing.assert_array_equal(from_dev(qs, np.uint32), es)
        gf = from_dev(f, np.uint8)
        np.testing.assert_array_equal(gf[es == S_NONE], ef[es == S_NONE])
        return es

    st = dev_empty(n, 32)
    t.insert(to_dev(keys), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(keys))
    q = np.concatenate([keys[::2], absent])
    check(q)
    er = keys[::3]
    est = dev_empty(len(er), 32)
    t.erase(to_dev(er), status_out=est)
    np.testing.assert_array_equal(from_dev(est, np.uint32), o.erase(er))
    back = er[: (cap - n) // 2]
    st = dev_empty(len(back), 32)
    t.insert(to_dev(back), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(back))
    check(q)
    assert t.peek_status() == o.status()
    rest = unique_keys(cap - n - len(back), kb, seed=73, exclude=np.concatenate([keys, absent]))
    st = dev_empty(len(rest), 32)
    t.insert(to_dev(rest), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(rest))
    assert t.size() == o.size()
    assert np.all(check(absent[:64]) == S_PLE)


@pytest.mark.parametrize("bits", [32, 64])
def test_unique_generator(cuda, oracle, bits):
    n = 1 << 20
    out = dev_empty(n, bits)
    fn = _lib().wc_gen_unique_u64 if bits == 64 else _lib().wc_gen_unique_u32
    assert fn(ctypes.c_void_p(out.data_ptr()), n, 12345, _stream()) == 0
    k = from_dev(out, dtype(bits))
    assert len(np.unique(k)) == n
    assert not np.any(k == 0) and not np.any(k == dtype(bits)(tomb(bits)))
    f = oracle.fmix64 if bits == 64 else oracle.fmix32
    m = (1 << bits) - 1
    for i in [0, 1, 777, n - 1]:
        assert int(k[i]) == f((12345 + i) & m)


def test_zipf_generator_distribution(cuda):
    n, d, s = 1 << 22, 1 << 12, 1.1
    keys = _zipf_keys(n, d, s, seed=99)
    ranks_key = np.array([_lib().wc_unique_key_u64(0x2000, r) for r in range(d)], np.uint64)
    assert np.all(np.isin(keys, ranks_key))
    counts = np.array([np.sum(keys == ranks_key[r]) for r in range(8)])
    pmf = (np.arange(1, d + 1) ** -s)
    pmf /= pmf.sum()
    exp = pmf[:8] * n
    assert np.all(np.abs(counts - exp) < 6 * np.sqrt(exp))
