"""SingleValueHashTable on the GPU vs the oracle (bit-exact).

Mirrors tests/src/single_value_hash_table_test.cu (size after insert, status
none, retrieved == inserted, for_each count == size) and extends it with the
oracle's per-key statuses, table status words, erase, invalid keys,
duplicates, probing-length exhaustion and retrieve_all.
"""
import numpy as np
import pytest
import torch

from datagen import dev_empty, dtype, from_dev, to_dev, tomb, uniform_values, unique_keys

pytestmark = pytest.mark.gpu

S_NONE, S_PLE, S_INVALID_KEY, S_DUP, S_NOT_FOUND, S_NOT_INIT = 0, 2, 8, 16, 32, 256
SENTINEL = 0xA5A5A5A5


def _tables(oracle, min_cap, kb, vb, cg, seed=0x5AD0DED):
    from warpcore_amd.tables import SingleValueHashTable
    t = SingleValueHashTable(min_cap, kb, vb, seed=seed, cg_size=cg)
    o = oracle.SVHT(min_cap, kb, vb, cg=cg, seed=seed)
    return t, o


def _retrieve(t, keys, vb, pl=None):
    kd = to_dev(keys)
    out = dev_empty(len(keys), vb, SENTINEL)
    st = dev_empty(len(keys), 32)
    if pl is None:
        t.retrieve(kd, out, status_out=st)
    else:
        t.retrieve(kd, out, status_out=st, probing_length=pl)
    torch.cuda.synchronize()
    return from_dev(out, dtype(vb)), from_dev(st, np.uint32)


def _oracle_retrieve(o, keys, vb, pl=None):
    out = np.full(len(keys), SENTINEL, dtype(vb))
    if pl is None:
        _, st = o.retrieve(keys, out)
    else:
        _, st = o.retrieve(keys, out, pl)
    return out, st


@pytest.mark.parametrize("kb,vb,cg,n", [
    (32, 32, 8, 1 << 20),     # configuration 1: 2^20 keys, min capacity 2^20 / 0.8
    (32, 32, 16, 1 << 18),
    (32, 32, 1, 1 << 16), (32, 32, 2, 1 << 16), (32, 32, 4, 1 << 16),
    (32, 32, 32, 1 << 16), (32, 32, 64, 1 << 16),
    (64, 64, 8, 1 << 18), (64, 64, 4, 1 << 16), (64, 64, 16, 1 << 16), (64, 64, 64, 1 << 16),
    (32, 64, 8, 1 << 16), (64, 32, 8, 1 << 16),
])
def test_insert_retrieve_parity(cuda, shape, oracle, kb, vb, cg, n):
    print(f"[svht parity] kb={kb} vb={vb} cg={cg} n={n}", flush=True)
    keys = unique_keys(n, kb, seed=0xC1)
    vals = uniform_values(n, vb, seed=0xC2)
    absent = unique_keys(n, kb, seed=0xC3, exclude=keys)
    min_cap = -(-n * 10 // 8)
    t, o = _tables(oracle, min_cap, kb, vb, cg)
    assert t.capacity() == o.capacity() == oracle.get_valid_capacity(min_cap, cg)
    assert t.size() == 0 and t.peek_status() == S_NONE

    st = dev_empty(n, 32, 0xFFFFFFFF)
    t.insert(to_dev(keys), to_dev(vals), status_out=st)
    st_o = o.insert(keys, vals)
    np.testing.assert_array_equal(from_dev(st, np.uint32), st_o)
    assert t.size() == o.size() == n
    assert t.peek_status() == o.status() == S_NONE

    query = np.concatenate([keys, absent])
    np.random.default_rng(7).shuffle(query)
    got, gst = _retrieve(t, query, vb)
    exp, est = _oracle_retrieve(o, query, vb)
    np.testing.assert_array_equal(gst, est)
    np.testing.assert_array_equal(got, exp)
    assert t.peek_status() == o.status() == S_NOT_FOUND
    assert t.pop_status() == S_NOT_FOUND
    assert t.peek_status() == S_NONE
    assert abs(t.load_factor() - n / t.capacity()) < 1e-6


@pytest.mark.parametrize("kb,vb,cg", [(32, 32, 8), (64, 64, 8), (32, 32, 16)])
def test_duplicates_one_none_per_key(cuda, shape, oracle, kb, vb, cg):
    n_keys = 1 << 14
    keys = unique_keys(n_keys, kb, seed=11)
    rng = np.random.default_rng(12)
    mult = rng.integers(1, 5, size=n_keys)
    batch = np.repeat(keys, mult)
    rng.shuffle(batch)
    vals = uniform_values(len(batch), vb, seed=13)
    t, _ = _tables(oracle, 2 * n_keys, kb, vb, cg)
    st = dev_empty(len(batch), 32)
    t.insert(to_dev(batch), to_dev(vals), status_out=st)
    s = from_dev(st, np.uint32)
    assert set(np.unique(s)) <= {S_NONE, S_DUP}
    # exactly one `none` per distinct key
    none_keys = batch[s == S_NONE]
    assert len(none_keys) == n_keys and len(np.unique(none_keys)) == n_keys
    assert t.size() == n_keys
    assert t.peek_status() == (S_DUP if len(batch) > n_keys else S_NONE)
    # the stored value is one of the values inserted with that key
    got, gst = _retrieve(t, keys, vb)
    assert np.all(gst == S_NONE)
    order = np.argsort(batch, kind="stable")
    sb, sv = batch[order], vals[order]
    lo = np.searchsorted(sb, keys, "left")
    hi = np.searchsorted(sb, keys, "right")
    ok = [got[i] in set(sv[lo[i]:hi[i]].tolist()) for i in range(0, n_keys, 7)]
    assert all(ok)


@pytest.mark.parametrize("kb,vb,cg", [(32, 32, 8), (64, 64, 8)])
def test_erase_and_reinsert(cuda, shape, oracle, kb, vb, cg):
    n = 1 << 16
    keys = unique_keys(n, kb, seed=21)
    vals = uniform_values(n, vb, seed=22)
    absent = unique_keys(1 << 12, kb, seed=23, exclude=keys)
    t, o = _tables(oracle, 2 * n, kb, vb, cg)
    t.insert(to_dev(keys), to_dev(vals)); o.insert(keys, vals)
    gone = np.concatenate([keys[: n // 3], absent])
    st = dev_empty(len(gone), 32)
    t.erase(to_dev(gone), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.erase(gone))
    assert t.size() == o.size() == n - n // 3
    # erase misses are not joined into the table status
    assert t.peek_status() == o.status() == S_NONE
    got, gst = _retrieve(t, keys, vb)
    exp, est = _oracle_retrieve(o, keys, vb)
    np.testing.assert_array_equal(gst, est)
    np.testing.assert_array_equal(got, exp)
    # re-insert erased keys with new values: tombstones are not reused but the keys come back
    t.pop_status(); o.clear_status()
    new_vals = uniform_values(n // 3, vb, seed=24)
    st = dev_empty(n // 3, 32)
    t.insert(to_dev(keys[: n // 3]), to_dev(new_vals), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(keys[: n // 3], new_vals))
    got, gst = _retrieve(t, keys, vb)
    exp, est = _oracle_retrieve(o, keys, vb)
    np.testing.assert_array_equal(gst, est)
    np.testing.assert_array_equal(got, exp)
    assert t.size() == o.size() == n


@pytest.mark.parametrize("kb", [32, 64])
def test_invalid_keys(cuda, shape, oracle, kb):
    keys = np.array([0, tomb(kb), 5, 0, 7], dtype(kb))
    vals = np.arange(5, dtype=dtype(kb))
    t, o = _tables(oracle, 1000, kb, kb, 8)
    st = dev_empty(5, 32)
    t.insert(to_dev(keys), to_dev(vals), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(keys, vals))
    assert t.peek_status() == o.status() == S_INVALID_KEY
    got, gst = _retrieve(t, keys, kb)
    exp, est = _oracle_retrieve(o, keys, kb)
    np.testing.assert_array_equal(gst, est)
    np.testing.assert_array_equal(got, exp)
    st = dev_empty(5, 32)
    t.erase(to_dev(keys), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.erase(keys))
    assert t.size() == o.size() == 0


@pytest.mark.parametrize("kb,vb,cg", [(32, 32, 1), (64, 64, 2)])
def test_full_table_probing_length_exceeded(cuda, shape, oracle, kb, vb, cg):
    """A table with no EMPTY slot: every new key and every miss exhausts probing.

    The reference would probe forever here (default probing length ~2^64);
    both the oracle and the GPU stop after one full cycle of windows.
    """
    cap = oracle.get_valid_capacity(1, cg)
    keys = unique_keys(cap, kb, seed=31)
    vals = uniform_values(cap, vb, seed=32)
    t, o = _tables(oracle, 1, kb, vb, cg)
    assert t.capacity() == cap
    st = dev_empty(cap, 32)
    t.insert(to_dev(keys), to_dev(vals), status_out=st)
    assert np.all(from_dev(st, np.uint32) == S_NONE)
    o.insert(keys, vals)
    assert t.size() == cap and t.load_factor() == 1.0
    extra = unique_keys(64, kb, seed=33, exclude=keys)
    xv = uniform_values(64, vb, seed=34)
    for pl in (None, cg, 5 * cg + 1):
        st = dev_empty(64, 32)
        if pl is None:
            t.insert(to_dev(extra), to_dev(xv), status_out=st)
            est = o.insert(extra, xv)
        else:
            t.insert(to_dev(extra), to_dev(xv), status_out=st, probing_length=pl)
            est = o.insert(extra, xv, pl)
        np.testing.assert_array_equal(from_dev(st, np.uint32), est)
        assert np.all(est == S_PLE)
        got, gst = _retrieve(t, extra, vb, pl)
        exp, ost = _oracle_retrieve(o, extra, vb, pl)
        np.testing.assert_array_equal(gst, ost)
        assert np.all(ost == S_PLE)
    assert t.peek_status() == o.status() == S_PLE
    # every stored key is still found
    got, gst = _retrieve(t, keys, vb)
    assert np.all(gst == S_NONE)
    np.testing.assert_array_equal(got, vals)


def test_no_init_is_inert(cuda):
    from warpcore_amd.tables import SingleValueHashTable
    t = SingleValueHashTable(1000, 32, 32, no_init=True)
    assert not t.is_initialized()
    keys = to_dev(np.arange(1, 101, dtype=np.uint32))
    st = dev_empty(100, 32, 0x77)
    t.insert(keys, keys, status_out=st)
    torch.cuda.synchronize()
    assert np.all(from_dev(st, np.uint32) == 0x77)  # nothing launched
    assert t.size() == 0
    t.init()
    assert t.is_initialized()
    t.insert(keys, keys)
    assert t.size() == 100


@pytest.mark.parametrize("kb,vb", [(32, 32), (64, 64)])
def test_retrieve_all_and_size(cuda, shape, oracle, kb, vb):
    n = 50000
    keys = unique_keys(n, kb, seed=41)
    vals = uniform_values(n, vb, seed=42)
    t, _ = _tables(oracle, 3 * n, kb, vb, 8)
    t.insert(to_dev(keys), to_dev(vals))
    ko = dev_empty(n, kb)
    vo = dev_empty(n, vb)
    m = t.retrieve_all(ko, vo)
    assert m == n == t.size()
    k, v = from_dev(ko, dtype(kb)), from_dev(vo, dtype(vb))
    order = np.argsort(k)
    ref = np.argsort(keys)
    np.testing.assert_array_equal(k[order], keys[ref])
    np.testing.assert_array_equal(v[order], vals[ref])


def test_bool_handler(cuda, shape, oracle):
    keys = np.array([3, 0, 9, 3], np.uint32)
    vals = np.array([1, 2, 3, 4], np.uint32)
    t, o = _tables(oracle, 1000, 32, 32, 8)
    f = dev_empty(4, 8, 0x55)
    t.insert_bool(to_dev(keys), to_dev(vals), f)
    flags = from_dev(f, np.uint8)
    # ReturnBoolean<> writes has_any(~none()) with the reference's
    # operator~ (status.cuh:146-150,316-333): ~none() is 0xFFFFF000, so
    # invalid_key and duplicate_key leave the flag false
    np.testing.assert_array_equal(flags, oracle.return_boolean(o.insert(keys, vals)))
    assert flags.tolist() == [0, 0, 0, 0]


@pytest.mark.parametrize("kb,vb,cg", [(64, 64, 8), (32, 32, 16), (64, 64, 4)])
def test_overflow_filter_lookups(cuda, shape, oracle, kb, vb, cg):
    """Lookups at high load, where most misses meet full windows: the
    lane-owner lookups end such a miss at the first full window whose
    overflow-filter bit is clear. Statuses and values equal the oracle's full
    walk at load 0.97, after erasing a third and re-inserting some of it
    (tombstones are never reused, so the re-inserts take EMPTY slots further
    along), and once the last EMPTY slot is taken (no EMPTY slot: every miss
    walks the whole cycle and reports probing_length_exceeded)."""
    print(f"[svht filter] kb={kb} vb={vb} cg={cg} shape={shape}", flush=True)
    cap = oracle.get_valid_capacity(1 << 16, cg)
    n = int(cap * 0.97)
    keys = unique_keys(n, kb, seed=61)
    vals = uniform_values(n, vb, seed=62)
    absent = unique_keys(n, kb, seed=63, exclude=keys)
    t, o = _tables(oracle, 1 << 16, kb, vb, cg)
    st = dev_empty(n, 32)
    t.insert(to_dev(keys), to_dev(vals), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(keys, vals))
    q = np.concatenate([keys[::2], absent])
    got, gst = _retrieve(t, q, vb)
    exp, ost = _oracle_retrieve(o, q, vb)
    np.testing.assert_array_equal(gst, ost)
    np.testing.assert_array_equal(got, exp)
    # erase a third, re-insert some of those with new values (fewer than the
    # EMPTY slots left)
    er = keys[::3]
    est = dev_empty(len(er), 32)
    t.erase(to_dev(er), status_out=est)
    np.testing.assert_array_equal(from_dev(est, np.uint32), o.erase(er))
    free = cap - n
    back = er[: free // 2]
    bv = uniform_values(len(back), vb, seed=64)
    st = dev_empty(len(back), 32)
    t.insert(to_dev(back), to_dev(bv), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(back, bv))
    got, gst = _retrieve(t, q, vb)
    exp, ost = _oracle_retrieve(o, q, vb)
    np.testing.assert_array_equal(gst, ost)
    np.testing.assert_array_equal(got, exp)
    assert t.peek_status() == o.status()
    # take the last EMPTY slots: no lookup may end early any more
    rest = unique_keys(free - len(back), kb, seed=65, exclude=np.concatenate([keys, absent]))
    rv = uniform_values(len(rest), vb, seed=66)
    st = dev_empty(len(rest), 32)
    t.insert(to_dev(rest), to_dev(rv), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(rest, rv))
    assert t.size() == o.size()
    got, gst = _retrieve(t, absent[:64], vb)
    exp, ost = _oracle_retrieve(o, absent[:64], vb)
    np.testing.assert_array_equal(gst, ost)
    assert np.all(ost == S_PLE)
