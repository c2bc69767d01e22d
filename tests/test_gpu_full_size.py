"""The configurations at their full sizes, checked through properties that do
not need a CPU replay of the whole run: round trips, counts and multisets
against a plain library routine (torch.unique / torch.sort on the same
device data), the no-false-negative property and the Bloom bit array word
for word against the oracle. Sizes: SVHT<u64,u64> 2^28 keys at load factor
0.9 (4.8 GB), SVHT<u32,u32> 2^20 keys at capacity 2^20/0.8,
CountingHashTable<u64> over 2^28 Zipf(1.1) occurrences of 2^24 keys,
MultiValueHashTable<u32,u32> 2^26 pairs with geometric multiplicity of mean
8, BucketListHashTable<u32,u32> 2^27 values on 2^24 keys, HashSet<u32> 2^27
keys, BloomFilter 2^30 bits with k=8 and 2^27 keys."""
import ctypes
import math

import numpy as np
import pytest
import torch

from datagen import dev_empty, from_dev, to_dev, unique_keys

pytestmark = pytest.mark.gpu

S_NONE, S_NOT_FOUND = 0, 32


def _lib():
    from warpcore_amd import _capi
    return _capi.lib()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _gen_unique(n, bits, offset):
    t = torch.empty(n, dtype=torch.int64 if bits == 64 else torch.int32, device="cuda")
    assert getattr(_lib(), f"wc_gen_unique_u{bits}")(_p(t), n, offset, None) == 0
    return t


def _gen_uniform(n, bits, seed):
    t = torch.empty(n, dtype=torch.int64 if bits == 64 else torch.int32, device="cuda")
    assert getattr(_lib(), f"wc_gen_uniform_u{bits}")(_p(t), n, seed, None) == 0
    return t


def test_full_svht_u64_lf09_round_trip(cuda):
    from warpcore_amd.tables import SingleValueHashTable
    n = 1 << 28
    print("[full] svht u64 2^28 LF 0.9", flush=True)
    keys = _gen_unique(n, 64, 0x1000)
    vals = _gen_uniform(n, 64, 0xC1)
    t = SingleValueHashTable(math.ceil(n / 0.9), 64, 64)
    assert t.capacity() == 298876376
    st = torch.empty(n, dtype=torch.int32, device="cuda")
    t.insert(keys, vals, status_out=st)
    torch.cuda.synchronize()
    assert int((st != S_NONE).sum()) == 0
    assert t.size() == n and int(t.peek_status()) == S_NONE
    out = torch.zeros_like(vals)
    t.retrieve(keys, out, status_out=st)
    torch.cuda.synchronize()
    assert torch.equal(out, vals) and int((st != S_NONE).sum()) == 0
    del out
    absent = _gen_unique(1 << 26, 64, 0x1000 + n + 7)
    out = torch.full((1 << 26,), 0x5A, dtype=torch.int64, device="cuda")
    st2 = torch.empty(1 << 26, dtype=torch.int32, device="cuda")
    t.retrieve(absent, out, status_out=st2)
    torch.cuda.synchronize()
    assert int((st2 != S_NOT_FOUND).sum()) == 0 and int((out != 0x5A).sum()) == 0   # misses write nothing
    # erase every fourth key: size drops by that many, they are gone, the rest remain
    gone = keys[::4].contiguous()
    t.erase(gone)
    assert t.size() == n - gone.numel()
    probe = torch.cat([gone[:1 << 20], keys[1:1 << 22:4]])
    out = torch.zeros_like(probe)
    st3 = torch.empty(probe.numel(), dtype=torch.int32, device="cuda")
    t.retrieve(probe, out, status_out=st3)
    torch.cuda.synchronize()
    assert int((st3[: 1 << 20] != S_NOT_FOUND).sum()) == 0
    assert int((st3[1 << 20:] != S_NONE).sum()) == 0
    assert torch.equal(out[1 << 20:], vals[1:1 << 22:4])


def test_full_svht_u32_config0(cuda):
    from warpcore_amd.tables import SingleValueHashTable
    n = 1 << 20
    print("[full] svht u32 2^20 cap 2^20/0.8", flush=True)
    rng = np.random.default_rng(0xC0)
    keys = unique_keys(n, 32, seed=0xC0)
    vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    absent = unique_keys(n, 32, seed=0xC1, exclude=keys)
    t = SingleValueHashTable(math.ceil(n / 0.8), 32, 32)
    assert t.capacity() == 2097496   # the capacity rule rounds up to a prime number of windows
    t.insert(to_dev(keys), to_dev(vals))
    q = np.concatenate([keys, absent])
    out, st = dev_empty(2 * n, 32, 0x77), dev_empty(2 * n, 32)
    t.retrieve(to_dev(q), out, status_out=st)
    got, gst = from_dev(out, np.uint32), from_dev(st, np.uint32)
    np.testing.assert_array_equal(got[:n], vals)
    assert np.all(gst[:n] == S_NONE) and np.all(gst[n:] == S_NOT_FOUND)
    assert np.all(got[n:] == 0x77)
    assert t.size() == n


def test_full_counting_zipf_exact(cuda):
    from warpcore_amd.tables import CountingHashTable
    n, d = 1 << 28, 1 << 24
    print("[full] counting u64 2^28 Zipf(1.1) over 2^24", flush=True)
    occ = torch.empty(n, dtype=torch.int64, device="cuda")
    assert _lib().wc_gen_zipf_u64(_p(occ), n, d, 1.1, 0xC3, 0x2000, None) == 0
    t = CountingHashTable(math.ceil(d / 0.9), 64)
    t.insert(occ)
    ref_keys, ref_counts = torch.unique(occ, return_counts=True)   # sorted
    del occ
    size = t.size()
    assert size == ref_keys.numel()
    keys_out = torch.empty(size, dtype=torch.int64, device="cuda")
    counts_out = torch.empty(size, dtype=torch.int64, device="cuda")
    assert t.retrieve_all(keys_out, counts_out) == size
    order = torch.argsort(keys_out)
    assert torch.equal(keys_out[order], ref_keys)
    assert torch.equal(counts_out[order], ref_counts)
    assert int(counts_out.sum()) == n
    got = torch.empty_like(ref_keys)
    t.retrieve(ref_keys, got)
    assert torch.equal(got, ref_counts)


def test_full_mvht_geometric_multisets(cuda):
    from warpcore_amd.tables import MultiValueHashTable
    n = 1 << 26
    print("[full] mvht u32 2^26 pairs, geometric mean 8", flush=True)
    rng = np.random.default_rng(0xC4)
    mult = rng.geometric(1 / 8, size=n // 8 + 4096)
    mult = mult[: np.searchsorted(np.cumsum(mult), n) + 1]
    keys = unique_keys(len(mult), 32, seed=0xC4)
    pk = np.repeat(keys, mult)[:n]
    pv = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    perm = rng.permutation(n)
    pk, pv = pk[perm], pv[perm]
    t = MultiValueHashTable(math.ceil(n / 0.9), 32, 32)
    dk, dv = to_dev(pk), to_dev(pv)
    t.insert(dk, dv)
    nk = t.num_keys()
    assert nk == len(np.unique(pk))
    kout = torch.empty(nk, dtype=torch.int32, device="cuda")
    beg = torch.empty(nk, dtype=torch.int64, device="cuda")
    end = torch.empty(nk, dtype=torch.int64, device="cuda")
    vout = torch.empty(n, dtype=torch.int32, device="cuda")
    got_k, got_v = t.retrieve_all(kout, beg, end, vout)
    assert got_k == nk and got_v == n
    per_key = (end - beg)
    rk = torch.repeat_interleave(kout, per_key)
    # (key, value) multisets: pack each pair into one 64-bit word and sort
    pack = lambda k, v: (k.to(torch.int64) & 0xFFFFFFFF) << 32 | (v.to(torch.int64) & 0xFFFFFFFF)
    a = torch.sort(pack(rk, vout)).values
    b = torch.sort(pack(dk, dv)).values
    assert torch.equal(a, b)


def test_full_bucketlist_multisets(cuda):
    """2^24 keys, each 8 times (2^27 pairs, shuffled), pool 2^28 cells in 256
    arenas, buckets growing by 1.1 from 1 value: every key's buckets are
    1, 2, 3 and 4 values long (13 cells with the three headers), so the
    pool usage is exact; the (key, value) multisets equal the inputs'."""
    from warpcore_amd.tables import BucketListHashTable
    d, mult = 1 << 24, 8
    n = d * mult
    print("[full] bucketlist u32 2^27 values on 2^24 keys", flush=True)
    ukeys = _gen_unique(d, 32, 0x6000)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0xC6)
    dk = ukeys.repeat(mult)[torch.randperm(n, device="cuda", generator=gen)].contiguous()
    dv = _gen_uniform(n, 32, 0xC7)
    t = BucketListHashTable(math.ceil(d / 0.9), 1 << 28, 32, 32, grow_factor=1.1, min_bucket_size=1, cg_size=8)
    st = torch.empty(n, dtype=torch.int32, device="cuda")
    t.insert(dk, dv, status_out=st)
    assert int((st == S_NONE).sum()) == d                # one new key per distinct key
    assert int((st & ~16).count_nonzero()) == 0          # the rest duplicate_key only
    assert t.num_keys() == d and t.num_values() == n
    assert t.value_cells_used() == 13 * d
    kout = torch.empty(d, dtype=torch.int32, device="cuda")
    beg = torch.empty(d, dtype=torch.int64, device="cuda")
    end = torch.empty(d, dtype=torch.int64, device="cuda")
    vout = torch.empty(n, dtype=torch.int32, device="cuda")
    got_k, got_v = t.retrieve_all(kout, beg, end, vout)
    assert got_k == d and got_v == n
    per_key = end - beg
    assert bool((per_key == mult).all()) and int(beg[0]) == 0 and bool((beg[1:] == end[:-1]).all())
    rk = torch.repeat_interleave(kout, per_key)
    pack = lambda k, v: (k.to(torch.int64) & 0xFFFFFFFF) << 32 | (v.to(torch.int64) & 0xFFFFFFFF)
    a = torch.sort(pack(rk, vout)).values
    b = torch.sort(pack(dk, dv)).values
    assert torch.equal(a, b)


def test_full_hashset_u32(cuda):
    from warpcore_amd.tables import HashSet
    n = 1 << 27
    print("[full] hashset u32 2^27", flush=True)
    keys = _gen_unique(n, 32, 0x3000)
    s = HashSet(math.ceil(n / 0.9), 32)
    assert s.capacity() == 151011248
    s.insert(keys)
    assert s.size() == n
    flags = torch.empty(n, dtype=torch.uint8, device="cuda")
    s.contains(keys, flags)
    assert int((flags != 1).sum()) == 0
    absent = _gen_unique(1 << 22, 32, 0x3000 + n + 5)
    f2 = torch.empty(1 << 22, dtype=torch.uint8, device="cuda")
    s.contains(absent, f2)
    assert int(f2.sum()) == 0


def test_full_bloom_2p30_bits(cuda, oracle):
    from warpcore_amd.tables import BloomFilter
    num_bits, k, n = 1 << 30, 8, 1 << 27
    print("[full] bloom 2^30 bits k=8 2^27 keys", flush=True)
    keys = _gen_unique(n, 64, 0x4000)
    bf = BloomFilter(num_bits, k, 64)
    bf.insert(keys)
    hk = from_dev(keys, np.uint64)
    ob = oracle.Bloom(num_bits, k, 64, 64, 1)
    ob.insert(hk)
    np.testing.assert_array_equal(from_dev(bf.words(), np.uint64), ob.words())
    flags = torch.empty(n, dtype=torch.uint8, device="cuda")
    bf.retrieve(keys, flags)
    assert int((flags != 1).sum()) == 0   # no false negatives
    absent = _gen_unique(n, 64, 0x4000 + n + 9)
    bf.retrieve(absent, flags)
    measured = float(flags.float().mean())
    sample = from_dev(absent[: 1 << 20], np.uint64)
    np.testing.assert_array_equal(from_dev(flags[: 1 << 20], np.uint8), ob.query(sample))
    model = bf.fpr(n)
    # the independent-bits model against the measured rate (0.0456 measured,
    # 0.0472 model on MI355X): close, the model does not capture that the
    # word index and the first pattern bit share a hash
    print(f"[full] bloom fpr measured {measured:.5f} model {model:.5f}", flush=True)
    assert 0.9 * model < measured < 2 * model
