"""Pins the oracle before it is trusted (CPU only).

Each check is chosen so that a plausible slip in the restatement (a wrong
shift or constant, a dropped term, an off-by-one window count, a wrong
status join) fails it:
  * hashers: closed-form golden values, an independent pure-Python MurmurHash3
    finalizer, and invertibility (fmix is a bijection);
  * capacity rule: the reference's own prime table (count, ends, SHA-256,
    samples) and the configuration capacities;
  * probing: the worked DoubleHashing example and the cycle-free invariant of
    tests/src/probing_schemes_test.cu:44-68 for every cg 1..64;
  * tables: brute-force dict/set/Counter models on small random op streams;
  * Bloom: closed-form words, no false negatives, FPR model vs the
    reference's own product form where the latter is finite.
"""
import collections
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
CF = json.loads((GOLD / "closed_form.json").read_text())
PR = json.loads((GOLD / "primes_fixture.json").read_text())
M32, M64 = (1 << 32) - 1, (1 << 64) - 1


def py_fmix32(x):
    x ^= x >> 16; x = (x * 0x85EBCA6B) & M32; x ^= x >> 13; x = (x * 0xC2B2AE35) & M32; x ^= x >> 16
    return x


def py_fmix64(x):
    x ^= x >> 33; x = (x * 0xFF51AFD7ED558CCD) & M64; x ^= x >> 33; x = (x * 0xC4CEB9FE1A85EC53) & M64
    x ^= x >> 33
    return x


def inv_xorshift(y, s, bits):
    x = y
    for _ in range(bits // s + 1):
        x = y ^ (x >> s)
    return x


def inv_fmix64(y):
    y = inv_xorshift(y, 33, 64); y = (y * pow(0xC4CEB9FE1A85EC53, -1, 1 << 64)) & M64
    y = inv_xorshift(y, 33, 64); y = (y * pow(0xFF51AFD7ED558CCD, -1, 1 << 64)) & M64
    return inv_xorshift(y, 33, 64)


def inv_fmix32(y):
    y = inv_xorshift(y, 16, 32); y = (y * pow(0xC2B2AE35, -1, 1 << 32)) & M32
    y = inv_xorshift(y, 13, 32); y = (y * pow(0x85EBCA6B, -1, 1 << 32)) & M32
    return inv_xorshift(y, 16, 32)


def test_fmix_golden():
    for k, v in CF["fmix32"].items():
        assert O.fmix32(int(k)) == v
    for k, v in CF["fmix64"].items():
        assert O.fmix64(int(k)) == v


def test_fmix_independent_and_bijective():
    rng = np.random.default_rng(1)
    for x in rng.integers(0, 1 << 32, 2000, dtype=np.uint64).tolist():
        assert O.fmix32(x) == py_fmix32(x)
        assert inv_fmix32(O.fmix32(x)) == x
    for x in rng.integers(0, M64, 2000, dtype=np.uint64).tolist():
        assert O.fmix64(x) == py_fmix64(x)
        assert inv_fmix64(O.fmix64(x)) == x


def test_other_hashers_closed_form():
    lib = O.lib()
    rng = np.random.default_rng(2)
    for x in rng.integers(0, 1 << 32, 500, dtype=np.uint64).tolist():
        y = x
        y = ((y + 0x7ED55D16) + (y << 12)) & M32
        y = ((y ^ 0xC761C23C) ^ (y >> 19)) & M32
        y = ((y + 0x165667B1) + (y << 5)) & M32
        y = ((y + 0xD3A2646C) ^ (y << 9)) & M32
        y = ((y + 0xFD7046C5) + (y << 3)) & M32
        y = ((y ^ 0xB55A4F09) ^ (y >> 16)) & M32
        assert lib.wco_nvidia_hash(x) == y
        z = (((x >> 16) ^ x) * 0x45D9F3B) & M32
        z = (((z >> 16) ^ z) * 0x45D9F3B) & M32
        z = (z >> 16) ^ z
        assert lib.wco_mueller_hash(x) == z


def test_prime_table_matches_reference():
    p = O.primes()
    assert len(p) == PR["count"] == 131036
    assert int(p[0]) == PR["first"] and int(p[-1]) == PR["last"]
    for i, v in PR["samples"].items():
        assert int(p[int(i)]) == v
    assert hashlib.sha256(p.astype("<u8").tobytes()).hexdigest() == PR["sha256_le_u64"]


def test_prime_rule_brute_force_head():
    # first entries by trial division: first prime >= 131072, then next >= prev + 131072
    def isp(n):
        if n < 2:
            return False
        d = 2
        while d * d <= n:
            if n % d == 0:
                return False
            d += 1
        return True
    p = O.primes()
    x = 131072
    for i in range(6):
        while not isp(x):
            x += 1
        assert int(p[i]) == x
        x += 131072


def test_capacity_rule():
    for mn, cg, cap in CF["capacity"]:
        assert O.get_valid_capacity(mn, cg) == cap
    assert O.get_valid_capacity(17177758133 * 8, 8) == 17177758133 * 8
    assert O.get_valid_capacity(17177758133 * 8 + 1, 8) == 0   # beyond the table


def _is_prime(n):
    if n < 2:
        return False
    d = 2
    while d * d <= n:
        if n % d == 0:
            return False
        d += 1
    return True


def test_batch_capacity_rule():
    """TableBatch's small-table rule by trial division: the smallest prime
    >= max(2, ceil(min / window)) windows of `window` slots."""
    rng = np.random.default_rng(9)
    cases = [(1, 4), (4, 4), (8, 4), (9, 4), (64, 16), (65, 16), (100, 32), (65536, 16)]
    cases += [(int(m), int(w)) for m, w in zip(np.exp2(rng.uniform(0, 17, 200)).astype(np.int64),
                                               rng.choice([4, 8, 16], 200))]
    for mn, w in cases:
        x = max(2, -(-mn // w))
        while not _is_prime(x):
            x += 1
        assert O.batch_capacity(mn, w) == x * w, (mn, w)


@pytest.mark.parametrize("kb,window,mn", [(32, 4, 64), (32, 8, 100), (64, 16, 300), (32, 16, 64)])
def test_sized_svht_fills_every_slot(kb, window, mn):
    """A table of exactly batch_capacity slots: double hashing over a prime
    number of windows reaches every window, so exactly `capacity` distinct
    keys are stored and every further key reports probing_length_exceeded;
    stored keys read back their values."""
    cap = O.batch_capacity(mn, window)
    o = O.SVHT(0, kb, kb, cg=window, exact_capacity=cap)
    assert o.capacity() == cap
    keys = np.arange(1, cap + 31, dtype=np.uint32 if kb == 32 else np.uint64) * 7919
    st = o.insert(keys, keys + 1)
    assert np.sum(st == O.S_NONE) == cap and np.sum(st == O.S_PLE) == 30
    assert np.all(st[:cap] == O.S_NONE)   # sequential: the first `capacity` keys fit
    assert o.size() == cap and o.status() == O.S_PLE
    out, rst = o.retrieve(keys)
    assert np.all(rst[:cap] == O.S_NONE) and np.all(out[:cap] == keys[:cap] + 1)
    assert np.all(rst[cap:] == O.S_PLE)   # full table: no EMPTY slot ends the walk
    for key in (1, 42, 1337):
        seq = O.probe_sequence(O.AL_DOUBLE, kb, key, O.DEFAULT_SEED, cap, window, cap, cap // window + 1)
        assert len(np.unique(seq)) == cap


def test_double_hashing_worked_example():
    c = CF["double_hashing_ref"]
    for case in c["cases"]:
        assert O.fmix32((case["key"] + c["seed"]) & M32) == case["h1"]
        seq = O.probe_sequence(O.REF_DOUBLE, 32, case["key"], c["seed"], c["cap"], c["cg"], 2 * c["cg"], 2)
        assert int(seq[0, 0]) == case["lane0"]
        assert (int(seq[1, 0]) - int(seq[0, 0])) % c["cap"] == case["step"]
        assert int(seq[1, 0]) == case["next_lane0"]
        # lanes are consecutive slots in the reference's unaligned window
        assert all(int(seq[0, r]) == (case["lane0"] + r) % c["cap"] for r in range(c["cg"]))


@pytest.mark.parametrize("scheme", [O.REF_DOUBLE, O.REF_LINEAR, O.AL_DOUBLE, O.AL_LINEAR])
@pytest.mark.parametrize("cg", [1, 2, 4, 8, 16, 32, 64])
def test_cycle_free(scheme, cg):
    """tests/src/probing_schemes_test.cu:44-68: one sequence of probing_length
    = capacity slots visits every slot exactly once."""
    for mn in (12345, 42424, 6969):
        cap = O.get_valid_capacity(mn, cg)
        for key in (1, 42, 69, 1337):
            for seed in (5, 42):
                seq = O.probe_sequence(scheme, 32, key, seed, cap, cg, cap, cap // cg + 1)
                assert seq.size == cap
                assert len(np.unique(seq)) == cap


@pytest.mark.parametrize("cg", [1, 8, 16])
def test_aligned_windows_are_aligned(cg):
    cap = O.get_valid_capacity(1 << 20, cg)
    seq = O.probe_sequence(O.AL_DOUBLE, 64, 12345, O.DEFAULT_SEED, cap, cg, 40 * cg, 40)
    assert seq.shape == (40, cg)
    assert np.all(seq[:, 0] % cg == 0)
    assert np.all(np.diff(seq, axis=1) == 1)


def test_window_budget():
    cap = O.get_valid_capacity(1, 8)
    for pl, windows in [(0, 1), (1, 1), (8, 1), (9, 2), (16, 2), (17, 3), (O.DEFAULT_PL, cap // 8)]:
        seq = O.probe_sequence(O.AL_DOUBLE, 32, 7, 5, cap, 8, pl, cap)
        assert seq.shape[0] == windows


# ------------------------------------------------------ brute-force models --
def _random_ops(rng, n, universe, kb):
    keys = rng.integers(0, universe, n).astype(np.uint64)
    keys[rng.random(n) < 0.02] = 0
    keys[rng.random(n) < 0.02] = (1 << kb) - 1
    return keys.astype(np.uint32 if kb == 32 else np.uint64)


@pytest.mark.parametrize("kb,vb,cg", [(32, 32, 8), (64, 64, 4), (32, 64, 16)])
def test_svht_matches_dict_model(kb, vb, cg):
    rng = np.random.default_rng(kb + vb + cg)
    o = O.SVHT(3000, kb, vb, cg=cg)
    model, status = {}, 0
    tomb = (1 << kb) - 1
    for _ in range(30):
        op = rng.integers(0, 3)
        keys = _random_ops(rng, 200, 400, kb)
        if op == 0:
            vals = rng.integers(0, 1 << 31, len(keys)).astype(np.uint32 if vb == 32 else np.uint64)
            st = o.insert(keys, vals)
            for k, v, s in zip(keys.tolist(), vals.tolist(), st.tolist()):
                if k in (0, tomb):
                    exp = O.S_INVALID_KEY
                elif k in model:
                    exp = O.S_DUP
                else:
                    exp = O.S_NONE
                assert s == exp
                status |= exp
                if exp in (O.S_NONE, O.S_DUP):
                    model[k] = v
        elif op == 1:
            out, st = o.retrieve(keys)
            for k, v, s in zip(keys.tolist(), out.tolist(), st.tolist()):
                exp = O.S_INVALID_KEY if k in (0, tomb) else (O.S_NONE if k in model else O.S_NOT_FOUND)
                assert s == exp
                status |= exp
                if exp == O.S_NONE:
                    assert v == model[k]
        else:
            st = o.erase(keys)
            for k, s in zip(keys.tolist(), st.tolist()):
                if k in (0, tomb):
                    exp = O.S_INVALID_KEY
                    status |= exp
                elif k in model:
                    exp = O.S_NONE
                    del model[k]
                else:
                    exp = O.S_NOT_FOUND          # not joined
                assert s == exp
        assert o.size() == len(model)
        assert o.status() == status


def test_counting_matches_counter():
    rng = np.random.default_rng(5)
    o = O.Counting(5000, 64, cg=4, max_count=7)
    c = collections.Counter()
    for _ in range(10):
        keys = rng.integers(1, 600, 500).astype(np.uint64)
        st = o.insert(keys)
        for k, s in zip(keys.tolist(), st.tolist()):
            assert s == (O.S_DUP if c[k] else O.S_NONE)
            c[k] += 1
    q = np.arange(1, 700, dtype=np.uint64)
    out, st = o.retrieve(q)
    for k, v, s in zip(q.tolist(), out.tolist(), st.tolist()):
        if c[k]:
            assert s == O.S_NONE and v == min(c[k], 7)
        else:
            assert s == O.S_NOT_FOUND
    assert o.size() == len(c)


def test_hashset_matches_set():
    rng = np.random.default_rng(6)
    o = O.HashSet(4000, 32, cg=16)
    model = set()
    for _ in range(20):
        keys = _random_ops(rng, 300, 900, 32)
        kind = rng.integers(0, 3)
        if kind == 0:
            st = o.insert(keys)
            for k, s in zip(keys.tolist(), st.tolist()):
                if k in (0, M32):
                    assert s == O.S_INVALID_KEY
                else:
                    assert s == (O.S_DUP if k in model else O.S_NONE)
                    model.add(k)
        elif kind == 1:
            f, st = o.contains(keys)
            for k, fl, s in zip(keys.tolist(), f.tolist(), st.tolist()):
                if k in (0, M32):
                    assert s == O.S_INVALID_KEY
                else:
                    assert s == O.S_NONE and fl == (k in model)
        else:
            st = o.erase(keys)
            for k, s in zip(keys.tolist(), st.tolist()):
                assert s == (O.S_INVALID_KEY if k in (0, M32) else O.S_NONE)
                model.discard(k)
        assert o.size() == len(model)


def test_mvht_matches_multiset_model():
    rng = np.random.default_rng(8)
    o = O.MVHT(6000, 32, 32, cg=8, max_values=5)
    model = collections.defaultdict(list)
    for _ in range(8):
        keys = rng.integers(1, 300, 400).astype(np.uint32)
        vals = rng.integers(0, 1 << 30, 400).astype(np.uint32)
        st = o.insert(keys, vals)
        for k, v, s in zip(keys.tolist(), vals.tolist(), st.tolist()):
            n = len(model[k])
            if n >= 5:
                assert s == O.S_DUP | O.S_MAXV
            else:
                assert s == (O.S_DUP if n else O.S_NONE)
                model[k].append(v)
    q = np.arange(1, 320, dtype=np.uint32)
    total, per, st = o.num_values(q)
    assert total == sum(len(model[k]) for k in q.tolist() if k in model)
    begin, end, vals, st2 = o.retrieve(q)
    for i, k in enumerate(q.tolist()):
        exp = sorted(model.get(k, []))
        assert per[i] == len(exp) == end[i] - begin[i]
        assert sorted(vals[begin[i]:end[i]].tolist()) == exp
        assert st2[i] == (O.S_NONE if exp else O.S_NOT_FOUND)
    assert o.num_keys() == len([k for k in model if model[k]])
    assert o.size() == sum(len(v) for v in model.values())


def bucket_pool_use(count, min_bucket, max_bucket, grow):
    """Pool slots used by one key holding `count` values: a first bucket of
    min_bucket slots, then buckets of min(max, ceil(prev * grow)) values plus
    one header slot each (storage.cuh:1113-1284, float32 growth)."""
    if count == 0:
        return 0
    used, size, held = min_bucket, min_bucket, min_bucket
    while held < count:
        size = int(min(np.float32(max_bucket), np.ceil(np.float32(size) * np.float32(grow))))
        used += size + 1
        held += size
    return used


@pytest.mark.parametrize("grow,minb,maxb", [(1.1, 1, 1023), (2.0, 4, 64), (1.5, 3, 10)])
def test_bucketlist_matches_model(grow, minb, maxb):
    rng = np.random.default_rng(9)
    o = O.BucketList(4000, 1 << 20, 32, 32, cg=8, grow=grow, min_bucket=minb, max_bucket=maxb, max_values=200)
    model = collections.defaultdict(list)
    for _ in range(6):
        keys = rng.integers(1, 200, 1500).astype(np.uint32)
        vals = rng.integers(0, 1 << 30, 1500).astype(np.uint32)
        st = o.insert(keys, vals)
        for k, v, s in zip(keys.tolist(), vals.tolist(), st.tolist()):
            n = len(model[k])
            base = O.S_DUP if n else O.S_NONE
            if n >= 200:
                assert s == base | O.S_MAXV
            else:
                assert s == base
                model[k].append(v)
    q = np.arange(1, 220, dtype=np.uint32)
    begin, end, vals, st = o.retrieve(q)
    for i, k in enumerate(q.tolist()):
        exp = sorted(model.get(k, []))
        assert sorted(vals[begin[i]:end[i]].tolist()) == exp
        assert st[i] == (O.S_NONE if exp else O.S_NOT_FOUND)
    assert o.next_free() == sum(bucket_pool_use(len(v), minb, maxb, grow) for v in model.values())


def test_bucketlist_out_of_memory():
    o = O.BucketList(1000, 20, 32, 32, cg=8, min_bucket=2)
    keys = np.repeat(np.arange(1, 9, dtype=np.uint32), 3)
    st = o.insert(keys, keys)
    assert np.any(st & O.S_OOM)
    assert o.status() & O.S_OOM


def test_bloom_closed_form_and_no_false_negatives():
    b = CF["bloom"]
    bf = O.Bloom(b["num_bits"], b["k"], 64, 64, b["cg"], b["seed"])
    for case in b["cases"]:
        assert bf.locate(case["key"]) == (case["slot"], case["word"])
    # independent restatement of slot/pattern for random keys
    rng = np.random.default_rng(10)
    for key in rng.integers(1, M64, 200, dtype=np.uint64).tolist():
        seeded = (key + b["seed"]) & M64
        slot = py_fmix64(seeded) % (b["num_bits"] // 64)
        word = 0
        for j in range(b["k"]):
            word |= 1 << (py_fmix64((seeded + j) & M64) % 64)
        assert bf.locate(key) == (slot, word)
    small = O.Bloom(1 << 16, 6, 64, 64, 1, b["seed"])
    keys = rng.integers(1, M64, 3000, dtype=np.uint64)
    small.insert(keys)
    assert np.all(small.query(keys) == 1)
    w = small.words()
    exp = np.zeros_like(w)
    for key in keys.tolist():
        s, p = small.locate(key)
        exp[s] |= np.uint64(p)
    np.testing.assert_array_equal(w, exp)


@pytest.mark.parametrize("cg,slot_bits", [(1, 64), (4, 32), (2, 64)])
def test_bloom_cooperative_blocks(cg, slot_bits):
    rng = np.random.default_rng(cg * slot_bits)
    bf = O.Bloom(1 << 18, 7, 32, slot_bits, cg, 77)
    keys = rng.integers(1, M32, 5000, dtype=np.uint64).astype(np.uint32)
    bf.insert(keys)
    assert np.all(bf.query(keys) == 1)
    absent = rng.integers(1, M32, 5000, dtype=np.uint64).astype(np.uint32)
    fp = bf.query(absent[~np.isin(absent, keys)]).mean()
    assert fp < 0.2


def test_bloom_fpr_model():
    b = CF["bloom"]
    bf = O.Bloom(b["num_bits"], b["k"], 64, 64, 1, b["seed"])
    assert abs(bf.fpr_stable(1 << 27) - b["fpr_blocked_n_2_27"]) < 5e-5
    assert bf.fpr_naive(1 << 27) == float("inf")   # the reference's product overflows (F3)
    for bits, k, n in [(1 << 16, 4, 2000), (1 << 18, 6, 9000), (1 << 20, 8, 40000)]:
        small = O.Bloom(bits, k, 64, 64, 1, 1)
        a, c = small.fpr_stable(n), small.fpr_naive(n)
        assert np.isfinite(c) and abs(a - c) <= 1e-9 * max(c, 1e-12)
