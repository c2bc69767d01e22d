"""BloomFilter on the GPU vs the oracle: the bit array word for word, query
flags bit-exact, insert_and_query, ragged geometries and the FPR model."""
import numpy as np
import pytest
import torch

from datagen import dev_empty, dtype, from_dev, to_dev, unique_keys

pytestmark = pytest.mark.gpu


def _words(bf):
    return from_dev(bf.words(), dtype(bf.slot_bits()))


@pytest.mark.parametrize("kb,sb,cg,k,num_bits", [
    (64, 64, 1, 6, 1 << 20),
    (32, 64, 1, 4, 1 << 18),
    (64, 32, 1, 8, 1 << 20),
    (32, 32, 2, 5, 1 << 19),
    (64, 64, 4, 7, 1 << 21),
    (64, 32, 8, 3, 3 << 17),
    (32, 64, 16, 9, 1 << 20),
    (64, 64, 32, 12, 1 << 21),
    (64, 64, 64, 16, 1 << 22),
    (64, 64, 4, 5, 64 * 10 + 3),     # 11 slots, 3 blocks of 4: the last block wraps to slot 0
    (32, 32, 8, 6, 32 * 1001 + 7),   # ragged bit count
])
def test_bloom_words_and_flags_match_oracle(cuda, oracle, kb, sb, cg, k, num_bits):
    from warpcore_amd.tables import BloomFilter
    print(f"[bloom] kb={kb} sb={sb} cg={cg} k={k} bits={num_bits}", flush=True)
    n = min(1 << 16, max(num_bits // 12, 8))
    keys = unique_keys(n, kb, seed=31 + cg + k)
    bf = BloomFilter(num_bits, k, kb, seed=0x5EED + cg, slot_bits=sb, cg_size=cg)
    ob = oracle.Bloom(num_bits, k, kb, sb, cg, 0x5EED + cg)
    assert bf.num_slots() == len(ob.words())
    assert bf.num_blocks() == -(-bf.num_slots() // cg)
    assert bf.bytes_total() == bf.num_slots() * sb // 8
    assert np.all(_words(bf) == 0)
    bf.insert(to_dev(keys))
    ob.insert(keys)
    np.testing.assert_array_equal(_words(bf), ob.words())
    absent = unique_keys(4 * n, kb, seed=97 + k, exclude=keys)
    q = np.concatenate([keys, absent])
    flags = dev_empty(len(q), 8, 0x55)
    bf.retrieve(to_dev(q), flags)
    got = from_dev(flags, np.uint8)
    np.testing.assert_array_equal(got, ob.query(q))
    assert np.all(got[:n] == 1)   # no false negatives


@pytest.mark.parametrize("kb", [32, 64])
def test_bloom_insert_order_and_duplicates_do_not_matter(cuda, oracle, kb):
    from warpcore_amd.tables import BloomFilter
    keys = unique_keys(20000, kb, seed=4)
    dup = np.concatenate([keys, keys[::3], keys[::7]])
    np.random.default_rng(9).shuffle(dup)
    a = BloomFilter(1 << 18, 6, kb, cg_size=4)
    b = BloomFilter(1 << 18, 6, kb, cg_size=4)
    a.insert(to_dev(keys))
    b.insert(to_dev(dup[: len(dup) // 2]))
    b.insert(to_dev(dup[len(dup) // 2:]))
    np.testing.assert_array_equal(_words(a), _words(b))


def test_bloom_empty_input_and_reinit(cuda, oracle):
    from warpcore_amd.tables import BloomFilter
    bf = BloomFilter(1 << 16, 5, 64)
    empty = torch.empty(0, dtype=torch.int64, device="cuda")
    bf.insert(empty)
    bf.retrieve(empty, torch.empty(0, dtype=torch.uint8, device="cuda"))
    assert np.all(_words(bf) == 0)
    keys = unique_keys(3000, 64, seed=12)
    bf.insert(to_dev(keys))
    assert np.any(_words(bf) != 0)
    bf.init(seed=1234)
    assert bf.seed() == 1234
    assert np.all(_words(bf) == 0)
    bf.insert(to_dev(keys))
    ob = oracle.Bloom(1 << 16, 5, 64, 64, 1, 1234)
    ob.insert(keys)
    np.testing.assert_array_equal(_words(bf), ob.words())


def test_bloom_bad_configurations_rejected(cuda):
    from warpcore_amd.tables import BloomFilter
    for kwargs in ({"num_bits": 0, "k": 3}, {"num_bits": 1024, "k": 0},
                   {"num_bits": 1024, "k": 3, "cg_size": 3}, {"num_bits": 1024, "k": 3, "slot_bits": 16}):
        with pytest.raises(ValueError):
            BloomFilter(**kwargs)


@pytest.mark.parametrize("kb", [32, 64])
def test_bloom_insert_and_query(cuda, oracle, kb):
    from warpcore_amd.tables import BloomFilter
    a_keys = unique_keys(30000, kb, seed=21)
    b_keys = unique_keys(30000, kb, seed=22, exclude=a_keys)
    bf = BloomFilter(1 << 20, 7, kb, slot_bits=64, cg_size=1)
    bf.insert(to_dev(a_keys))
    flags = dev_empty(len(a_keys), 8, 0x55)
    bf.insert_and_query(to_dev(a_keys), flags)
    assert np.all(from_dev(flags, np.uint8) == 1)     # every bit already set
    flags = dev_empty(len(b_keys), 8, 0x55)
    bf.insert_and_query(to_dev(b_keys), flags)
    got = from_dev(flags, np.uint8)
    assert set(np.unique(got)) <= {0, 1}
    ob = oracle.Bloom(1 << 20, 7, kb, 64, 1)
    ob.insert(a_keys)
    before = ob.query(b_keys)
    # a key whose bits were all set by A reports true whatever the order;
    # a key reported false had some bit unset by A
    assert np.all(got[before == 1] == 1)
    ob.insert(b_keys)
    np.testing.assert_array_equal(_words(bf), ob.words())
    # sequential single-key calls are deterministic
    bf2 = BloomFilter(1 << 12, 4, kb, cg_size=1)
    ob2 = oracle.Bloom(1 << 12, 4, kb, 64, 1)
    for key in b_keys[:40]:
        f = dev_empty(1, 8, 0x55)
        bf2.insert_and_query(to_dev(np.array([key], dtype(kb))), f)
        expect = ob2.query(np.array([key], dtype(kb)))[0]
        ob2.insert(np.array([key], dtype(kb)))
        assert from_dev(f, np.uint8)[0] == expect
    big = BloomFilter(1 << 12, 4, kb, cg_size=4)
    with pytest.raises(RuntimeError):
        big.insert_and_query(to_dev(a_keys[:4]), dev_empty(4, 8))


def test_bloom_fpr_model_and_measured_rate(cuda, oracle):
    from warpcore_amd.tables import BloomFilter
    num_bits, k, n = 1 << 22, 6, 1 << 18
    bf = BloomFilter(num_bits, k, 64, cg_size=1)
    ob = oracle.Bloom(num_bits, k, 64, 64, 1)
    assert bf.fpr(n) == pytest.approx(ob.fpr_stable(n), rel=1e-12)
    assert bf.fpr(n) == pytest.approx(ob.fpr_naive(n), rel=1e-9)
    # at 2^27 keys the reference's factorial product overflows; the model stays finite
    assert np.isfinite(bf.fpr(1 << 27))
    keys = unique_keys(n, 64, seed=3)
    bf.insert(to_dev(keys))
    ob.insert(keys)
    probe = unique_keys(1 << 20, 64, seed=4, exclude=keys)
    flags = dev_empty(len(probe), 8)
    bf.retrieve(to_dev(probe), flags)
    got = from_dev(flags, np.uint8)
    np.testing.assert_array_equal(got, ob.query(probe))
    # the block index and the first pattern bit come from the same hash, so
    # the measured rate sits above the independent-bits model here
    assert bf.fpr(n) < got.mean() < 2 * bf.fpr(n)
