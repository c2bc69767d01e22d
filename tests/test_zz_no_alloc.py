"""wc_counting_u64_insert_noalloc (the insert that allocates nothing: its
combining buffers are a fixed device array of the library) against
wc_counting_u64_insert on the same inputs: counts, the multiset of
(key, status) pairs of every call (which occurrence of a new key reports
`none` is unspecified), sizes and table status words, over two calls in a
row, the second smaller; inputs over the limit are refused and change
nothing."""
import ctypes

import numpy as np
import pytest
import torch

from datagen import dev_empty, from_dev, to_dev, unique_keys

pytestmark = pytest.mark.gpu

S_NONE, S_DUP = 0, 16


def _zipf_batch(n, distinct, seed):
    rng = np.random.default_rng(seed)
    uni = unique_keys(distinct, 64, seed=seed)
    w = 1.0 / np.arange(1, distinct + 1) ** 1.1
    w /= w.sum()
    return uni[rng.choice(distinct, size=n, p=w)]


def _pairs(keys, st):
    order = np.lexsort((st, keys))
    return np.stack([keys[order], st[order].astype(np.uint64)])


def test_noalloc_insert_matches_insert(cuda):
    from warpcore_amd import _capi
    lib = _capi.lib()
    limit = lib.wc_counting_u64_insert_noalloc_max()
    assert limit == 1 << 20
    a = lib.wc_counting_u64_create(1 << 21, 0x5ad0ded, (1 << 64) - 1, 0, 4)
    b = lib.wc_counting_u64_create(1 << 21, 0x5ad0ded, (1 << 64) - 1, 0, 4)
    try:
        first = _zipf_batch(limit, 1 << 18, seed=7)
        second = np.concatenate([_zipf_batch(300_000, 1 << 17, seed=8), first[:5000]])   # smaller, overlapping
        for call, batch in enumerate([first, second]):
            print(f"[no_alloc] call {call}: {len(batch)} occurrences", flush=True)
            k = to_dev(batch)
            sa, sb = dev_empty(len(batch), 32, 0x77), dev_empty(len(batch), 32, 0x66)
            torch.cuda.synchronize()
            assert lib.wc_counting_u64_insert_noalloc(a, k.data_ptr(), len(batch), None, (1 << 64) - 1 - 1024,
                                                      sa.data_ptr()) == 0
            assert lib.wc_counting_u64_insert(b, k.data_ptr(), len(batch), None, (1 << 64) - 1 - 1024,
                                              sb.data_ptr()) == 0
            torch.cuda.synchronize()
            got_a, got_b = from_dev(sa, np.uint32), from_dev(sb, np.uint32)
            assert set(np.unique(got_a)) <= {S_NONE, S_DUP}
            np.testing.assert_array_equal(_pairs(batch, got_a), _pairs(batch, got_b))
            assert lib.wc_counting_u64_size(a, None) == lib.wc_counting_u64_size(b, None)
            assert lib.wc_counting_u64_peek_status(a, None) == lib.wc_counting_u64_peek_status(b, None)
        q = to_dev(np.unique(np.concatenate([first, second])))
        ca, cb = dev_empty(q.numel(), 64, 0x11), dev_empty(q.numel(), 64, 0x22)
        torch.cuda.synchronize()
        assert lib.wc_counting_u64_retrieve(a, q.data_ptr(), q.numel(), ca.data_ptr(), None,
                                            (1 << 64) - 1 - 1024, None) == 0
        assert lib.wc_counting_u64_retrieve(b, q.data_ptr(), q.numel(), cb.data_ptr(), None,
                                            (1 << 64) - 1 - 1024, None) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_dev(ca, np.uint64), from_dev(cb, np.uint64))
        # exact counts against the inputs
        uq, cnt = np.unique(np.concatenate([first, second]), return_counts=True)
        np.testing.assert_array_equal(from_dev(ca, np.uint64), cnt.astype(np.uint64))
        # over the limit: refused, nothing changes
        big = to_dev(_zipf_batch(limit + 1, 1 << 10, seed=9))
        size_before = lib.wc_counting_u64_size(a, None)
        torch.cuda.synchronize()
        assert lib.wc_counting_u64_insert_noalloc(a, big.data_ptr(), limit + 1, None, (1 << 64) - 1 - 1024,
                                                  None) == -4
        torch.cuda.synchronize()
        assert lib.wc_counting_u64_size(a, None) == size_before
    finally:
        lib.wc_counting_u64_destroy(a)
        lib.wc_counting_u64_destroy(b)
