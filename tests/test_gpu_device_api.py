"""Device-side tile API (tables passed by value into user kernels), driven
through tests/device_api/libwc_device_api.so: the filtered histogram of
examples/src/advanced_usage_from_device.cu, and insert / retrieve / erase /
multi-value round trips issued from user kernels, checked against the
oracle."""
import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

from datagen import dev_empty, from_dev, to_dev, unique_keys

pytestmark = pytest.mark.gpu

LIB = Path(__file__).resolve().parent / "device_api" / "libwc_device_api.so"
S_NONE, S_DUP, S_NOT_FOUND = 0, 16, 32


@pytest.fixture(scope="module")
def dlib(cuda):
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is missing: run the build first")
    lib = ctypes.CDLL(str(LIB))
    P, U64, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    lib.wcx_filter_and_count_u32.argtypes = [P, U64, U64, P, P, P]
    lib.wcx_svht_create.restype = P
    lib.wcx_svht_create.argtypes = [U64]
    lib.wcx_svht_destroy.argtypes = [P]
    lib.wcx_svht_phase.argtypes = [P, I, P, P, U64, P, P]
    lib.wcx_mvht_roundtrip_u32.argtypes = [P, P, U64, P, U64, U64, P, P, U64, P, P]
    lib.wcx_bucketlist_roundtrip_u32.argtypes = [P, P, U64, P, U64, U64, U64, P, P, U64, P, P]
    return lib


_keep = []


def _p(t):
    """Device pointer of t; t is kept alive until the test module ends (a
    temporary tensor freed before the call could be handed to the next
    allocation)."""
    if t is None:
        return None
    _keep.append(t)
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(autouse=True)
def _release():
    yield
    torch.cuda.synchronize()
    _keep.clear()


def test_filtered_histogram_from_device(dlib):
    rng = np.random.default_rng(1)
    uniq = unique_keys(1 << 18, 32, seed=2)
    repeated = unique_keys(1 << 16, 32, seed=3, exclude=uniq)
    keys = np.concatenate([uniq, np.repeat(repeated, 4)])
    rng.shuffle(keys)
    n = len(keys)
    ko, co = dev_empty(n, 32), dev_empty(n, 64)
    num = ctypes.c_uint64(0)
    assert dlib.wcx_filter_and_count_u32(_p(to_dev(keys)), n, 1 << 26, _p(ko), _p(co), ctypes.byref(num)) == 0
    got = dict(zip(from_dev(ko, np.uint32)[: num.value].tolist(), from_dev(co, np.uint64)[: num.value].tolist()))
    # every repeated key is counted 3 times (4 occurrences minus the first),
    # or 4 when the filter gave a false positive on its first occurrence
    rep = set(repeated.tolist())
    exact = sum(1 for k in rep if got.get(k) == 3)
    assert all(got.get(k) in (3, 4) for k in rep)
    extra = [k for k in got if k not in rep]
    # unique keys appear only through false positives (count 1); FPR here is ~1e-6
    assert all(got[k] == 1 for k in extra)
    assert len(extra) + (len(repeated) - exact) <= 8


def test_svht_device_insert_retrieve_erase(dlib, oracle):
    n = 40000
    keys = unique_keys(n, 64, seed=11)
    vals = unique_keys(n, 64, seed=12)
    h = dlib.wcx_svht_create(int(n / 0.8))
    try:
        st = dev_empty(n, 32, 0x77)
        assert dlib.wcx_svht_phase(h, 0, _p(to_dev(keys)), _p(to_dev(vals)), n, None, _p(st)) == 0
        assert np.all(from_dev(st, np.uint32) == S_NONE)
        # re-insert half with new values: duplicates, values overwritten
        vals2 = vals.copy()
        vals2[: n // 2] ^= np.uint64(0xFFFF)
        assert dlib.wcx_svht_phase(h, 0, _p(to_dev(keys[: n // 2])), _p(to_dev(vals2[: n // 2])), n // 2, None,
                                   _p(st)) == 0
        assert np.all(from_dev(st, np.uint32)[: n // 2] == S_DUP)
        absent = unique_keys(1000, 64, seed=13, exclude=keys)
        q = np.concatenate([keys, absent])
        out, qst = dev_empty(len(q), 64, 0x55), dev_empty(len(q), 32)
        assert dlib.wcx_svht_phase(h, 1, _p(to_dev(q)), None, len(q), _p(out), _p(qst)) == 0
        o = oracle.SVHT(int(n / 0.8), 64, 64, cg=8)
        o.insert(keys, vals)
        o.insert(keys[: n // 2], vals2[: n // 2])
        ov, ost = o.retrieve(q)
        np.testing.assert_array_equal(from_dev(qst, np.uint32), ost)
        got = from_dev(out, np.uint64)
        np.testing.assert_array_equal(got[:n], ov[:n])
        # erase a third, then retrieve again
        er = keys[::3]
        est = dev_empty(len(er), 32)
        assert dlib.wcx_svht_phase(h, 2, _p(to_dev(er)), None, len(er), None, _p(est)) == 0
        assert np.all(from_dev(est, np.uint32) == S_NONE)
        o.erase(er)
        out, qst = dev_empty(len(q), 64, 0x55), dev_empty(len(q), 32)
        assert dlib.wcx_svht_phase(h, 1, _p(to_dev(q)), None, len(q), _p(out), _p(qst)) == 0
        ov, ost = o.retrieve(q)
        np.testing.assert_array_equal(from_dev(qst, np.uint32), ost)
        hit = ost == S_NONE
        np.testing.assert_array_equal(from_dev(out, np.uint64)[hit], ov[hit])
    finally:
        dlib.wcx_svht_destroy(h)


def _multi_pairs(num_keys, mean, seed):
    rng = np.random.default_rng(seed)
    keys = unique_keys(num_keys, 32, seed)
    mult = rng.geometric(1.0 / mean, size=num_keys)
    k = np.repeat(keys, mult)
    v = unique_keys(len(k), 32, seed + 1)
    perm = rng.permutation(len(k))
    return keys, mult, k[perm], v[perm]


@pytest.mark.parametrize("kind", ["mvht", "bucketlist"])
def test_multi_value_device_roundtrip(dlib, kind):
    keys, mult, k, v = _multi_pairs(3000, 5.0, seed=21)
    absent = unique_keys(200, 32, seed=23, exclude=keys)
    q = np.concatenate([keys, absent])
    stride = int(mult.max())
    ist = dev_empty(len(k), 32, 0x77)
    vals = dev_empty(len(q) * stride, 32, 0xAB)
    counts, qst = dev_empty(len(q), 64), dev_empty(len(q), 32)
    if kind == "mvht":
        rc = dlib.wcx_mvht_roundtrip_u32(_p(to_dev(k)), _p(to_dev(v)), len(k), _p(to_dev(q)), len(q),
                                         2 * len(k), _p(ist), _p(vals), stride, _p(counts), _p(qst))
    else:
        rc = dlib.wcx_bucketlist_roundtrip_u32(_p(to_dev(k)), _p(to_dev(v)), len(k), _p(to_dev(q)), len(q),
                                               2 * len(keys), 4 * len(k), _p(ist), _p(vals), stride, _p(counts),
                                               _p(qst))
    assert rc == 0
    s = from_dev(ist, np.uint32)
    assert set(np.unique(s)) <= {S_NONE, S_DUP}
    assert np.sum(s == S_NONE) == len(keys)
    c = from_dev(counts, np.uint64)
    np.testing.assert_array_equal(c[: len(keys)], mult.astype(np.uint64))
    assert np.all(c[len(keys):] == 0)
    qs = from_dev(qst, np.uint32)
    assert np.all(qs[: len(keys)] == S_NONE) and np.all(qs[len(keys):] == S_NOT_FOUND)
    got = from_dev(vals, np.uint32).reshape(len(q), stride)
    exp = {}
    for key, val in zip(k.tolist(), v.tolist()):
        exp.setdefault(key, []).append(val)
    for i, key in enumerate(keys.tolist()):
        assert sorted(got[i, : mult[i]].tolist()) == sorted(exp[key])


@pytest.mark.parametrize("max_count,preset", [((1 << 32) - 1, (1 << 32) - 10), (1000, 990), (1000, 1000)])
def test_counting_combined_add_saturates(dlib, max_count, preset):
    """A combined count (one add of all occurrences of a key in a tile) ends
    where single increments would (counting_hash_table.cuh:154-170: each adds
    one while the count is below max_count): min(preset + occurrences,
    max_count), with no index_overflow even where one add would pass 2^32-1."""
    dlib.wcx_counting_u32_saturation.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                 ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p]
    rng = np.random.default_rng(31)
    others = unique_keys(8192, 32, seed=32)
    hot = others[0]
    batch = np.concatenate([np.full(40, hot, np.uint32), others[1:]])
    rng.shuffle(batch)
    batch = np.concatenate([[hot], batch]).astype(np.uint32)   # keys[0] is the preset key
    n = len(batch)
    st, counts = dev_empty(n, 32, 0x77), dev_empty(n, 32, 0x55)
    word = ctypes.c_uint32(0)
    assert dlib.wcx_counting_u32_saturation(_p(to_dev(batch)), n, max_count, preset, _p(st), _p(counts),
                                            ctypes.byref(word)) == 0
    got_s, got_c = from_dev(st, np.uint32), from_dev(counts, np.uint32)
    is_hot = batch == hot
    assert np.all(got_s[is_hot] == S_DUP)            # the key was in the table before the batch
    assert np.all(got_s[~is_hot] == S_NONE)
    assert np.all(got_c[is_hot] == min(preset + int(is_hot.sum()), max_count))
    assert np.all(got_c[~is_hot] == 1)
    assert word.value == S_DUP                        # duplicate_key joined, no index_overflow (64)
