"""Template choices other than the defaults, on the GPU, against the oracle.

- Status handlers: ReturnBoolean<Ignore> (the reference's formula,
  include/warpcore/status.cuh:146-150,316-333) and ReturnAnyFlag<Ignore> as
  compiled for the device, over every 13-bit status word; insert_bool /
  retrieve_bool through the C ABI against the oracle's ReturnBoolean model
  applied to the oracle's per-key statuses.
- Probing schemes and hashers (probing_schemes.cuh:22-302, hashers.cuh:14-175):
  the device iterators' window sequences against the oracle's, and tables
  built with them (wc_svht_*_create_custom) against the oracle: statuses and
  values, and, with probing limited to one window, the number of keys that
  each first window admits (which depends on the hasher and the scheme).
- SoAStore tables (storage.cuh:218-432) against the oracle, and the store
  invariants of tests/src/key_value_store_test.cu:26-160 for both stores.
- for_each: "for_each count == size" (tests/src/single_value_hash_table_test.cu:138-154)
  for SingleValueHashTable and HashSet.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

from datagen import dev_empty, dtype, from_dev, to_dev, uniform_values, unique_keys

pytestmark = pytest.mark.gpu

LIB = Path(__file__).resolve().parent / "device_api" / "libwc_device_api.so"
S_NONE, S_PLE, S_INVALID_KEY, S_DUP, S_NOT_FOUND = 0, 2, 8, 16, 32
SENTINEL = 0xA5A5A5A5

SCHEMES = {"double": 0, "linear": 1, "quadratic": 2}
HASHERS = {"murmur": 0, "mueller": 1, "nvidia": 2, "identity": 3}


@pytest.fixture(scope="module")
def dlib(cuda):
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is missing: run the build first")
    lib = ctypes.CDLL(str(LIB))
    P, U64, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    lib.wcx_status_handlers.argtypes = [P, U64, I, P, P]
    lib.wcx_probe_windows.argtypes = [I, I, I, P, U64, U64, U64, U64, U64, P, P]
    lib.wcx_store_check.argtypes = [I, U64, P]
    return lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _oracle_scheme(oracle, probing, hasher):
    base = {"double": oracle.AL_DOUBLE, "linear": oracle.AL_LINEAR, "quadratic": oracle.AL_QUADRATIC}[probing]
    return oracle.with_hasher(base, HASHERS[hasher])


# ------------------------------------------------------------------ status --
@pytest.mark.parametrize("which,ignore", [(0, 0), (1, S_DUP), (2, S_DUP | S_NOT_FOUND), (3, 1 << 12)])
def test_device_status_handlers(dlib, oracle, which, ignore):
    rng = np.random.default_rng(which)
    st = np.concatenate([np.arange(1 << 13, dtype=np.uint32),
                         rng.integers(0, 1 << 32, 4096, dtype=np.uint64).astype(np.uint32)])
    d = to_dev(st)
    rb, anyf = dev_empty(len(st), 8, 0x77), dev_empty(len(st), 8, 0x77)
    assert dlib.wcx_status_handlers(_p(d), len(st), which, _p(rb), _p(anyf)) == 0
    np.testing.assert_array_equal(from_dev(rb, np.uint8), oracle.return_boolean(st, ignore))
    exp_any = ((st & np.uint32(0x1FFF & ~ignore)) != 0).astype(np.uint8)
    np.testing.assert_array_equal(from_dev(anyf, np.uint8), exp_any)


@pytest.mark.parametrize("kb,vb", [(32, 32), (64, 64)])
def test_bool_handlers_through_capi(cuda, oracle, kb, vb):
    """insert_bool / retrieve_bool write ReturnBoolean<>'s flag: with the
    reference's operator~ none of the flags these calls produce (duplicate,
    invalid key, not found) sets it."""
    from warpcore_amd.tables import SingleValueHashTable
    n = 1 << 14
    keys = unique_keys(n, kb, seed=71)
    batch = np.concatenate([keys, keys[:100], np.zeros(5, dtype(kb))])
    vals = uniform_values(len(batch), vb, seed=72)
    t = SingleValueHashTable(2 * n, kb, vb)
    o = oracle.SVHT(2 * n, kb, vb, cg=8)
    f = dev_empty(len(batch), 8, 0x55)
    t.insert_bool(to_dev(batch), to_dev(vals), f)
    ost = o.insert(batch, vals)
    np.testing.assert_array_equal(from_dev(f, np.uint8), oracle.return_boolean(ost))
    q = np.concatenate([keys, unique_keys(1000, kb, seed=73, exclude=keys), np.zeros(3, dtype(kb))])
    out = dev_empty(len(q), vb, SENTINEL)
    f = dev_empty(len(q), 8, 0x55)
    t.retrieve_bool(to_dev(q), out, f)
    _, qst = o.retrieve(q)
    np.testing.assert_array_equal(from_dev(f, np.uint8), oracle.return_boolean(qst))
    assert set(np.unique(qst).tolist()) == {S_NONE, S_NOT_FOUND, S_INVALID_KEY}


# ---------------------------------------------------------- probe windows --
COMBOS = [(s, h, kb) for s in SCHEMES for h in HASHERS for kb in (32, 64)
          if not (kb == 64 and h in ("mueller", "nvidia"))]


@pytest.mark.parametrize("probing,hasher,kb", COMBOS)
def test_device_probe_windows(dlib, oracle, probing, hasher, kb):
    cg, n, mw = 8, 512, 24
    cap = oracle.get_valid_capacity(3_000_000, cg)
    keys = unique_keys(n, kb, seed=81)
    wins, cnt = dev_empty(n * mw, 64, 0xEE), dev_empty(n, 32)
    pl = 20 * cg - 3   # 20 windows
    rc = dlib.wcx_probe_windows(SCHEMES[probing], HASHERS[hasher], kb, _p(to_dev(keys)), n, oracle.DEFAULT_SEED,
                                cap, pl, mw, _p(wins), _p(cnt))
    assert rc == 0
    got = from_dev(wins, np.uint64).reshape(n, mw)
    counts = from_dev(cnt, np.uint32)
    assert np.all(counts == 20)
    sch = _oracle_scheme(oracle, probing, hasher)
    for i, k in enumerate(keys.tolist()):
        seq = oracle.probe_sequence(sch, kb, k, oracle.DEFAULT_SEED, cap, cg, pl, mw)
        assert seq.shape[0] == 20
        np.testing.assert_array_equal(got[i, :20], seq[:, 0] // cg)


# ---------------------------------------------------------- custom tables --
TABLES = [("linear", "murmur", "aos"), ("quadratic", "murmur", "aos"), ("double", "mueller", "aos"),
          ("double", "nvidia", "aos"), ("double", "identity", "aos"), ("linear", "identity", "aos"),
          ("quadratic", "identity", "aos"), ("double", "murmur", "soa")]


def _custom(oracle, min_cap, kb, probing, hasher, storage):
    from warpcore_amd.tables import SingleValueHashTable
    t = SingleValueHashTable(min_cap, kb, kb, probing=probing, hasher=hasher, storage=storage)
    o = oracle.SVHT(min_cap, kb, kb, cg=8, scheme=_oracle_scheme(oracle, probing, hasher))
    return t, o


@pytest.mark.parametrize("probing,hasher,storage", TABLES)
@pytest.mark.parametrize("kb", [32, 64])
def test_custom_table_parity(cuda, shape, oracle, probing, hasher, storage, kb):
    if kb == 64 and hasher in ("mueller", "nvidia"):
        pytest.skip("32-bit hashers")
    print(f"[custom] {probing} {hasher} {storage} kb={kb} shape={shape}", flush=True)
    n = 1 << 17
    keys = unique_keys(n, kb, seed=91)
    vals = uniform_values(n, kb, seed=92)
    absent = unique_keys(n // 2, kb, seed=93, exclude=keys)
    min_cap = int(n / 0.85)
    t, o = _custom(oracle, min_cap, kb, probing, hasher, storage)
    assert t.capacity() == o.capacity()
    st = dev_empty(n, 32, 0xFFFFFFFF)
    t.insert(to_dev(keys), to_dev(vals), status_out=st)
    np.testing.assert_array_equal(from_dev(st, np.uint32), o.insert(keys, vals))
    assert t.size() == o.size()
    q = np.concatenate([keys, absent])
    np.random.default_rng(94).shuffle(q)
    out, qst = dev_empty(len(q), kb, SENTINEL), dev_empty(len(q), 32)
    t.retrieve(to_dev(q), out, status_out=qst)
    exp = np.full(len(q), SENTINEL, dtype(kb))
    _, est = o.retrieve(q, exp)
    np.testing.assert_array_equal(from_dev(qst, np.uint32), est)
    np.testing.assert_array_equal(from_dev(out, dtype(kb)), exp)
    gone = keys[::5]
    est = dev_empty(len(gone), 32)
    t.erase(to_dev(gone), status_out=est)
    np.testing.assert_array_equal(from_dev(est, np.uint32), o.erase(gone))
    out, qst = dev_empty(len(q), kb, SENTINEL), dev_empty(len(q), 32)
    t.retrieve(to_dev(q), out, status_out=qst)
    exp = np.full(len(q), SENTINEL, dtype(kb))
    _, est = o.retrieve(q, exp)
    np.testing.assert_array_equal(from_dev(qst, np.uint32), est)
    np.testing.assert_array_equal(from_dev(out, dtype(kb)), exp)
    assert t.size() == o.size()
    assert int(t.peek_status()) == o.status()
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    t.for_each_count(counter)
    assert int(counter.item()) == t.size()


def _first_windows(oracle, hasher, keys, p):
    """h1(key + seed) % p for 32-bit keys, vectorised (probing_schemes.cuh:79,
    :176, :267: every scheme starts at the first hash)."""
    x = (keys.astype(np.uint64) + oracle.DEFAULT_SEED) & np.uint64(0xFFFFFFFF)
    if hasher == "murmur":
        out = np.empty(len(x), np.uint32)
        oracle.lib().wco_fmix32_array(x.astype(np.uint32).ctypes.data_as(ctypes.c_void_p),
                                      out.ctypes.data_as(ctypes.c_void_p), len(x))
        h = out.astype(np.uint64)
    elif hasher == "identity":
        h = x
    else:
        h = np.array([(oracle.lib().wco_mueller_hash if hasher == "mueller" else oracle.lib().wco_nvidia_hash)(int(v))
                      for v in x.tolist()], np.uint64)
    return h % np.uint64(p)


@pytest.mark.parametrize("probing,hasher,storage", TABLES)
def test_custom_table_one_window(cuda, oracle, probing, hasher, storage):
    """Probing limited to one window: each first window admits exactly
    min(keys starting there, cg) keys, whatever their order, so the counts
    pin the device hasher inside the table's own kernels."""
    cg = 8
    t, o = _custom(oracle, 1, 32, probing, hasher, storage)
    cap = t.capacity()
    n = cap + cap // 8
    keys = unique_keys(n, 32, seed=95)
    st = dev_empty(n, 32)
    t.insert(to_dev(keys), to_dev(keys), status_out=st, probing_length=cg)
    got = from_dev(st, np.uint32)
    assert set(np.unique(got).tolist()) <= {S_NONE, S_PLE}
    w = _first_windows(oracle, hasher, keys, cap // cg)
    admitted = np.bincount(w.astype(np.int64), weights=(got == S_NONE), minlength=cap // cg)
    starting = np.bincount(w.astype(np.int64), minlength=cap // cg)
    np.testing.assert_array_equal(admitted.astype(np.int64), np.minimum(starting, cg))
    ost = o.insert(keys, keys, pl=cg)
    assert (got == S_NONE).sum() == (ost == S_NONE).sum() == t.size() == o.size()


@pytest.mark.parametrize("probing,hasher,kb", COMBOS)
def test_device_window_cursor(dlib, oracle, probing, hasher, kb):
    """The window cursor the bulk kernels step with (detail::WindowCursor)
    against the oracle, as for the iterator classes above."""
    cg, n, mw = 8, 256, 24
    cap = oracle.get_valid_capacity(2_000_000, cg)
    keys = unique_keys(n, kb, seed=83)
    wins, cnt = dev_empty(n * mw, 64, 0xEE), dev_empty(n, 32)
    rc = dlib.wcx_probe_windows(SCHEMES[probing] | 16, HASHERS[hasher], kb, _p(to_dev(keys)), n,
                                oracle.DEFAULT_SEED, cap, mw * cg, mw, _p(wins), _p(cnt))
    assert rc == 0
    got = from_dev(wins, np.uint64).reshape(n, mw)
    assert np.all(from_dev(cnt, np.uint32) == mw)
    sch = _oracle_scheme(oracle, probing, hasher)
    for i, k in enumerate(keys.tolist()):
        seq = oracle.probe_sequence(sch, kb, k, oracle.DEFAULT_SEED, cap, cg, mw * cg, mw)
        np.testing.assert_array_equal(got[i], seq[:, 0] // cg)


# ------------------------------------------------------------------ stores --
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("capacity", [123456, 42424242, 69696969])
def test_store_invariants(dlib, kind, capacity):
    errors = torch.zeros(5, dtype=torch.int64, device="cuda")
    assert dlib.wcx_store_check(kind, capacity, _p(errors)) == 0
    assert errors.cpu().tolist() == [0, 0, 0, 0, 0]


# ---------------------------------------------------------------- for_each --
@pytest.mark.parametrize("kb,vb,cg", [(32, 32, 8), (64, 64, 8), (32, 32, 16), (32, 64, 8)])
def test_svht_for_each_count_equals_size(cuda, kb, vb, cg):
    from warpcore_amd.tables import SingleValueHashTable
    n = 100000
    keys = unique_keys(n, kb, seed=101)
    t = SingleValueHashTable(2 * n, kb, vb, cg_size=cg)
    t.insert(to_dev(keys), to_dev(uniform_values(n, vb, seed=102)))
    t.erase(to_dev(keys[: n // 4]))
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    t.for_each_count(counter)
    assert int(counter.item()) == t.size() == n - n // 4


@pytest.mark.parametrize("kb", [32, 64])
def test_hashset_for_each_count_equals_size(cuda, kb):
    from warpcore_amd.tables import HashSet
    n = 100000
    keys = unique_keys(n, kb, seed=103)
    s = HashSet(2 * n, kb)
    s.insert(to_dev(keys))
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    s.for_each_count(counter)
    assert int(counter.item()) == s.size() == n
