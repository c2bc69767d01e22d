"""Status algebra and the alternative hashers/probing schemes (CPU only).

Status: the known-answer checks of tests/src/status_test.cu:8-45 restated,
plus closed forms of the reference's formulas (include/warpcore/status.cuh:
mask 0xFFF at :207, ~s = (~s) ^ mask at :146-150, ReturnBoolean writes
has_any(~Status(Ignore)) at :316-333), evaluated three ways that must
agree: the literals below (worked out by hand from those lines), the
oracle's restatement, and the library's own Status compiled for the host
(wc_status_* through the C ABI, no GPU call).

Hashers/schemes: the oracle's probe sequences for MuellerHash, NvidiaHash
and IdentityMap under double, linear and quadratic probing against an
independent Python restatement of probing_schemes.cuh (aligned windows),
and the cycle-free invariant of tests/src/probing_schemes_test.cu:44-68 for
the permutation hashers.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from warpcore_amd import _capi

M32 = (1 << 32) - 1
NONE, UNKNOWN, PLE, INVALID_CONFIG, INVALID_KEY, DUP, NOT_FOUND = 0, 1, 2, 4, 8, 16, 32
INDEX_OVERFLOW, OOM, NOT_INIT, DRY_RUN, PHASE, MAXV, INVALID_VALUE = 64, 128, 256, 512, 1024, 2048, 4096
ERRORS = UNKNOWN | INVALID_CONFIG | INDEX_OVERFLOW | OOM | NOT_INIT | PHASE


@pytest.fixture(scope="module")
def wc():
    if not _capi.LIB_PATH.exists():
        from warpcore_amd.build import build
        build()
    return _capi.lib()


# ------------------------------------------------------------------ status --
def test_status_closed_forms_oracle():
    assert O.status_all() == 0xFFF
    assert O.status_error_mask() == ERRORS == 0x5C5
    assert O.status_warning_mask() == 0xFFF & ~ERRORS == 0xA3A
    assert O.status_not(NONE) == 0xFFFFF000
    assert O.status_not(INVALID_KEY) == 0xFFFFF008
    assert O.status_not(0xFFF) == 0xFFFFFFFF
    assert O.status_not(INVALID_VALUE) == 0xFFFFE000
    # ReturnBoolean<>: only invalid_value (and bits above the 13 flags) count
    assert list(O.return_boolean([NONE, DUP, NOT_FOUND, INVALID_KEY, PLE, INVALID_VALUE, DUP | INVALID_VALUE])) == \
        [0, 0, 0, 0, 0, 1, 1]
    # ReturnBoolean<duplicate_key>: the ignored flag is the one that counts
    assert list(O.return_boolean([NONE, DUP, NOT_FOUND, DUP | NOT_FOUND], DUP)) == [0, 1, 0, 1]


def test_status_test_cu_known_answers(wc):
    """tests/src/status_test.cu:8-45, section "general", against the library's
    host Status."""
    s = NONE
    assert s == NONE
    s = wc.wc_status_join(s, UNKNOWN)
    assert s == UNKNOWN
    assert wc.wc_status_has_any(s, UNKNOWN)
    assert wc.wc_status_has_any(s, wc.wc_status_all())
    assert not NONE
    s = wc.wc_status_join(s, NOT_FOUND)
    assert wc.wc_status_has_any(s, NOT_FOUND) and wc.wc_status_has_any(s, UNKNOWN)
    assert wc.wc_status_has_any(s, wc.wc_status_all())
    assert not wc.wc_status_has_all(s, wc.wc_status_complement(INVALID_KEY))
    assert wc.wc_status_has_all(s, wc.wc_status_join(NOT_FOUND, UNKNOWN))
    s = wc.wc_status_remove(s, UNKNOWN)
    assert not wc.wc_status_has_any(s, UNKNOWN) and wc.wc_status_has_any(s, NOT_FOUND)
    s = wc.wc_status_remove(s, NOT_FOUND)
    assert s == NONE
    assert wc.wc_status_complement(NONE) != s
    err = wc.wc_status_error_mask()
    assert wc.wc_status_has_any_errors(err) and not wc.wc_status_has_any_warnings(err)
    warn = wc.wc_status_warning_mask()
    assert wc.wc_status_has_any_warnings(warn) and not wc.wc_status_has_any_errors(warn)
    assert not wc.wc_status_has_any(err, warn)


def test_status_library_matches_oracle_exhaustively(wc):
    assert wc.wc_status_all() == O.status_all()
    assert wc.wc_status_error_mask() == O.status_error_mask()
    assert wc.wc_status_warning_mask() == O.status_warning_mask()
    rng = np.random.default_rng(5)
    values = list(range(1 << 13)) + rng.integers(0, 1 << 32, 2000, dtype=np.uint64).tolist()
    for v in values:
        assert wc.wc_status_complement(v) == O.status_not(v)
        assert wc.wc_status_get_errors(v) == v & O.status_error_mask()
        assert wc.wc_status_get_warnings(v) == v & O.status_warning_mask()
    ignores = [NONE, DUP, NOT_FOUND, DUP | NOT_FOUND, INVALID_VALUE, 0xFFF]
    for ign in ignores:
        exp = O.return_boolean(values, ign)
        got = [wc.wc_status_return_boolean(v, ign) for v in values]
        assert got == exp.tolist()
        any_flag = [wc.wc_status_return_any_flag(v, ign) for v in values]
        assert any_flag == [int((v & 0x1FFF & ~ign) != 0) for v in values]
    for a, b in zip(values[:500], values[500:1000]):
        assert wc.wc_status_remove(a, b) == O.status_remove(a, b)
        assert wc.wc_status_join(a, b) == a | b


def test_status_to_string(wc):
    buf = ctypes.create_string_buffer(256)
    n = wc.wc_status_to_string(DUP | UNKNOWN | NOT_FOUND, buf, 256)
    # status.cuh:220-269: duplicate key first, then the flags in bit order
    assert buf.value.decode() == "[duplicate key, unknown error, key not found]" and n == len(buf.value)
    wc.wc_status_to_string(NONE, buf, 256)
    assert buf.value.decode() == "[]"


# ----------------------------------------------------------------- hashers --
def py_mueller(x):
    x = (((x >> 16) ^ x) * 0x45D9F3B) & M32
    x = (((x >> 16) ^ x) * 0x45D9F3B) & M32
    return (x >> 16) ^ x


def py_nvidia(x):
    x = ((x + 0x7ED55D16) + (x << 12)) & M32
    x = ((x ^ 0xC761C23C) ^ (x >> 19)) & M32
    x = ((x + 0x165667B1) + (x << 5)) & M32
    x = ((x + 0xD3A2646C) ^ (x << 9)) & M32
    x = ((x + 0xFD7046C5) + (x << 3)) & M32
    return ((x ^ 0xB55A4F09) ^ (x >> 16)) & M32


def py_windows(scheme, hasher, key, seed, cap, cg, n, bits=32):
    """Aligned windows (w_0 = h1 % p; double: + (h2 % (p-1) + 1); linear: + 1;
    quadratic: + 1, 2, 3, ...), restated from probing_schemes.cuh:79-100,
    173-192 and 264-285 with the window size as the unit."""
    mask = (1 << bits) - 1
    h = {O.H_MUELLER: py_mueller, O.H_NVIDIA: py_nvidia, O.H_IDENTITY: lambda x: x}[hasher]
    p = cap // cg
    w = h((key + seed) & mask) % p
    step = h((key + seed + 1) & mask) % (p - 1) + 1
    out = []
    for i in range(n):
        out.append(w)
        if scheme == O.AL_DOUBLE:
            w = (w + step) % p
        elif scheme == O.AL_LINEAR:
            w = (w + 1) % p
        else:
            w = (w + i + 1) % p
    return out


@pytest.mark.parametrize("scheme", [O.AL_DOUBLE, O.AL_LINEAR, O.AL_QUADRATIC])
@pytest.mark.parametrize("hasher", [O.H_MUELLER, O.H_NVIDIA, O.H_IDENTITY])
def test_probe_sequences_other_hashers(scheme, hasher):
    cg = 8
    cap = O.get_valid_capacity(200000, cg)
    rng = np.random.default_rng(scheme * 7 + hasher)
    for key in rng.integers(1, M32, 50, dtype=np.uint64).tolist():
        seq = O.probe_sequence(O.with_hasher(scheme, hasher), 32, key, O.DEFAULT_SEED, cap, cg, 12 * cg, 12)
        exp = py_windows(scheme, hasher, key, O.DEFAULT_SEED, cap, cg, 12)
        assert (seq[:, 0] // cg).tolist() == exp
    # 64-bit keys with the identity map: h1 = key + seed in 64 bits
    cap = O.get_valid_capacity(300000, cg)
    for key in (1, 2**40 + 3, 2**63 + 11):
        seq = O.probe_sequence(O.with_hasher(scheme, O.H_IDENTITY), 64, key, 5, cap, cg, 6 * cg, 6)
        assert (seq[:, 0] // cg).tolist() == py_windows(scheme, O.H_IDENTITY, key, 5, cap, cg, 6, bits=64)


@pytest.mark.parametrize("hasher", [O.H_MUELLER, O.H_IDENTITY])
@pytest.mark.parametrize("scheme", [O.AL_DOUBLE, O.AL_LINEAR])
def test_cycle_free_permutation_hashers(scheme, hasher):
    """probing_schemes_test.cu:44-68 with the true-permutation hashers."""
    for cg in (1, 8):
        cap = O.get_valid_capacity(4242, cg)
        for key in (1, 42, 1337):
            seq = O.probe_sequence(O.with_hasher(scheme, hasher), 32, key, 42, cap, cg, cap, cap // cg + 1)
            assert seq.size == cap and len(np.unique(seq)) == cap


def _np_hash(hasher, x):
    x = x.astype(np.uint64)
    if hasher == O.H_IDENTITY:
        return x
    if hasher == O.H_MURMUR:
        out = np.empty(len(x), np.uint32)
        O.lib().wco_fmix32_array(x.astype(np.uint32).ctypes.data_as(ctypes.c_void_p),
                                 out.ctypes.data_as(ctypes.c_void_p), len(x))
        return out.astype(np.uint64)
    x = (((x >> 16) ^ x) * 0x45D9F3B) & M32
    x = (((x >> 16) ^ x) * 0x45D9F3B) & M32
    return (x >> 16) ^ x


def test_oracle_tables_take_the_hasher():
    """With probing limited to one window (pl = cg) a key is stored iff fewer
    than cg earlier keys share its first window, so the statuses depend on
    the hasher; closed form computed here by ranking keys per window."""
    cg = 8
    cap = O.get_valid_capacity(1, cg)
    p = cap // cg
    n = cap + cap // 8
    keys = (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(2654435761)) & np.uint64(M32)
    keys = keys.astype(np.uint32)
    res = {}
    for hasher in (O.H_MURMUR, O.H_MUELLER, O.H_IDENTITY):
        o = O.SVHT(1, 32, 32, cg=cg, scheme=O.with_hasher(O.AL_DOUBLE, hasher))
        assert o.capacity() == cap
        st = o.insert(keys, keys, pl=cg)
        w = _np_hash(hasher, (keys.astype(np.uint64) + O.DEFAULT_SEED) & M32) % np.uint64(p)
        order = np.argsort(w, kind="stable")
        ws = w[order]
        first = np.searchsorted(ws, ws, "left")
        rank = np.empty(n, np.int64)
        rank[order] = np.arange(n) - first
        exp = np.where(rank < cg, O.S_NONE, O.S_PLE).astype(np.uint32)
        np.testing.assert_array_equal(st, exp)
        assert (exp == O.S_PLE).sum() > 1000
        res[hasher] = st
    assert not np.array_equal(res[O.H_MURMUR], res[O.H_IDENTITY])
    assert not np.array_equal(res[O.H_MURMUR], res[O.H_MUELLER])
