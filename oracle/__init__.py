"""TEST INFRASTRUCTURE ONLY -- the oracle.

A sequential CPU restatement of sleeepyjack/warpcore's hot-path semantics
(wc_oracle.hpp / wc_oracle.cpp, each function citing the reference file:line
it follows) plus the OpenMP baseline table (omp_table.cpp). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker or as the reported CPU baseline. The
product never routes through it.

Pinning (see tests/test_oracle_*.py): fmix32/fmix64 against the closed-form
golden values and against an independent pure-Python MurmurHash3 finalizer;
the capacity rule against the reference's own prime table (count, ends,
checksum, samples held in tests/golden/primes_fixture.json); probing schemes
against the cycle-free invariant of tests/src/probing_schemes_test.cu and
the worked DoubleHashing example; the tables against brute-force Python
models on small inputs; the Bloom filter against its closed form and the
FPR model against the reference's naive product where that stays finite.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

# probing schemes
REF_DOUBLE, REF_LINEAR, REF_QUADRATIC, AL_DOUBLE, AL_LINEAR, AL_QUADRATIC = range(6)
# hashers of a probing scheme: pass `scheme | (hasher << 4)` wherever a scheme is taken
H_MURMUR, H_MUELLER, H_NVIDIA, H_IDENTITY = range(4)


def with_hasher(scheme: int, hasher: int) -> int:
    return scheme | (hasher << 4)

S_NONE = 0
S_UNKNOWN = 1 << 0
S_PLE = 1 << 1
S_INVALID_CONFIG = 1 << 2
S_INVALID_KEY = 1 << 3
S_DUP = 1 << 4
S_NOT_FOUND = 1 << 5
S_INDEX_OVERFLOW = 1 << 6
S_OOM = 1 << 7
S_NOT_INIT = 1 << 8
S_DRY_RUN = 1 << 9
S_PHASE = 1 << 10
S_MAXV = 1 << 11

DEFAULT_PL = (1 << 64) - 1 - 1024
DEFAULT_SEED = 0x5AD0DED

_u32, _u64, _i, _vp, _f, _d = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_float, ctypes.c_double)

_SIGS = {
    "wco_fmix32": (_u32, [_u32]), "wco_fmix64": (_u64, [_u64]),
    "wco_nvidia_hash": (_u32, [_u32]), "wco_mueller_hash": (_u32, [_u32]),
    "wco_status_all": (_u32, []), "wco_status_error_mask": (_u32, []), "wco_status_warning_mask": (_u32, []),
    "wco_status_not": (_u32, [_u32]), "wco_status_remove": (_u32, [_u32, _u32]),
    "wco_status_return_boolean": (_i, [_u32, _u32]),
    "wco_fmix32_array": (None, [_vp, _vp, _u64]), "wco_fmix64_array": (None, [_vp, _vp, _u64]),
    "wco_num_primes": (_u64, []), "wco_primes": (None, [_vp, _u64]), "wco_is_prime": (_i, [_u64]),
    "wco_get_valid_capacity": (_u64, [_u64, _u64]),
    "wco_probe_sequence": (_u64, [_i, _i, _u64, _u64, _u64, _u64, _u64, _vp, _u64]),
    "wco_svht_create": (_vp, [_i, _i, _u64, _u64, _i, _u64]), "wco_svht_destroy": (None, [_vp]),
    "wco_svht_create_sized": (_vp, [_i, _i, _u64, _u64, _i, _u64]), "wco_batch_windows": (_u64, [_u64, _u64]),
    "wco_svht_capacity": (_u64, [_vp]), "wco_svht_size": (_u64, [_vp]), "wco_svht_status": (_u32, [_vp]),
    "wco_svht_clear_status": (None, [_vp]),
    "wco_svht_insert": (None, [_vp, _vp, _vp, _u64, _u64, _vp]),
    "wco_svht_retrieve": (None, [_vp, _vp, _u64, _vp, _u64, _vp]),
    "wco_svht_erase": (None, [_vp, _vp, _u64, _u64, _vp]),
    "wco_svht_slots": (None, [_vp, _vp, _vp]),
    "wco_counting_create": (_vp, [_i, _u64, _u64, _i, _u64, _u64]), "wco_counting_destroy": (None, [_vp]),
    "wco_counting_capacity": (_u64, [_vp]), "wco_counting_size": (_u64, [_vp]),
    "wco_counting_status": (_u32, [_vp]),
    "wco_counting_insert": (None, [_vp, _vp, _u64, _u64, _vp]),
    "wco_counting_retrieve": (None, [_vp, _vp, _u64, _vp, _u64, _vp]),
    "wco_hashset_create": (_vp, [_i, _u64, _u64, _i, _u64]), "wco_hashset_destroy": (None, [_vp]),
    "wco_hashset_capacity": (_u64, [_vp]), "wco_hashset_size": (_u64, [_vp]),
    "wco_hashset_status": (_u32, [_vp]),
    "wco_hashset_insert": (None, [_vp, _vp, _u64, _u64, _vp]),
    "wco_hashset_contains": (None, [_vp, _vp, _u64, _vp, _u64, _vp]),
    "wco_hashset_erase": (None, [_vp, _vp, _u64, _u64, _vp]),
    "wco_bloom_create": (_vp, [_i, _i, _u64, _u64, _u64, _u64]), "wco_bloom_destroy": (None, [_vp]),
    "wco_bloom_num_slots": (_u64, [_vp]), "wco_bloom_insert": (None, [_vp, _vp, _u64]),
    "wco_bloom_query": (None, [_vp, _vp, _u64, _vp]), "wco_bloom_words": (None, [_vp, _vp]),
    "wco_bloom_locate": (None, [_vp, _u64, _u64, _vp, _vp]),
    "wco_bloom_fpr_stable": (_d, [_vp, _u64]), "wco_bloom_fpr_naive": (_d, [_vp, _u64]),
    "wco_mvht_create": (_vp, [_i, _i, _u64, _u64, _i, _u64, _u64]), "wco_mvht_destroy": (None, [_vp]),
    "wco_mvht_capacity": (_u64, [_vp]), "wco_mvht_size": (_u64, [_vp]), "wco_mvht_num_keys": (_u64, [_vp]),
    "wco_mvht_status": (_u32, [_vp]),
    "wco_mvht_insert": (None, [_vp, _vp, _vp, _u64, _u64, _vp]),
    "wco_mvht_num_values": (_u64, [_vp, _vp, _u64, _vp, _u64, _vp]),
    "wco_mvht_retrieve": (_u64, [_vp, _vp, _u64, _vp, _vp, _vp, _u64, _vp]),
    "wco_bucketlist_create": (_vp, [_i, _i, _u64, _u64, _u64, _i, _u64, _f, _u64, _u64, _u64]),
    "wco_bucketlist_destroy": (None, [_vp]), "wco_bucketlist_key_capacity": (_u64, [_vp]),
    "wco_bucketlist_num_keys": (_u64, [_vp]), "wco_bucketlist_next_free": (_u64, [_vp]),
    "wco_bucketlist_status": (_u32, [_vp]),
    "wco_bucketlist_insert": (None, [_vp, _vp, _vp, _u64, _u64, _vp]),
    "wco_bucketlist_num_values": (_u64, [_vp, _vp, _u64, _vp, _u64, _vp]),
    "wco_bucketlist_retrieve": (_u64, [_vp, _vp, _u64, _vp, _vp, _vp, _u64, _vp]),
    "wco_omp_max_threads": (_i, []), "wco_omp_create": (_vp, [_i, _i, _u64, _u64, _u64]),
    "wco_omp_destroy": (None, [_vp]), "wco_omp_capacity": (_u64, [_vp]), "wco_omp_init": (None, [_vp]),
    "wco_omp_insert": (_d, [_vp, _vp, _vp, _u64, _vp]), "wco_omp_count": (_d, [_vp, _vp, _u64, _vp]),
    "wco_omp_retrieve": (_d, [_vp, _vp, _u64, _vp, _vp]),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from .build import build
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(_lib, name)
            f.restype, f.argtypes = res, args
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _dt(bits):
    return np.uint32 if bits == 32 else np.uint64


def fmix32(x: int) -> int:
    return lib().wco_fmix32(x)


def fmix64(x: int) -> int:
    return lib().wco_fmix64(x)


def status_all() -> int:
    return lib().wco_status_all()


def status_error_mask() -> int:
    return lib().wco_status_error_mask()


def status_warning_mask() -> int:
    return lib().wco_status_warning_mask()


def status_not(s: int) -> int:
    return lib().wco_status_not(s)


def status_remove(a: int, b: int) -> int:
    return lib().wco_status_remove(a, b)


def return_boolean(statuses, ignore: int = 0) -> np.ndarray:
    """Flags status_handlers::ReturnBoolean<ignore> writes for `statuses`."""
    mask = np.uint32(status_not(ignore))
    return ((np.asarray(statuses, np.uint32) & mask) != 0).astype(np.uint8)


def get_valid_capacity(min_capacity: int, cg: int) -> int:
    return lib().wco_get_valid_capacity(min_capacity, cg)


def batch_capacity(min_capacity: int, window: int) -> int:
    """Slots of a TableBatch table: the smallest prime >= max(2,
    SDIV(min_capacity, window)) windows of `window` slots."""
    return lib().wco_batch_windows(min_capacity, window) * window


def primes() -> np.ndarray:
    n = lib().wco_num_primes()
    out = np.empty(n, np.uint64)
    lib().wco_primes(_p(out), n)
    return out


def probe_sequence(scheme, key_bits, key, seed, cap, cg, pl, max_windows) -> np.ndarray:
    out = np.zeros(max_windows * cg, np.uint64)
    w = lib().wco_probe_sequence(scheme, key_bits, key, seed, cap, cg, pl, _p(out), max_windows)
    return out[: w * cg].reshape(w, cg)


class _Handle:
    _prefix = ""

    def _f(self, name):
        return getattr(lib(), self._prefix + name)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                self._f("destroy")(h)
            except TypeError:   # interpreter shutdown: the module's globals are gone
                pass
            self.h = None


class SVHT(_Handle):
    _prefix = "wco_svht_"

    def __init__(self, min_capacity, key_bits=32, value_bits=32, cg=8, scheme=AL_DOUBLE, seed=DEFAULT_SEED,
                 exact_capacity=None):
        """exact_capacity: exactly that many slots (a multiple of cg), the
        geometry of a TableBatch table, instead of the reference's rule."""
        self.kb, self.vb = key_bits, value_bits
        if exact_capacity is None:
            self.h = self._f("create")(key_bits, value_bits, min_capacity, cg, scheme, seed)
        else:
            self.h = self._f("create_sized")(key_bits, value_bits, exact_capacity, cg, scheme, seed)

    def insert(self, keys, values, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb)); values = np.ascontiguousarray(values, _dt(self.vb))
        st = np.zeros(len(keys), np.uint32)
        self._f("insert")(self.h, _p(keys), _p(values), len(keys), pl, _p(st))
        return st

    def retrieve(self, keys, values_out=None, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        out = np.zeros(len(keys), _dt(self.vb)) if values_out is None else values_out
        st = np.zeros(len(keys), np.uint32)
        self._f("retrieve")(self.h, _p(keys), len(keys), _p(out), pl, _p(st))
        return out, st

    def erase(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        st = np.zeros(len(keys), np.uint32)
        self._f("erase")(self.h, _p(keys), len(keys), pl, _p(st))
        return st

    def capacity(self):
        return self._f("capacity")(self.h)

    def size(self):
        return self._f("size")(self.h)

    def status(self):
        return self._f("status")(self.h)

    def clear_status(self):
        self._f("clear_status")(self.h)

    def slots(self):
        k = np.zeros(self.capacity(), _dt(self.kb)); v = np.zeros(self.capacity(), _dt(self.vb))
        self._f("slots")(self.h, _p(k), _p(v))
        return k, v


class Counting(_Handle):
    _prefix = "wco_counting_"

    def __init__(self, min_capacity, key_bits=64, cg=4, scheme=AL_DOUBLE, seed=DEFAULT_SEED,
                 max_count=(1 << 64) - 1):
        self.kb = key_bits
        self.h = self._f("create")(key_bits, min_capacity, cg, scheme, seed, max_count)

    def insert(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb)); st = np.zeros(len(keys), np.uint32)
        self._f("insert")(self.h, _p(keys), len(keys), pl, _p(st))
        return st

    def retrieve(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        out = np.zeros(len(keys), np.uint64); st = np.zeros(len(keys), np.uint32)
        self._f("retrieve")(self.h, _p(keys), len(keys), _p(out), pl, _p(st))
        return out, st

    def capacity(self):
        return self._f("capacity")(self.h)

    def size(self):
        return self._f("size")(self.h)

    def status(self):
        return self._f("status")(self.h)


class HashSet(_Handle):
    _prefix = "wco_hashset_"

    def __init__(self, min_capacity, key_bits=32, cg=16, scheme=AL_DOUBLE, seed=DEFAULT_SEED):
        self.kb = key_bits
        self.h = self._f("create")(key_bits, min_capacity, cg, scheme, seed)

    def insert(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb)); st = np.zeros(len(keys), np.uint32)
        self._f("insert")(self.h, _p(keys), len(keys), pl, _p(st))
        return st

    def contains(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        flags = np.zeros(len(keys), np.uint8); st = np.zeros(len(keys), np.uint32)
        self._f("contains")(self.h, _p(keys), len(keys), _p(flags), pl, _p(st))
        return flags, st

    def erase(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb)); st = np.zeros(len(keys), np.uint32)
        self._f("erase")(self.h, _p(keys), len(keys), pl, _p(st))
        return st

    def capacity(self):
        return self._f("capacity")(self.h)

    def size(self):
        return self._f("size")(self.h)

    def status(self):
        return self._f("status")(self.h)


class Bloom(_Handle):
    _prefix = "wco_bloom_"

    def __init__(self, num_bits, k, key_bits=64, slot_bits=64, cg=1, seed=DEFAULT_SEED):
        self.kb, self.sb = key_bits, slot_bits
        self.h = self._f("create")(key_bits, slot_bits, num_bits, k, seed, cg)

    def insert(self, keys):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        self._f("insert")(self.h, _p(keys), len(keys))

    def query(self, keys):
        keys = np.ascontiguousarray(keys, _dt(self.kb)); out = np.zeros(len(keys), np.uint8)
        self._f("query")(self.h, _p(keys), len(keys), _p(out))
        return out

    def words(self):
        out = np.zeros(self._f("num_slots")(self.h), _dt(self.sb))
        self._f("words")(self.h, _p(out))
        return out

    def locate(self, key, rank=0):
        s, p = ctypes.c_uint64(), ctypes.c_uint64()
        self._f("locate")(self.h, key, rank, ctypes.byref(s), ctypes.byref(p))
        return s.value, p.value

    def fpr_stable(self, n):
        return self._f("fpr_stable")(self.h, n)

    def fpr_naive(self, n):
        return self._f("fpr_naive")(self.h, n)


class MVHT(_Handle):
    _prefix = "wco_mvht_"

    def __init__(self, min_capacity, key_bits=32, value_bits=32, cg=8, scheme=AL_DOUBLE, seed=DEFAULT_SEED,
                 max_values=(1 << 64) - 1):
        self.kb, self.vb = key_bits, value_bits
        self.h = self._f("create")(key_bits, value_bits, min_capacity, cg, scheme, seed, max_values)

    def insert(self, keys, values, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb)); values = np.ascontiguousarray(values, _dt(self.vb))
        st = np.zeros(len(keys), np.uint32)
        self._f("insert")(self.h, _p(keys), _p(values), len(keys), pl, _p(st))
        return st

    def num_values(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        per = np.zeros(len(keys), np.uint64); st = np.zeros(len(keys), np.uint32)
        total = self._f("num_values")(self.h, _p(keys), len(keys), _p(per), pl, _p(st))
        return total, per, st

    def retrieve(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        total, per, _ = self.num_values(keys, pl)
        begin = np.zeros(len(keys), np.uint64); end = np.zeros(len(keys), np.uint64)
        vals = np.zeros(max(int(total), 1), _dt(self.vb)); st = np.zeros(len(keys), np.uint32)
        n = self._f("retrieve")(self.h, _p(keys), len(keys), _p(begin), _p(end), _p(vals), pl, _p(st))
        return begin, end, vals[:n], st

    def capacity(self):
        return self._f("capacity")(self.h)

    def size(self):
        return self._f("size")(self.h)

    def num_keys(self):
        return self._f("num_keys")(self.h)

    def status(self):
        return self._f("status")(self.h)


class BucketList(_Handle):
    _prefix = "wco_bucketlist_"

    def __init__(self, key_capacity, value_capacity, key_bits=32, value_bits=32, cg=8, scheme=AL_DOUBLE,
                 seed=DEFAULT_SEED, grow=1.1, min_bucket=1, max_bucket=1023, max_values=(1 << 20) - 1):
        self.kb, self.vb = key_bits, value_bits
        self.h = self._f("create")(key_bits, value_bits, key_capacity, value_capacity, cg, scheme, seed, grow,
                                   min_bucket, max_bucket, max_values)

    def insert(self, keys, values, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb)); values = np.ascontiguousarray(values, _dt(self.vb))
        st = np.zeros(len(keys), np.uint32)
        self._f("insert")(self.h, _p(keys), _p(values), len(keys), pl, _p(st))
        return st

    def num_values(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        per = np.zeros(len(keys), np.uint64); st = np.zeros(len(keys), np.uint32)
        total = self._f("num_values")(self.h, _p(keys), len(keys), _p(per), pl, _p(st))
        return total, per, st

    def retrieve(self, keys, pl=DEFAULT_PL):
        keys = np.ascontiguousarray(keys, _dt(self.kb))
        total, _, _ = self.num_values(keys, pl)
        begin = np.zeros(len(keys), np.uint64); end = np.zeros(len(keys), np.uint64)
        vals = np.zeros(max(int(total), 1), _dt(self.vb)); st = np.zeros(len(keys), np.uint32)
        n = self._f("retrieve")(self.h, _p(keys), len(keys), _p(begin), _p(end), _p(vals), pl, _p(st))
        return begin, end, vals[:n], st

    def key_capacity(self):
        return self._f("key_capacity")(self.h)

    def num_keys(self):
        return self._f("num_keys")(self.h)

    def next_free(self):
        return self._f("next_free")(self.h)

    def status(self):
        return self._f("status")(self.h)


class OmpTable(_Handle):
    """The OpenMP baseline table (reported CPU baseline)."""

    _prefix = "wco_omp_"

    def __init__(self, min_capacity, key_bits=64, value_bits=64, cg=8, seed=DEFAULT_SEED):
        self.kb, self.vb = key_bits, value_bits
        self.h = self._f("create")(key_bits, value_bits, min_capacity, cg, seed)

    def init(self):
        self._f("init")(self.h)

    def capacity(self):
        return self._f("capacity")(self.h)

    def insert(self, keys, values, status=True):
        st = np.zeros(len(keys), np.uint32) if status else None
        t = self._f("insert")(self.h, _p(keys), _p(values), len(keys), _p(st))
        return t, st

    def count(self, keys, status=True):
        st = np.zeros(len(keys), np.uint32) if status else None
        t = self._f("count")(self.h, _p(keys), len(keys), _p(st))
        return t, st

    def retrieve(self, keys, values_out, status=True):
        st = np.zeros(len(keys), np.uint32) if status else None
        t = self._f("retrieve")(self.h, _p(keys), len(keys), _p(values_out), _p(st))
        return t, st


def omp_threads() -> int:
    return lib().wco_omp_max_threads()
