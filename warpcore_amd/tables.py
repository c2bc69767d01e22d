"""Python mirror of the reference's host API over the C ABI.

Each class mirrors one class template of sleeepyjack/warpcore for a fixed
(key, value) instantiation: the same constructor arguments, the same member
functions with the same argument order (device buffers, stream, probing
length, optional per-key status output), and the same error model (nothing
raises for per-key outcomes; flags go to ``status_out`` and to the table's
status word). Buffers are CUDA tensors (torch on ROCm) of the matching
element width; signedness is ignored, bits are what count.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi

DEFAULT_SEED = 0x5AD0DED
DEFAULT_PROBING_LENGTH = (1 << 64) - 1 - 1024


class Status(int):
    """Status word (include/warpcore/status.cuh:44-68)."""

    NONE = 0
    UNKNOWN_ERROR = 1 << 0
    PROBING_LENGTH_EXCEEDED = 1 << 1
    INVALID_CONFIGURATION = 1 << 2
    INVALID_KEY = 1 << 3
    DUPLICATE_KEY = 1 << 4
    KEY_NOT_FOUND = 1 << 5
    INDEX_OVERFLOW = 1 << 6
    OUT_OF_MEMORY = 1 << 7
    NOT_INITIALIZED = 1 << 8
    DRY_RUN = 1 << 9
    INVALID_PHASE_OVERLAP = 1 << 10
    MAX_VALUES_FOR_KEY_REACHED = 1 << 11
    INVALID_VALUE = 1 << 12
    ALL = 0xFFF   # status.cuh:207: invalid_value is outside all()

    def has_any(self, mask: int = ALL) -> bool:
        return bool(self & mask)

    def __invert__(self) -> "Status":
        """(~s) ^ 0xFFF as a 32-bit word (status.cuh:146-150)."""
        return Status(((~int(self)) ^ self.ALL) & 0xFFFFFFFF)


# Probing scheme, hasher and slot storage of SingleValueHashTable.custom
# (include/warpcore/probing_schemes.cuh, hashers.cuh, storage.cuh).
PROBING = {"double": 0, "linear": 1, "quadratic": 2}
HASHER = {"murmur": 0, "mueller": 1, "nvidia": 2, "identity": 3}
STORAGE = {"aos": 0, "soa": 1}


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device buffers must be CUDA tensors")
    if not t.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    if isinstance(stream, torch.cuda.Stream):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


def _width(t, bits: int, what: str):
    if t is not None and t.element_size() * 8 != bits:
        raise ValueError(f"{what} must have {bits}-bit elements, got {t.dtype}")


class SingleValueHashTable:
    """SingleValueHashTable<K, V> (include/warpcore/single_value_hash_table.cuh)."""

    _NAMES = {(32, 32): "u32u32", (64, 64): "u64u64", (32, 64): "u32u64", (64, 32): "u64u32"}

    def __init__(self, min_capacity: int, key_bits: int = 32, value_bits: int = 32,
                 seed: int = DEFAULT_SEED, no_init: bool = False, cg_size: int = 0,
                 probing: str = "double", hasher: str = "murmur", storage: str = "aos"):
        """probing / hasher / storage other than the defaults select
        SingleValueHashTable<K, V, E, T, Probing<Hasher, Hasher, 8>, Storage>
        (wc_svht_*_create_custom; u32u32 and u64u64 only)."""
        self.key_bits, self.value_bits = key_bits, value_bits
        self._p = f"wc_svht_{self._NAMES[(key_bits, value_bits)]}_"
        self._lib = _capi.lib()
        if (probing, hasher, storage) == ("double", "murmur", "aos"):
            self._h = self._fn("create")(min_capacity, seed, int(no_init), cg_size)
        else:
            self._h = self._fn("create_custom")(min_capacity, seed, int(no_init), cg_size, PROBING[probing],
                                                HASHER[hasher], HASHER[hasher], STORAGE[storage])
        if not self._h:
            raise ValueError(f"unsupported configuration (cg_size={cg_size}, {probing}, {hasher}, {storage})")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def init(self, seed: int | None = None, stream=None):
        if seed is None:
            _check(self._fn("init")(self._h, _stream(stream)), "init")
        else:
            _check(self._fn("init_seed")(self._h, seed, _stream(stream)), "init")

    def insert(self, keys_in, values_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH,
               status_out=None, num_in: int | None = None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_in, self.value_bits, "values_in")
        _width(status_out, 32, "status_out")
        n = keys_in.numel() if num_in is None else num_in
        _check(self._fn("insert")(self._h, _ptr(keys_in), _ptr(values_in), n, _stream(stream),
                                  probing_length, _ptr(status_out)), "insert")

    def insert_bool(self, keys_in, values_in, flags_out, stream=None, probing_length=DEFAULT_PROBING_LENGTH):
        _check(self._fn("insert_bool")(self._h, _ptr(keys_in), _ptr(values_in), keys_in.numel(),
                                       _stream(stream), probing_length, _ptr(flags_out)), "insert")

    def retrieve(self, keys_in, values_out, stream=None, probing_length=DEFAULT_PROBING_LENGTH,
                 status_out=None, num_in: int | None = None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_out, self.value_bits, "values_out")
        _width(status_out, 32, "status_out")
        n = keys_in.numel() if num_in is None else num_in
        _check(self._fn("retrieve")(self._h, _ptr(keys_in), n, _ptr(values_out), _stream(stream),
                                    probing_length, _ptr(status_out)), "retrieve")

    def retrieve_bool(self, keys_in, values_out, flags_out, stream=None, probing_length=DEFAULT_PROBING_LENGTH):
        _check(self._fn("retrieve_bool")(self._h, _ptr(keys_in), keys_in.numel(), _ptr(values_out),
                                         _stream(stream), probing_length, _ptr(flags_out)), "retrieve")

    def erase(self, keys_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _check(self._fn("erase")(self._h, _ptr(keys_in), keys_in.numel(), _stream(stream),
                                 probing_length, _ptr(status_out)), "erase")

    def retrieve_all(self, keys_out, values_out, stream=None) -> int:
        n = ctypes.c_uint64(0)
        _check(self._fn("retrieve_all")(self._h, _ptr(keys_out), _ptr(values_out), ctypes.byref(n),
                                        _stream(stream)), "retrieve_all")
        return n.value

    def for_each_count(self, counter_out, stream=None):
        """for_each(f) with f adding one to counter_out[0] (a caller-zeroed
        64-bit device word) per stored pair."""
        _width(counter_out, 64, "counter_out")
        _check(self._fn("for_each_count")(self._h, _ptr(counter_out), _stream(stream)), "for_each")

    def size(self, stream=None) -> int:
        return self._fn("size")(self._h, _stream(stream))

    def load_factor(self, stream=None) -> float:
        return self._fn("load_factor")(self._h, _stream(stream))

    def capacity(self) -> int:
        return self._fn("capacity")(self._h)

    def bytes_total(self) -> int:
        return self._fn("bytes_total")(self._h)

    def peek_status(self, stream=None) -> Status:
        return Status(self._fn("peek_status")(self._h, _stream(stream)))

    def pop_status(self, stream=None) -> Status:
        return Status(self._fn("pop_status")(self._h, _stream(stream)))

    def is_initialized(self) -> bool:
        return bool(self._fn("is_initialized")(self._h))

    def cg_size(self) -> int:
        return self._fn("cg_size")(self._h)


class CountingHashTable:
    """CountingHashTable<K, uint64_t> (include/warpcore/counting_hash_table.cuh)."""

    def __init__(self, min_capacity: int, key_bits: int = 64, seed: int = DEFAULT_SEED,
                 max_count: int = (1 << 64) - 1, no_init: bool = False, cg_size: int = 0):
        self.key_bits = key_bits
        self._p = f"wc_counting_u{key_bits}_"
        self._lib = _capi.lib()
        self._h = self._fn("create")(min_capacity, seed, max_count, int(no_init), cg_size)
        if not self._h:
            raise ValueError(f"unsupported configuration (cg_size={cg_size})")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def init(self, seed: int | None = None, stream=None):
        if seed is None:
            _check(self._fn("init")(self._h, _stream(stream)), "init")
        else:
            _check(self._fn("init_seed")(self._h, seed, _stream(stream)), "init")

    def insert(self, keys_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _check(self._fn("insert")(self._h, _ptr(keys_in), keys_in.numel(), _stream(stream), probing_length,
                                  _ptr(status_out)), "insert")

    def retrieve(self, keys_in, counts_out, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(counts_out, 64, "counts_out")
        _check(self._fn("retrieve")(self._h, _ptr(keys_in), keys_in.numel(), _ptr(counts_out), _stream(stream),
                                    probing_length, _ptr(status_out)), "retrieve")

    def erase(self, keys_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _check(self._fn("erase")(self._h, _ptr(keys_in), keys_in.numel(), _stream(stream), probing_length,
                                 _ptr(status_out)), "erase")

    def retrieve_all(self, keys_out, counts_out, stream=None) -> int:
        n = ctypes.c_uint64(0)
        _check(self._fn("retrieve_all")(self._h, _ptr(keys_out), _ptr(counts_out), ctypes.byref(n),
                                        _stream(stream)), "retrieve_all")
        return n.value

    def size(self, stream=None) -> int:
        return self._fn("size")(self._h, _stream(stream))

    def load_factor(self, stream=None) -> float:
        return self._fn("load_factor")(self._h, _stream(stream))

    def capacity(self) -> int:
        return self._fn("capacity")(self._h)

    def bytes_total(self) -> int:
        return self._fn("bytes_total")(self._h)

    def peek_status(self, stream=None) -> Status:
        return Status(self._fn("peek_status")(self._h, _stream(stream)))

    def pop_status(self, stream=None) -> Status:
        return Status(self._fn("pop_status")(self._h, _stream(stream)))

    def cg_size(self) -> int:
        return self._fn("cg_size")(self._h)


class HashSet:
    """HashSet<K> (include/warpcore/hash_set.cuh)."""

    def __init__(self, min_capacity: int, key_bits: int = 32, seed: int = DEFAULT_SEED, cg_size: int = 0):
        self.key_bits = key_bits
        self._p = f"wc_hashset_u{key_bits}_"
        self._lib = _capi.lib()
        self._h = self._fn("create")(min_capacity, seed, cg_size)
        if not self._h:
            raise ValueError(f"unsupported configuration (cg_size={cg_size})")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def init(self, seed: int | None = None, stream=None):
        if seed is None:
            _check(self._fn("init")(self._h, _stream(stream)), "init")
        else:
            _check(self._fn("init_seed")(self._h, seed, _stream(stream)), "init")

    def insert(self, keys_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _check(self._fn("insert")(self._h, _ptr(keys_in), keys_in.numel(), _stream(stream), probing_length,
                                  _ptr(status_out)), "insert")

    def contains(self, keys_in, flags_out, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(flags_out, 8, "flags_out")
        _check(self._fn("contains")(self._h, _ptr(keys_in), keys_in.numel(), _ptr(flags_out), _stream(stream),
                                    probing_length, _ptr(status_out)), "contains")

    retrieve = contains

    def erase(self, keys_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _check(self._fn("erase")(self._h, _ptr(keys_in), keys_in.numel(), _stream(stream), probing_length,
                                 _ptr(status_out)), "erase")

    def for_each_count(self, counter_out, stream=None):
        """for_each(f) with f adding one to counter_out[0] per stored key."""
        _width(counter_out, 64, "counter_out")
        _check(self._fn("for_each_count")(self._h, _ptr(counter_out), _stream(stream)), "for_each")

    def retrieve_all(self, keys_out, stream=None) -> int:
        n = ctypes.c_uint64(0)
        _check(self._fn("retrieve_all")(self._h, _ptr(keys_out), ctypes.byref(n), _stream(stream)), "retrieve_all")
        return n.value

    def size(self, stream=None) -> int:
        return self._fn("size")(self._h, _stream(stream))

    def load_factor(self, stream=None) -> float:
        return self._fn("load_factor")(self._h, _stream(stream))

    def capacity(self) -> int:
        return self._fn("capacity")(self._h)

    def bytes_total(self) -> int:
        return self._fn("bytes_total")(self._h)

    def peek_status(self, stream=None) -> Status:
        return Status(self._fn("peek_status")(self._h, _stream(stream)))

    def pop_status(self, stream=None) -> Status:
        return Status(self._fn("pop_status")(self._h, _stream(stream)))

    def cg_size(self) -> int:
        return self._fn("cg_size")(self._h)


class BloomFilter:
    """BloomFilter<K, MurmurHash<K>, Slot, CG> (include/warpcore/bloom_filter.cuh)."""

    def __init__(self, num_bits: int, k: int, key_bits: int = 64, seed: int = DEFAULT_SEED,
                 slot_bits: int = 64, cg_size: int = 1):
        self.key_bits = key_bits
        self._p = f"wc_bloom_u{key_bits}_"
        self._lib = _capi.lib()
        self._h = self._fn("create")(num_bits, k, seed, slot_bits, cg_size)
        if not self._h:
            raise ValueError(f"unsupported configuration (num_bits={num_bits}, k={k}, "
                             f"slot_bits={slot_bits}, cg_size={cg_size})")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def init(self, seed: int | None = None, stream=None):
        if seed is None:
            _check(self._fn("init")(self._h, _stream(stream)), "init")
        else:
            _check(self._fn("init_seed")(self._h, seed, _stream(stream)), "init")

    def insert(self, keys_in, stream=None):
        _width(keys_in, self.key_bits, "keys_in")
        _check(self._fn("insert")(self._h, _ptr(keys_in), keys_in.numel(), _stream(stream)), "insert")

    def retrieve(self, keys_in, flags_out, stream=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(flags_out, 8, "flags_out")
        _check(self._fn("retrieve")(self._h, _ptr(keys_in), keys_in.numel(), _ptr(flags_out), _stream(stream)),
               "retrieve")

    def insert_and_query(self, keys_in, flags_out, stream=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(flags_out, 8, "flags_out")
        _check(self._fn("insert_and_query")(self._h, _ptr(keys_in), keys_in.numel(), _ptr(flags_out),
                                            _stream(stream)), "insert_and_query")

    def words(self, stream=None):
        """Copy of the bit array as a CUDA tensor of num_slots words."""
        dt = torch.int64 if self.slot_bits() == 64 else torch.int32
        out = torch.empty(max(self.num_slots(), 1), dtype=dt, device="cuda")[: self.num_slots()]
        _check(self._fn("copy_words")(self._h, _ptr(out), _stream(stream)), "copy_words")
        return out

    def num_bits(self) -> int:
        return self._fn("num_bits")(self._h)

    def num_slots(self) -> int:
        return self._fn("num_slots")(self._h)

    def num_blocks(self) -> int:
        return self._fn("num_blocks")(self._h)

    def k(self) -> int:
        return self._fn("k")(self._h)

    def seed(self) -> int:
        return self._fn("seed")(self._h)

    def slot_bits(self) -> int:
        return self._fn("slot_bits")(self._h)

    def cg_size(self) -> int:
        return self._fn("cg_size")(self._h)

    def bytes_total(self) -> int:
        return self._fn("bytes_total")(self._h)

    def fpr(self, n: int) -> float:
        return self._fn("fpr")(self._h, n)


class MultiValueHashTable:
    """MultiValueHashTable<K, V> (include/warpcore/multi_value_hash_table.cuh)."""

    _NAMES = {(32, 32): "u32u32", (64, 64): "u64u64"}

    def __init__(self, min_capacity: int, key_bits: int = 32, value_bits: int = 32, seed: int = DEFAULT_SEED,
                 max_values_per_key: int = (1 << 64) - 1, no_init: bool = False, cg_size: int = 0):
        self.key_bits, self.value_bits = key_bits, value_bits
        self._p = f"wc_mvht_{self._NAMES[(key_bits, value_bits)]}_"
        self._lib = _capi.lib()
        self._h = self._fn("create")(min_capacity, seed, max_values_per_key, int(no_init), cg_size)
        if not self._h:
            raise ValueError(f"unsupported configuration (cg_size={cg_size})")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def init(self, stream=None):
        _check(self._fn("init")(self._h, _stream(stream)), "init")

    def insert(self, keys_in, values_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_in, self.value_bits, "values_in")
        _width(status_out, 32, "status_out")
        _check(self._fn("insert")(self._h, _ptr(keys_in), _ptr(values_in), keys_in.numel(), _stream(stream),
                                  probing_length, _ptr(status_out)), "insert")

    def retrieve(self, keys_in, begin_offsets_out, end_offsets_out, values_out, stream=None,
                 probing_length=DEFAULT_PROBING_LENGTH, status_out=None) -> int:
        """Two-pass retrieve; returns the total number of values. values_out=None is a dry run."""
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_out, self.value_bits, "values_out")
        _width(begin_offsets_out, 64, "begin_offsets_out")
        _width(end_offsets_out, 64, "end_offsets_out")
        n = ctypes.c_uint64(0)
        _check(self._fn("retrieve")(self._h, _ptr(keys_in), keys_in.numel(), _ptr(begin_offsets_out),
                                    _ptr(end_offsets_out), _ptr(values_out), ctypes.byref(n), _stream(stream),
                                    probing_length, _ptr(status_out)), "retrieve")
        return n.value

    def num_values(self, keys_in, num_per_key_out=None, stream=None, probing_length=DEFAULT_PROBING_LENGTH,
                   status_out=None) -> int:
        _width(keys_in, self.key_bits, "keys_in")
        _width(num_per_key_out, 64, "num_per_key_out")
        n = ctypes.c_uint64(0)
        _check(self._fn("num_values")(self._h, _ptr(keys_in), keys_in.numel(), ctypes.byref(n),
                                      _ptr(num_per_key_out), _stream(stream), probing_length, _ptr(status_out)),
               "num_values")
        return n.value

    def retrieve_all_keys(self, keys_out, stream=None) -> int:
        n = ctypes.c_uint64(0)
        _check(self._fn("retrieve_all_keys")(self._h, _ptr(keys_out), ctypes.byref(n), _stream(stream)),
               "retrieve_all_keys")
        return n.value

    def retrieve_all(self, keys_out, begin_offsets_out, end_offsets_out, values_out, stream=None):
        nk, nv = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(self._fn("retrieve_all")(self._h, _ptr(keys_out), ctypes.byref(nk), _ptr(begin_offsets_out),
                                        _ptr(end_offsets_out), _ptr(values_out), ctypes.byref(nv),
                                        _stream(stream)), "retrieve_all")
        return nk.value, nv.value

    def num_keys(self, stream=None) -> int:
        return self._fn("num_keys")(self._h, _stream(stream))

    def size(self, stream=None) -> int:
        return self._fn("size")(self._h, _stream(stream))

    def load_factor(self, stream=None) -> float:
        return self._fn("load_factor")(self._h, _stream(stream))

    def storage_density(self, stream=None) -> float:
        return self._fn("storage_density")(self._h, _stream(stream))

    def capacity(self) -> int:
        return self._fn("capacity")(self._h)

    def bytes_total(self) -> int:
        return self._fn("bytes_total")(self._h)

    def max_values_per_key(self) -> int:
        return self._fn("max_values_per_key")(self._h)

    def is_initialized(self) -> bool:
        return bool(self._fn("is_initialized")(self._h))

    def peek_status(self, stream=None) -> Status:
        return Status(self._fn("peek_status")(self._h, _stream(stream)))

    def pop_status(self, stream=None) -> Status:
        return Status(self._fn("pop_status")(self._h, _stream(stream)))

    def cg_size(self) -> int:
        return self._fn("cg_size")(self._h)


class BucketListHashTable:
    """BucketListHashTable<K, V> (include/warpcore/bucket_list_hash_table.cuh)."""

    _NAMES = {(32, 32): "u32u32", (64, 64): "u64u64"}

    def __init__(self, key_capacity: int, value_capacity: int, key_bits: int = 32, value_bits: int = 32,
                 seed: int = DEFAULT_SEED, grow_factor: float = 1.1, min_bucket_size: int = 1,
                 max_bucket_size: int = 1023, max_values_per_key: int = (1 << 20) - 1, no_init: bool = False,
                 cg_size: int = 0):
        self.key_bits, self.value_bits = key_bits, value_bits
        self._p = f"wc_bucketlist_{self._NAMES[(key_bits, value_bits)]}_"
        self._lib = _capi.lib()
        self._h = self._fn("create")(key_capacity, value_capacity, seed, grow_factor, min_bucket_size,
                                     max_bucket_size, max_values_per_key, int(no_init), cg_size)
        if not self._h:
            raise ValueError(f"unsupported configuration (cg_size={cg_size})")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def init(self, seed: int | None = None, stream=None):
        if seed is None:
            _check(self._fn("init")(self._h, _stream(stream)), "init")
        else:
            _check(self._fn("init_seed")(self._h, seed, _stream(stream)), "init")

    def insert(self, keys_in, values_in, stream=None, probing_length=DEFAULT_PROBING_LENGTH, status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_in, self.value_bits, "values_in")
        _width(status_out, 32, "status_out")
        _check(self._fn("insert")(self._h, _ptr(keys_in), _ptr(values_in), keys_in.numel(), _stream(stream),
                                  probing_length, _ptr(status_out)), "insert")

    def retrieve(self, keys_in, begin_offsets_out, end_offsets_out, values_out, stream=None,
                 probing_length=DEFAULT_PROBING_LENGTH, status_out=None) -> int:
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_out, self.value_bits, "values_out")
        n = ctypes.c_uint64(0)
        _check(self._fn("retrieve")(self._h, _ptr(keys_in), keys_in.numel(), _ptr(begin_offsets_out),
                                    _ptr(end_offsets_out), _ptr(values_out), ctypes.byref(n), _stream(stream),
                                    probing_length, _ptr(status_out)), "retrieve")
        return n.value

    def num_values(self, keys_in=None, num_per_key_out=None, stream=None, probing_length=DEFAULT_PROBING_LENGTH,
                   status_out=None) -> int:
        """Values of keys_in (with per-key counts), or of the whole table when keys_in is None."""
        if keys_in is None:
            return self._fn("total_values")(self._h, _stream(stream))
        _width(keys_in, self.key_bits, "keys_in")
        n = ctypes.c_uint64(0)
        _check(self._fn("num_values")(self._h, _ptr(keys_in), keys_in.numel(), ctypes.byref(n),
                                      _ptr(num_per_key_out), _stream(stream), probing_length, _ptr(status_out)),
               "num_values")
        return n.value

    def retrieve_all_keys(self, keys_out, stream=None) -> int:
        n = ctypes.c_uint64(0)
        _check(self._fn("retrieve_all_keys")(self._h, _ptr(keys_out), ctypes.byref(n), _stream(stream)),
               "retrieve_all_keys")
        return n.value

    def retrieve_all(self, keys_out, begin_offsets_out, end_offsets_out, values_out, stream=None):
        nk, nv = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(self._fn("retrieve_all")(self._h, _ptr(keys_out), ctypes.byref(nk), _ptr(begin_offsets_out),
                                        _ptr(end_offsets_out), _ptr(values_out), ctypes.byref(nv),
                                        _stream(stream)), "retrieve_all")
        return nk.value, nv.value

    def num_keys(self, stream=None) -> int:
        return self._fn("num_keys")(self._h, _stream(stream))

    def key_load_factor(self, stream=None) -> float:
        return self._fn("key_load_factor")(self._h, _stream(stream))

    def value_load_factor(self, stream=None) -> float:
        return self._fn("value_load_factor")(self._h, _stream(stream))

    def storage_density(self, stream=None) -> float:
        return self._fn("storage_density")(self._h, _stream(stream))

    def relative_storage_density(self, stream=None) -> float:
        return self._fn("relative_storage_density")(self._h, _stream(stream))

    def key_capacity(self) -> int:
        return self._fn("key_capacity")(self._h)

    def value_capacity(self) -> int:
        return self._fn("value_capacity")(self._h)

    def value_cells_used(self, stream=None) -> int:
        return self._fn("value_cells_used")(self._h, _stream(stream))

    def bytes_total(self) -> int:
        return self._fn("bytes_total")(self._h)

    def max_values_per_key(self) -> int:
        return self._fn("max_values_per_key")(self._h)

    def is_initialized(self) -> bool:
        return bool(self._fn("is_initialized")(self._h))

    def peek_status(self, stream=None) -> Status:
        return Status(self._fn("peek_status")(self._h, _stream(stream)))

    def pop_status(self, stream=None) -> Status:
        return Status(self._fn("pop_status")(self._h, _stream(stream)))

    def cg_size(self) -> int:
        return self._fn("cg_size")(self._h)


class TableBatch:
    """Many small SingleValueHashTables in one arena (no reference
    counterpart; see include/warpcore_c.h). Keys of all tables are passed
    concatenated with a device array of num_tables + 1 offsets."""

    _NAMES = {(32, 32): "u32u32", (64, 64): "u64u64"}

    def __init__(self, min_capacities, key_bits: int = 32, value_bits: int = 32, seed: int = DEFAULT_SEED,
                 cg_size: int = 16):
        import numpy as np
        self.key_bits, self.value_bits = key_bits, value_bits
        self._p = f"wc_batch_{self._NAMES[(key_bits, value_bits)]}_"
        self._lib = _capi.lib()
        caps = np.ascontiguousarray(min_capacities, dtype=np.uint64)
        self._h = self._fn("create")(caps.ctypes.data_as(ctypes.c_void_p), len(caps), seed, cg_size)
        if not self._h:
            raise ValueError(f"unsupported configuration (cg_size={cg_size})")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def init(self, stream=None):
        _check(self._fn("init")(self._h, _stream(stream)), "init")

    def insert(self, keys_in, values_in, offsets, stream=None, probing_length=DEFAULT_PROBING_LENGTH,
               status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_in, self.value_bits, "values_in")
        _width(offsets, 64, "offsets")
        _check(self._fn("insert")(self._h, _ptr(keys_in), _ptr(values_in), _ptr(offsets), keys_in.numel(),
                                  _stream(stream), probing_length, _ptr(status_out)), "insert")

    def retrieve(self, keys_in, offsets, values_out, stream=None, probing_length=DEFAULT_PROBING_LENGTH,
                 status_out=None):
        _width(keys_in, self.key_bits, "keys_in")
        _width(values_out, self.value_bits, "values_out")
        _check(self._fn("retrieve")(self._h, _ptr(keys_in), _ptr(offsets), keys_in.numel(), _ptr(values_out),
                                    _stream(stream), probing_length, _ptr(status_out)), "retrieve")

    def set_launch(self, item_keys: int, blocks_per_cu: int, block_size: int = 256):
        _check(self._fn("set_launch")(self._h, item_keys, blocks_per_cu, block_size), "set_launch")

    def set_lds_insert(self, on: bool):
        """Inserts and lookups in tables of at most 128 KiB of slots run in
        LDS, one block per table (default), or through the work items
        (False)."""
        _check(self._fn("set_lds_insert")(self._h, int(bool(on))), "set_lds_insert")

    def lds_tables(self) -> int:
        """how many tables the LDS insert takes"""
        return self._fn("lds_tables")(self._h)

    def get_launch(self):
        """(keys per work item, blocks per CU, threads per block)"""
        out = (ctypes.c_uint64 * 3)()
        _check(self._fn("get_launch")(self._h, out), "get_launch")
        return tuple(out)

    def tune(self, keys_in, values_in, offsets, queries, query_offsets, values_out, stream=None):
        """Autotuner: per (capacity, fill) class configuration, then launch.
        Returns (launch, tuned ms of the whole workload); leaves the batch
        empty."""
        out = (ctypes.c_uint64 * 3)()
        ms = ctypes.c_double(0)
        _check(self._fn("tune")(self._h, _ptr(keys_in), _ptr(values_in), _ptr(offsets), keys_in.numel(),
                                _ptr(queries), _ptr(query_offsets), queries.numel(), _ptr(values_out),
                                _stream(stream), out, ctypes.byref(ms)), "tune")
        return tuple(out), ms.value

    REPORT_FIELDS = ("class_id", "tables", "keys", "queries", "fixed_config", "chosen_config", "fixed_ns",
                     "chosen_ns", "fixed_slots", "chosen_slots")

    def tune_report(self):
        """Per class of the last tune: dicts of REPORT_FIELDS plus the
        (lanes, window) of both configurations; and the whole-workload
        (fixed ms, tuned ms)."""
        n = self._fn("tune_report")(self._h, None, 0, None)
        _check(0 if n >= 0 else n, "tune_report")
        rows = (ctypes.c_uint64 * (10 * max(n, 1)))()
        totals = (ctypes.c_double * 2)()
        _check(0 if self._fn("tune_report")(self._h, rows, n, totals) >= 0 else -2, "tune_report")
        shapes = self.configs()
        out = []
        for i in range(n):
            r = dict(zip(self.REPORT_FIELDS, rows[10 * i: 10 * i + 10]))
            r["capacity_band"], r["fill_band"] = r["class_id"] // 4, r["class_id"] % 4
            r["fixed_shape"], r["chosen_shape"] = shapes[r["fixed_config"]], shapes[r["chosen_config"]]
            out.append(r)
        return out, (totals[0], totals[1])

    def configs(self):
        """(lanes, window) of every configuration index"""
        out = []
        for c in range(self._fn("num_configs")()):
            lw = (ctypes.c_uint32 * 2)()
            _check(self._fn("config_shape")(c, lw), "config_shape")
            out.append((lw[0], lw[1]))
        return out

    def set_configs(self, configs):
        """Per-table configuration indices; lays the tables out again and
        empties the batch."""
        import numpy as np
        c = np.ascontiguousarray(configs, dtype=np.uint8)
        _check(self._fn("set_configs")(self._h, c.ctypes.data_as(ctypes.c_void_p), len(c)), "set_configs")

    def table_config(self, table: int) -> int:
        c = self._fn("table_config")(self._h, table)
        _check(0 if c >= 0 else c, "table_config")
        return c

    def window(self, table: int) -> int:
        return self.configs()[self.table_config(table)][1]

    def status_words(self, stream=None):
        out = torch.empty(self.num_tables(), dtype=torch.int32, device="cuda")
        _check(self._fn("copy_status_words")(self._h, _ptr(out), _stream(stream)), "copy_status_words")
        return out

    def num_tables(self) -> int:
        return self._fn("num_tables")(self._h)

    def capacity(self, table: int) -> int:
        return self._fn("capacity")(self._h, table)

    def total_slots(self) -> int:
        return self._fn("total_slots")(self._h)

    def bytes_total(self) -> int:
        return self._fn("bytes_total")(self._h)

    def cg_size(self) -> int:
        return self._fn("cg_size")(self._h)


class Server:
    """Serving loop over device-resident SingleValueHashTables (no
    reference counterpart; see include/warpcore_c.h)."""

    _NAMES = {(64, 64): "u64u64", (32, 32): "u32u32"}
    INSERT, RETRIEVE, ERASE = 0, 1, 2

    def __init__(self, table_capacities, key_bits: int = 64, value_bits: int = 64, max_requests: int = 0,
                 item_keys: int = 0, max_load: float = 0.9, use_graph: bool = True, blocks_per_cu: int = 0,
                 max_step_keys: int = 0, direct_keys: int = 0):
        """direct_keys: while the backlog is short, requests of at least this
        many keys run alone on their table's bulk kernels (0: 2^18 keys,
        None: never)."""
        import numpy as np
        self.key_bits, self.value_bits = key_bits, value_bits
        self._p = f"wc_server_{self._NAMES[(key_bits, value_bits)]}_"
        self._lib = _capi.lib()
        caps = np.ascontiguousarray(table_capacities, dtype=np.uint64)
        self._h = self._fn("create")(caps.ctypes.data_as(ctypes.c_void_p), len(caps), max_requests, item_keys,
                                     max_step_keys, max_load, int(use_graph), blocks_per_cu,
                                     (1 << 64) - 1 if direct_keys is None else direct_keys)
        if not self._h:
            raise ValueError("server could not be created")

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._fn("destroy")(h)

    def submit(self, op: int, table: int, keys_in, values_in=None, values_out=None, status_out=None) -> int:
        rid = self._fn("submit")(self._h, op, table, _ptr(keys_in), _ptr(values_in), _ptr(values_out),
                                 _ptr(status_out), keys_in.numel())
        if rid < 0:
            raise ValueError("request refused at submission")
        return rid

    SERVED, REFUSED, FAILED, UNKNOWN_ID = 1, 0, -6, -7

    def wait(self, rid: int) -> bool:
        """True when served, False when refused (out of budget); raises when
        a step of the request failed or the id is unknown."""
        rc = self._fn("wait")(self._h, rid)
        if rc == self.SERVED:
            return True
        if rc == self.REFUSED:
            return False
        raise RuntimeError(f"request {rid}: wait returned {rc}")

    def drain(self):
        _check(self._fn("drain")(self._h), "drain")

    def latency_us(self, rid: int) -> float:
        return self._fn("latency_us")(self._h, rid)

    def step_of(self, rid: int) -> int:
        return self._fn("step_of")(self._h, rid)

    def stats(self):
        out = (ctypes.c_uint64 * 3)()
        _check(self._fn("stats")(self._h, out), "stats")
        c = (ctypes.c_uint64 * 4)()
        _check(self._fn("counters")(self._h, c), "counters")
        return {"requests": out[0], "steps": out[1], "keys": out[2],
                "direct_steps": c[0], "compactions": c[1], "failed": c[2]}

    def table_reserved(self, table: int) -> int:
        return self._fn("table_reserved")(self._h, table)

    def table_budget(self, table: int) -> int:
        return self._fn("table_budget")(self._h, table)

    def fail_step(self, step: int):
        """Failure handling from the host: step `step` (1, 2, ...) is not
        launched and reports a launch failure."""
        _check(self._fn("fail_step")(self._h, step), "fail_step")

    def wait_code(self, rid: int) -> int:
        """wait() as the C ABI reports it: 1 served, 0 refused, -6 failed,
        -7 unknown id."""
        return self._fn("wait")(self._h, rid)

    def table_size(self, table: int) -> int:
        return self._fn("table_size")(self._h, table)

    def table_capacity(self, table: int) -> int:
        return self._fn("table_capacity")(self._h, table)

    def table_status(self, table: int) -> Status:
        return Status(self._fn("table_status")(self._h, table))

    def drive(self, mode: int, ops, tables, sizes, arrival_us, keys, values, values_out, status_out=None):
        """Replay a request list (numpy ops/tables/sizes/arrival_us; device
        keys/values/values_out/status_out concatenated). mode 0: server,
        mode 1: one plain launch per request. Returns (latencies_us, makespan_us)."""
        import numpy as np
        ops = np.ascontiguousarray(ops, np.uint32)
        tables = np.ascontiguousarray(tables, np.uint32)
        sizes = np.ascontiguousarray(sizes, np.uint64)
        arrival = np.ascontiguousarray(arrival_us, np.float64)
        lat = np.zeros(len(ops), np.float64)
        mk = ctypes.c_double(0)
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _check(self._fn("drive")(self._h, mode, vp(ops), vp(tables), vp(sizes), vp(arrival), len(ops), _ptr(keys),
                                 _ptr(values), _ptr(values_out), _ptr(status_out), vp(lat), ctypes.byref(mk)),
               "drive")
        return lat, mk.value


def load_binary(path, key_bits: int = 64, max_keys: int = 1 << 27):
    """Keys of a raw little-endian key file (the reference benchmarks'
    load_keys input format, benchmarks/include/benchmark_common.cuh:62-75),
    at most max_keys from the start, as a numpy array (host)."""
    import numpy as np
    dt = {32: np.uint32, 64: np.uint64}[key_bits]
    out = np.empty(max_keys, dt)
    fn = getattr(_capi.lib(), f"wc_load_binary_u{key_bits}")
    n = fn(str(path).encode(), max_keys, out.ctypes.data_as(ctypes.c_void_p))
    if n < 0:
        raise OSError(f"cannot read key file {path}")
    return out[:n]


def load_keys(path, key_bits: int = 64, max_keys: int = 1 << 27, stream=None):
    """The same keys copied to a device tensor of the keys read (int32 or
    int64 holding the key bits)."""
    import torch
    buf = torch.empty(max_keys, dtype={32: torch.int32, 64: torch.int64}[key_bits], device="cuda")
    num = ctypes.c_uint64(0)
    fn = getattr(_capi.lib(), f"wc_load_keys_u{key_bits}")
    _check(fn(str(path).encode(), max_keys, _ptr(buf), ctypes.byref(num), _stream(stream)), "load_keys")
    return buf[: num.value]
