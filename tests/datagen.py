"""Seeded synthetic inputs shared by the tests (numpy, host side)."""
from __future__ import annotations

import numpy as np

EMPTY = 0


def tomb(bits: int) -> int:
    return (1 << bits) - 1


def dtype(bits: int):
    return np.uint32 if bits == 32 else np.uint64


def unique_keys(n: int, bits: int, seed: int, exclude: np.ndarray | None = None) -> np.ndarray:
    """n distinct keys uniform on [1, 2^bits - 2] (no EMPTY, no TOMBSTONE), shuffled."""
    rng = np.random.default_rng(seed)
    dt = dtype(bits)
    hi = tomb(bits)  # exclusive upper bound keeps ~0 out
    out = np.empty(0, dt)
    while len(out) < n:
        need = int((n - len(out)) * 1.1) + 16
        draw = rng.integers(1, hi, size=need, dtype=dt, endpoint=False)
        out = np.unique(np.concatenate([out, draw]))
        if exclude is not None and len(exclude):
            out = out[~np.isin(out, exclude)]
    out = out[:n]
    rng.shuffle(out)
    return out


def uniform_values(n: int, bits: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << bits, size=n, dtype=dtype(bits), endpoint=False)


def zipf_indices(n: int, universe: int, s: float, seed: int) -> np.ndarray:
    """n draws from Zipf(s) over {0..universe-1} by inverse CDF on the exact pmf."""
    rng = np.random.default_rng(seed)
    ranks = np.arange(1, universe + 1, dtype=np.float64)
    w = ranks ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = rng.random(n)
    return np.minimum(np.searchsorted(cdf, u, side="right"), universe - 1)


def to_dev(a: np.ndarray):
    import torch
    if a.dtype == np.uint32:
        v = a.view(np.int32)
    elif a.dtype == np.uint64:
        v = a.view(np.int64)
    elif a.dtype == np.uint8 or a.dtype == np.bool_:
        v = a.view(np.uint8)
    else:
        v = a
    return torch.from_numpy(np.ascontiguousarray(v)).to("cuda")


def from_dev(t, dt) -> np.ndarray:
    return t.cpu().numpy().view(dt)


def dev_empty(n: int, bits: int, fill: int = 0):
    import torch
    tdt = {8: torch.uint8, 32: torch.int32, 64: torch.int64}[bits]
    t = torch.empty(max(n, 1), dtype=tdt, device="cuda")
    if fill == 0:
        t.zero_()
    else:
        # fill with the bit pattern `fill`
        a = np.full(max(n, 1), fill, dtype=dtype(bits) if bits in (32, 64) else np.uint8)
        t.copy_(to_dev(a))
    return t[:n] if n > 0 else t[:0]


class kernel_shape:
    """Context manager selecting the bulk-kernel shape: 'lanes' (one lane per
    key), 'groups' (cooperative groups, streaming kernels), 'lockstep'
    (cooperative groups, lockstep kernels), 'refill' (cooperative groups,
    refill inserts and lookups), 'pipe' (refill inserts, pipelined refill
    lookups: each slot's result consumed as soon as its own request
    returns) or
    'owner' (lane-owner kernels: every lane owns one key, groups read their
    owners' windows in turn) or 'pairs' (lane-owner kernels reading packed
    8-byte slots two per lane, the default), by
    moving the table-size limit of
    the lane-per-key path and choosing the group kernels."""

    def __init__(self, shape: str):
        self.limit = (1 << 62) if shape == "lanes" else 0
        self.group_kernel = {"groups": 0, "refill": 12, "pipe": 28, "owner": 220, "pairs": 476}.get(shape, 3)

    def __enter__(self):
        from warpcore_amd import _capi
        self.lib = _capi.lib()
        self.saved = (self.lib.wc_get_lane_table_bytes(), self.lib.wc_get_group_kernel(),
                      self.lib.wc_get_lane_insert_table_bytes())
        self.lib.wc_set_lane_table_bytes(self.limit)
        self.lib.wc_set_lane_insert_table_bytes(self.limit)
        self.lib.wc_set_group_kernel(self.group_kernel)
        return self

    def __exit__(self, *exc):
        self.lib.wc_set_lane_table_bytes(self.saved[0])
        self.lib.wc_set_lane_insert_table_bytes(self.saved[2])
        self.lib.wc_set_group_kernel(self.saved[1])
        return False
