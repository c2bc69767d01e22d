"""ctypes binding of the C ABI declared in include/warpcore_c.h.

The prototypes are read from the header itself, so the Python side and the
C side cannot drift apart. Loading fails loudly when the HIP library has not
been built: there is no CPU fallback for any table operation.
"""
from __future__ import annotations

import ctypes
import re
from functools import lru_cache
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
HEADER = REPO / "include" / "warpcore_c.h"
LIB_PATH = PKG / "libwarpcore_amd.so"

_SCALARS = {
    "void": None,
    "int": ctypes.c_int,
    "uint32_t": ctypes.c_uint32,
    "uint64_t": ctypes.c_uint64,
    "int64_t": ctypes.c_int64,
    "float": ctypes.c_float,
    "double": ctypes.c_double,
    "bool": ctypes.c_bool,
}

_PROTO = re.compile(r"WC_API\s+([\w\s\*]+?)\s*\b(wc_\w+)\s*\(([^)]*)\)\s*;")


def _ctype(decl: str):
    decl = decl.replace("const", "").strip()
    if "*" in decl:
        return ctypes.c_void_p
    base = decl.split()[0] if decl else "void"
    if base not in _SCALARS:
        raise ValueError(f"unsupported C type in warpcore_c.h: {decl!r}")
    return _SCALARS[base]


def parse_header(text: str | None = None) -> dict[str, tuple]:
    """Return {name: (restype, [argtypes])} for every WC_API prototype."""
    text = HEADER.read_text() if text is None else text
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in _PROTO.finditer(" ".join(text.split())):
        ret, name, args = m.group(1), m.group(2), m.group(3).strip()
        argtypes = []
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                # drop the parameter name
                typ = re.sub(r"\b\w+\s*$", "", a) if not a.endswith("*") else a
                argtypes.append(_ctype(typ))
        out[name] = (_ctype(ret), argtypes)
    return out


@lru_cache(maxsize=1)
def lib() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: the HIP library must be built first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    handle = ctypes.CDLL(str(LIB_PATH))
    for name, (restype, argtypes) in parse_header().items():
        fn = getattr(handle, name)  # AttributeError = a declared symbol is not exported
        fn.restype = restype
        fn.argtypes = argtypes
    return handle
