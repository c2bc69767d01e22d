"""MI355X-native warpcore hash-table hot path.

The product is the header-only HIP C++ library under ``include/warpcore``
and the C ABI in ``libwarpcore_amd.so`` (declared in the repository's
``include/warpcore_c.h``). This package exposes that ABI to Python through
ctypes and mirrors the reference's host API in :mod:`warpcore_amd.tables`.
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
INCLUDE_DIR = PACKAGE_DIR / "include"
LIBRARY_PATH = PACKAGE_DIR / "libwarpcore_amd.so"

__all__ = ["PACKAGE_DIR", "INCLUDE_DIR", "LIBRARY_PATH"]
