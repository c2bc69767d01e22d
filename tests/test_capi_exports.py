"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes

from warpcore_amd import _capi


def _library_path():
    if not _capi.LIB_PATH.exists():
        from warpcore_amd.build import build
        build()
    return _capi.LIB_PATH


def test_header_parses():
    protos = _capi.parse_header()
    assert len(protos) > 50
    assert "wc_svht_u32u32_insert" in protos
    ret, args = protos["wc_svht_u64u64_retrieve"]
    assert ret is ctypes.c_int and len(args) == 7


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(_library_path()))
    missing = [n for n in _capi.parse_header() if not hasattr(lib, n)]
    assert missing == []


def test_binding_loads():
    _library_path()
    _capi.lib.cache_clear()
    lib = _capi.lib()
    assert lib.wc_svht_u32u32_create.restype is ctypes.c_void_p
