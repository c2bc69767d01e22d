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


def test_probes_library_exports_its_header():
    """The benchmark's measurement probes live in their own library, not in
    the table library's C ABI."""
    from warpcore_amd.build import PROBES, PROBES_LIB, build_probes
    if not PROBES_LIB.exists():
        build_probes()
    protos = _capi.parse_header((PROBES / "warpcore_probes.h").read_text())
    assert set(protos) == {"wc_util_copy", "wc_util_gather_lines", "wc_util_claim_lines"}
    lib = ctypes.CDLL(str(PROBES_LIB))
    assert all(hasattr(lib, n) for n in protos)
    assert not any(n.startswith("wc_util_") for n in _capi.parse_header())


def test_noalloc_entry_point_declared():
    """The insert that allocates nothing is declared beside the others, with
    its size limit exported."""
    protos = _capi.parse_header()
    ret, args = protos["wc_counting_u64_insert_noalloc"]
    assert ret is ctypes.c_int and len(args) == 6
    lib = ctypes.CDLL(str(_library_path()))
    lib.wc_counting_u64_insert_noalloc_max.restype = ctypes.c_uint64
    assert lib.wc_counting_u64_insert_noalloc_max() == 1 << 20
