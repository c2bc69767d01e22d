"""bench.py's host-side helpers (no GPU): algorithmic bytes per operation,
the kernel each selection launches, and that a traffic figure is only read
for the kernel it was measured on."""
import bench


class _Lib:
    def __init__(self, mask):
        self.mask = mask

    def wc_get_group_kernel(self):
        return self.mask


def test_algorithmic_bytes():
    assert bench.algorithmic_bytes("retrieve", 64, 64) == 144
    assert bench.algorithmic_bytes("insert", 64, 64) == 272
    assert bench.algorithmic_bytes("insert", 32, 32) == 264


def test_kernel_family_follows_selection():
    assert bench.kernel_family(_Lib(220), "insert") == "insert_owner"
    assert bench.kernel_family(_Lib(220), "retrieve") == "retrieve_owner"
    assert bench.kernel_family(_Lib(28), "insert") == "insert_refill"
    assert bench.kernel_family(_Lib(28), "retrieve") == "retrieve_refill_pipe"
    assert bench.kernel_family(_Lib(12), "retrieve") == "retrieve_refill"


def test_traffic_only_for_the_measured_kernel():
    assert bench.read_traffic("no_such_kernel") is None
    assert bench.read_traffic("insert_owner", log2n=24) is None


def test_probe_library_binding():
    from warpcore_amd.build import PROBES_LIB, build_probes
    if not PROBES_LIB.exists():
        build_probes()
    lib = bench.probes_lib()
    assert hasattr(lib, "wc_util_gather_lines")
