"""Key files: raw little-endian key arrays (the reference benchmarks'
load_keys format, benchmarks/include/benchmark_common.cuh:62-75), read
through the C ABI and compared with numpy's reading of the same bytes."""
import numpy as np
import pytest


@pytest.mark.parametrize("bits", [32, 64])
def test_load_binary_reads_le_keys(tmp_path, bits):
    from warpcore_amd.tables import load_binary
    dt = {32: np.uint32, 64: np.uint64}[bits]
    keys = np.random.default_rng(bits).integers(0, np.iinfo(dt).max, 1000, dtype=dt)
    f = tmp_path / "keys.bin"
    keys.astype(np.dtype(dt).newbyteorder("<")).tofile(f)
    np.testing.assert_array_equal(load_binary(f, bits, 1 << 12), keys)          # file shorter than max
    np.testing.assert_array_equal(load_binary(f, bits, 100), keys[:100])         # max shorter than file
    assert len(load_binary(f, bits, 0)) == 0
    # a trailing partial key is ignored
    with open(f, "ab") as fh:
        fh.write(b"\x01\x02\x03")
    np.testing.assert_array_equal(load_binary(f, bits, 1 << 12), keys)
    # byte layout: little-endian
    raw = f.read_bytes()
    assert int.from_bytes(raw[: bits // 8], "little") == int(keys[0])


def test_load_binary_missing_file(tmp_path):
    from warpcore_amd.tables import load_binary
    with pytest.raises(OSError):
        load_binary(tmp_path / "absent.bin", 64, 10)


@pytest.mark.gpu
def test_load_keys_to_device_and_insert(cuda, tmp_path):
    from warpcore_amd.tables import SingleValueHashTable, load_keys
    from datagen import from_dev, unique_keys
    keys = unique_keys(5000, 64, seed=9)
    f = tmp_path / "keys.bin"
    keys.tofile(f)
    d = load_keys(f, 64, 1 << 13)
    assert d.numel() == 5000
    np.testing.assert_array_equal(from_dev(d, np.uint64), keys)
    t = SingleValueHashTable(1 << 14, 64, 64)
    t.insert(d, d)
    out = d.clone()
    t.retrieve(d, out)
    np.testing.assert_array_equal(from_dev(out, np.uint64), keys)
    assert t.size() == 5000
