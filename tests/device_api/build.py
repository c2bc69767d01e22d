"""Builds tests/device_api/libwc_device_api.so: user kernels that call the
tables' device API (test infrastructure, gfx950, header-only library)."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
LIB = HERE / "libwc_device_api.so"


def build() -> Path:
    from warpcore_amd.build import HIPCC, HIP_FLAGS
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", str(HERE / "device_api.hip"), "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for device_api.hip:\n{proc.stdout}{proc.stderr}")
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    sys.path.insert(0, str(REPO))
    print(build())
