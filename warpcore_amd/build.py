"""Build libwarpcore_amd.so (gfx950) in-tree.

Every ``csrc/*.hip`` translation unit is compiled with hipcc for gfx950 in
parallel and linked into ``warpcore_amd/libwarpcore_amd.so``. Object files
go to ``warpcore_amd/_build/``. The measurement probes of the benchmark
(``probes/*.hip``) go into their own ``warpcore_amd/probes/
libwarpcore_probes.so``. A build always recompiles.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG / "include"
BUILD = PKG / "_build"
LIB = PKG / "libwarpcore_amd.so"
PROBES = PKG / "probes"
PROBES_LIB = PROBES / "libwarpcore_probes.so"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    "-Wall",
    "-Wno-unused-function",
]


def _compile(src: Path) -> tuple[Path, str]:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    log = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{log}")
    return obj, log


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    """Compile all HIP sources and link the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, sources))
    objs = [o for o, _ in results]
    if verbose:
        for (_, log), src in zip(results, sources):
            if log.strip():
                print(f"[{src.name}]\n{log}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_probes() -> Path:
    """Compile the benchmark's measurement probes into their own library."""
    tmp = PROBES_LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", *map(str, sorted(PROBES.glob("*.hip"))), "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for the probes:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, PROBES_LIB)
    return PROBES_LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
    print(build_probes())
