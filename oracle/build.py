"""Build the oracle (test infrastructure): oracle/liboracle.so via make."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    subprocess.run(["make", "-B", "-C", str(HERE), "liboracle.so"], check=True,
                   capture_output=True, text=True)
    return HERE / "liboracle.so"


if __name__ == "__main__":
    print(build())
