"""Build the CPU oracle `oracle/liboracle_sv.so` (test infrastructure only).

g++ with OpenMP; x86-64-v3 so the library also runs on the GPU hosts' CPUs.
The reference itself (Rust) has no toolchain here and is not built.
"""

from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ["sv_oracle.cpp"]
LIB = HERE / "liboracle_sv.so"


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall",
           "-o", str(tmp), *[str(HERE / s) for s in SOURCES]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())
