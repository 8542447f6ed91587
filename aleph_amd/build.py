"""Build the gfx950 shared library `aleph_amd/libaleph_hip.so` with hipcc.

Every source compiles for gfx950 only, in-tree; objects go to
`aleph_amd/_build/`.  Packed FP32 instructions are disabled for device
code (`-packed-fp32-ops`), and floating-point atomics use the hardware
instructions.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import time
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libaleph_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["capi.cpp", "engine.cpp", "batch.cpp", "jit.cpp", "k_gates.hip", "k_fused.hip", "k_kq_mfma.hip", "k_batch.hip", "k_mlp.hip", "k_readout.hip", "noise.cpp", "k_noise.hip", "k_exchange.hip"]

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    "-ffp-contract=fast",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unknown-pragmas",
]


def _newest_input() -> float:
    """Newest modification time among the headers every source may include."""
    paths = [*CSRC.glob("*.hpp"), *(PKG.parent / "include").glob("*.h"), BUILD / "jit_headers.inc", Path(__file__)]
    return max(p.stat().st_mtime for p in paths if p.exists())


def _compile(src: str, incremental: bool = False, headers_mtime: float = 0.0) -> tuple[str, str, float]:
    out = BUILD / (src + ".o")
    if incremental and out.exists():
        built = out.stat().st_mtime
        if built > (CSRC / src).stat().st_mtime and built > headers_mtime:
            return str(out), "", 0.0
    cmd = [HIPCC, *COMMON, "-I", str(PKG.parent / "include"), "-I", str(BUILD), "-c", str(CSRC / src), "-o", str(out)]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return str(out), res.stderr, time.time() - t0


# Device headers the generated pass kernels compile against (hiprtc sees
# only these), embedded into the library as string literals.
JIT_HEADERS = ["device_types.hpp", "plan.hpp", "fused_chunk.hpp"]


def _write_jit_headers() -> None:
    delim = "ALEPH_JIT_HDR"
    lines = [f"static const int kJitHeaderCount = {len(JIT_HEADERS)};",
             "static const char* kJitHeaderNames[] = {" + ", ".join(f'"{h}"' for h in JIT_HEADERS) + "};",
             "static const char* kJitHeaderSources[] = {"]
    for h in JIT_HEADERS:
        text = (CSRC / h).read_text()
        if f"){delim}" in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter")
        lines.append(f'R"{delim}({text}){delim}",')
    lines.append("};")
    out = BUILD / "jit_headers.inc"
    new = "\n".join(lines) + "\n"
    if not out.exists() or out.read_text() != new:
        out.write_text(new)


def build(verbose: bool = False, incremental: bool = False) -> Path:
    """Compile every source (or, with `incremental`, only those older than
    their inputs) and link the library."""
    BUILD.mkdir(exist_ok=True)
    _write_jit_headers()
    hdr = _newest_input()
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(lambda s: _compile(s, incremental, hdr), SOURCES))
    objs = [r[0] for r in results]
    if verbose:
        for src, (_, err, secs) in zip(SOURCES, results):
            if err.strip():
                print(err, file=sys.stderr)
            print(f"{src}: {secs:.1f} s", file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
