"""Generated pass kernels, host side: the source written for every fused
pass compiles with hiprtc for gfx950 against the device headers the library
embeds, and lists exactly the pass's ops in order (one call per op, stage
exchanges between them).  Compilation needs no GPU.
"""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from aleph_amd import workloads as W
from aleph_amd.plan_tools import HostPlan
from tests.test_oracle import random_circuit

HIPRTC = Path("/opt/rocm/lib/libhiprtc.so")
CSRC = Path(__file__).resolve().parent.parent / "aleph_amd" / "csrc"
HEADERS = ["device_types.hpp", "plan.hpp", "fused_chunk.hpp"]
OP_CALL = {1: "op_1q<", 2: "op_2q<", 3: "op_diag<", 4: "op_sign<", 5: "op_phase<"}


def _compile(src: str) -> tuple[int, str]:
    rtc = C.CDLL(str(HIPRTC))
    prog = C.c_void_p()
    names = (C.c_char_p * len(HEADERS))(*[h.encode() for h in HEADERS])
    texts = (C.c_char_p * len(HEADERS))(*[(CSRC / h).read_bytes() for h in HEADERS])
    assert rtc.hiprtcCreateProgram(C.byref(prog), src.encode(), b"aleph_pass.hip", len(HEADERS), texts, names) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-ffp-contract=fast",
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    rc = rtc.hiprtcCompileProgram(prog, len(opts), (C.c_char_p * len(opts))(*opts))
    n = C.c_size_t()
    rtc.hiprtcGetProgramLogSize(prog, C.byref(n))
    log = C.create_string_buffer(n.value + 1)
    rtc.hiprtcGetProgramLog(prog, log)
    rtc.hiprtcDestroyProgram(C.byref(prog))
    return rc, log.value.decode(errors="replace")


def _plans():
    rng = np.random.default_rng(11)
    return [
        HostPlan(W.brickwall(16, 8, seed=3), precision=64),
        HostPlan(W.hea(14, 3, seed=1), precision=32),
        HostPlan(random_circuit(12, 60, rng, kq=False), precision=64, local_qubits=10),
        HostPlan(random_circuit(13, 60, rng, kq=False), precision=32, reg_bits=5, nontemporal=True),
    ]


def test_sources_list_the_pass_ops_in_order():
    for hp in _plans():
        for i, p in enumerate(hp.passes):
            src = hp.jit_sources[i]
            stages = hp.stages[p.stage_begin:p.stage_begin + p.n_stages]
            want = [OP_CALL[hp.ops[o].type] for sd in stages for o in range(sd.op_begin, sd.op_end)]
            got = re.findall(r"(op_\w+<)", src)
            assert got == want
            assert src.count("io.load(a)") == 1 and src.count("io.store(a)") == 1
            if hp.precision == 64 and p.n_stages > 1:
                # FP64: real and imaginary parts exchanged in turn through half the LDS
                assert src.count("io.exchange_split(a, ") == p.n_stages - 1
                assert src.count("io.enter(") == 1 and "io.put(" not in src
                assert f"__launch_bounds__({1 << (p.m - p.reg_bits)}, 4)" in src
            else:
                assert src.count("io.enter(") == p.n_stages
                assert src.count("io.put(a)") == p.n_stages - 1 == src.count("io.get(a)")
                assert f"__launch_bounds__({1 << (p.m - p.reg_bits)}, 2)" in src



@pytest.mark.skipif(not HIPRTC.exists(), reason="hiprtc not installed")
def test_generated_sources_compile_for_gfx950():
    seen = set()
    for hp in _plans():
        for src in hp.jit_sources:
            if src in seen or "op_phase<" in src:
                continue
            seen.add(src)
            if len(seen) > 6:
                break
            rc, log = _compile(src)
            assert rc == 0, log
    assert len(seen) >= 4
