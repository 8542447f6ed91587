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
OP_CALL = {1: "op_1q<", 2: "op_2q<", 3: "op_diag<", 4: "op_sign<", 5: "phase"}
# A phase op is written out as a block of per-class factors (phase_factor,
# from the workgroup's uniform table) and slot multiplies, or, with terms on
# three or more register bits, as a call of the generic op_phase.
# A 4x4 op whose matrix is a unit permutation (SWAP, CNOT) renames register
# slots instead ("// slot renaming").
OP_RE = r"(op_\w+<|const DevOp\* po|// slot renaming)"
# With ALEPH_HIP_FP32_MFMA=1, FP32 uncontrolled 4x4 ops run on the matrix cores (op_2q_mfma).


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
        HostPlan(W.qft(16), precision=64),
        HostPlan(W.qft(13), precision=64, fuse_2q=False),
        HostPlan(W.qft(14), precision=32, reg_bits=5),
        HostPlan(W.qaoa_maxcut(14, W.random_3_regular(14, 5), [0.4, 1.1], [0.7, 0.2]), precision=64),
    ]


def test_sources_list_the_pass_ops_in_order():
    for hp in _plans():
        for i, p in enumerate(hp.passes):
            src = hp.jit_sources[i]
            stages = hp.stages[p.stage_begin:p.stage_begin + p.n_stages]
            want = [OP_CALL[hp.ops[o].type] for sd in stages for o in range(sd.op_begin, sd.op_end)]
            alias = {"op_phase<": "phase", "const DevOp* po": "phase", "// slot renaming": "op_2q<",
                     "op_2q_mfma<": "op_2q<"}
            got = [alias.get(g, g) for g in re.findall(OP_RE, src)]
            assert got == want
            n_fast = sum(1 for sd in stages for o in range(sd.op_begin, sd.op_end)
                         if hp.ops[o].type == 5 and hp.ops[o].b == 0)
            assert src.count("const DevOp* po") == n_fast
            assert (n_fast > 0) == ("phase_class_uniform<" in src)
            assert src.count("io.load(a)") == 1 and src.count("io.store(a)") == 1
            if hp.precision == 64 and p.n_stages > 1:
                # FP64: real and imaginary parts exchanged in turn through half the LDS
                assert src.count("io.exchange_split(a, ") == p.n_stages - 1
                assert src.count("io.enter(") == 1 and "io.put(" not in src
                # four waves per SIMD (128 registers), two for passes with phase ops
                waves = 2 if n_fast else 4
                assert f"__launch_bounds__({1 << (p.m - p.reg_bits)}, {waves})" in src
            else:
                assert src.count("io.enter(") == p.n_stages
                assert src.count("io.put(a)") == p.n_stages - 1 == src.count("io.get(a)")
                assert f"__launch_bounds__({1 << (p.m - p.reg_bits)}, 2)" in src



@pytest.mark.skipif(not HIPRTC.exists(), reason="hiprtc not installed")
def test_generated_sources_compile_for_gfx950():
    seen = set()
    for hp in _plans():
        for src in hp.jit_sources:
            if src in seen:
                continue
            seen.add(src)
            if len(seen) > 9:
                break
            rc, log = _compile(src)
            assert rc == 0, log
    assert len(seen) >= 4
    assert any("phase_class_uniform<" in s for s in seen)
