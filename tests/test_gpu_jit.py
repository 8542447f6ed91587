"""Generated pass kernels on the device: a plan whose fused passes run as
kernels generated and compiled per pass (ALEPH_HIP_PLAN_JIT) must give the
same amplitudes as the compiled op-loop kernel for the same plan (to
rounding: 1e-14 FP64, 1e-6 FP32), and the
unfused oracle result (FP64 1e-12, FP32 1e-5), across planner options and
controlled ops.
"""

import numpy as np
import pytest

from aleph_amd import workloads as W
from aleph_amd.backend import HipSvBackend
from oracle.oracle import OracleSv
from tests.helpers import PARITY_TOL, max_abs
from tests.test_oracle import random_circuit

pytestmark = pytest.mark.gpu


def _run(b, c, psi, **kw):
    plan = b.compile(c, **kw)
    st = b.allocate(c.num_qubits)
    st.load(psi)
    plan.execute(st)
    out = st.amplitudes()
    plan.free()
    st.free()
    return out


# (name, seed of the random-circuit generator, builder, planner options)
CASES = [
    ("brickwall", 11, lambda rng: W.brickwall(18, 10, seed=4), {}),
    ("hea", 12, lambda rng: W.hea(16, 4, seed=2), {}),
    ("qaoa", 13, lambda rng: W.qaoa_maxcut(14, W.random_3_regular(14, 5), [0.4, 1.1], [0.7, 0.2]), {}),
    ("random_ctl", 14, lambda rng: random_circuit(14, 120, rng, kq=False), {}),
    ("random_m10", 15, lambda rng: random_circuit(13, 120, rng, kq=False), {"local_qubits": 10}),
    ("random_g5", 16, lambda rng: random_circuit(14, 120, rng, kq=False), {"reg_bits": 5, "local_qubits": 12}),
    ("brickwall_nt", 17, lambda rng: W.brickwall(17, 8, seed=9), {"nontemporal": True, "fuse_2q": False}),
    ("hea_m10", 18, lambda rng: W.hea(15, 4, seed=6), {"local_qubits": 10}),
]


@pytest.mark.parametrize("precision,tol", [(64, PARITY_TOL), (32, 1e-5)])
@pytest.mark.parametrize("name,seed,make,kw", CASES, ids=[c[0] for c in CASES])
def test_generated_kernels_equal_compiled_kernels(precision, tol, name, seed, make, kw):
    rng = np.random.default_rng(seed)
    c = make(rng)
    psi = W.gaussian_state(c.num_qubits, 7)
    b = HipSvBackend(seed=3, precision=precision)
    jit = _run(b, c, psi, jit=True, **kw)
    ref = _run(b, c, psi, **kw)
    # Same arithmetic up to the compiler's FMA contraction choices.
    assert max_abs(jit, ref) <= (1e-14 if precision == 64 else 1e-6), f"{name} seed={seed}"
    o = OracleSv(cap=30)
    st = o.allocate(c.num_qubits)
    st.load(psi)
    o.run(c, st)
    assert max_abs(jit, st.amplitudes()) < tol, f"{name} seed={seed}"


def test_fp32_matrix_core_4x4_ops_equal_vector_ops(monkeypatch):
    """ALEPH_HIP_FP32_MFMA=1: FP32 4x4 ops of generated kernels on the matrix
    cores (v_mfma_f32_4x4x1f32, each lane's quartet its own column) give the
    vector-ALU result to FP32 rounding and the oracle to 1e-5."""
    c = W.brickwall(16, 8, seed=21)
    psi = W.gaussian_state(16, 3)
    b = HipSvBackend(seed=3, precision=32)
    ref = _run(b, c, psi, jit=True)
    monkeypatch.setenv("ALEPH_HIP_FP32_MFMA", "1")
    mf = _run(b, c, psi, jit=True)
    assert max_abs(mf, ref) <= 1e-6
    o = OracleSv(cap=30)
    st = o.allocate(16)
    st.load(psi)
    o.run(c, st)
    assert max_abs(mf, st.amplitudes()) < 1e-5
