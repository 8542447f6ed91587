"""Dense k-qubit blocks on the device against the CPU oracle.

Widths 4..6 run on the matrix cores once the state has at least k + 4
qubits (16 groups per wavefront tile); smaller states and widths 2..3 use
the vector kernel.  Every width, qubit order (including non-adjacent and
descending targets) and control configuration is checked at 1e-12 (FP64)
and 1e-5 (FP32) against the oracle's gather/matvec/scatter restatement.
"""

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd import workloads as W
from aleph_amd.circuit import Circuit, Gate, GateInstance
from tests.helpers import PARITY_TOL, embed, max_abs, random_unitary

pytestmark = pytest.mark.gpu


def oracle_apply(orc, n, gates, psi):
    o = orc.OracleSv()
    st = o.allocate(n)
    st.load(psi)
    for g in gates:
        o.apply_gate(st, g)
    return st.amplitudes()


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("n", [9, 12])
def test_dense_blocks_match_oracle(hip, hip32, oracle_lib, k, n):
    if k > n:
        pytest.skip("block wider than state")
    rng = np.random.default_rng(100 * k + n)
    gates = []
    for rep in range(3):
        qs = [int(q) for q in rng.choice(n, k, replace=False)]
        if rep == 1:
            qs = sorted(qs, reverse=True)
        gates.append(GateInstance(Gate.unitary_kq(random_unitary(1 << k, rng)), qs))
    psi = W.gaussian_state(n, k)
    want = oracle_apply(oracle_lib, n, gates, psi)
    for backend, tol in ((hip, PARITY_TOL), (hip32, 1e-5)):
        st = backend.allocate(n)
        st.load(psi)
        for g in gates:
            backend.apply_gate(st, g)
        assert max_abs(st.amplitudes(), want) < tol, (backend.precision, k, n)
        st.free()


@pytest.mark.parametrize("k", [4, 5, 6])
def test_controlled_dense_blocks_match_brute_force(hip, hip32, k):
    """Controlled blocks through the low-level entry point (the gate-level
    API rejects controls on UnitaryKq, as the reference does): the state
    after the block equals the full 2^n operator built by enumerating basis
    states, applied in numpy."""
    n = k + 5
    rng = np.random.default_rng(k)
    qs = [int(q) for q in rng.choice(n, k, replace=False)]
    ctl = [int(q) for q in rng.choice([q for q in range(n) if q not in qs], 2, replace=False)]
    U = random_unitary(1 << k, rng)
    psi = W.gaussian_state(n, 11)
    want = embed(U, qs, ctl, n) @ psi
    mat = np.ascontiguousarray(U.reshape(-1).view(np.float64))
    qarr = np.ascontiguousarray(np.array(qs, dtype=np.uint32))
    cmask = sum(1 << c for c in ctl)
    for backend, tol in ((hip, PARITY_TOL), (hip32, 1e-5)):
        st = backend.allocate(n)
        st.load(psi)
        backend._check(L.load().aleph_hip_apply_kq(backend._ctx, st.handle, k, qarr.ctypes.data_as(L.PU32), cmask,
                                                   mat.ctypes.data_as(L.PF64)))
        assert max_abs(st.amplitudes(), want) < tol, (backend.precision, k)
        st.free()


def test_dense_block_plan_step_at_n22(hip, oracle_lib):
    """A k=6 block inside a planned circuit at a size where every
    wavefront tile is full."""
    n = 22
    rng = np.random.default_rng(7)
    c = Circuit(n)
    for q in range(n):
        c.h(q)
    c.gate(Gate.unitary_kq(random_unitary(64, rng)), [3, 17, 0, 9, 21, 12])
    c.gate(Gate.unitary_kq(random_unitary(16, rng)), [20, 1, 5, 14])
    psi = W.gaussian_state(n, 3)
    o = oracle_lib.OracleSv()
    ost = o.allocate(n)
    ost.load(psi)
    o.run(c, ost)
    st = hip.allocate(n)
    st.load(psi)
    plan = hip.compile(c)
    plan.execute(st)
    assert max_abs(st.amplitudes(), ost.amplitudes()) < PARITY_TOL
    plan.free()
    st.free()
