"""The CPU pipeline restatement (oracle/pipeline.py) against the oracle's
gate-by-gate run: the reference's own acceptance criterion for its fusion
passes (fused == unfused, crates/aleph-ir/tests/fuse_kq_oracle.rs)."""

import numpy as np
import pytest

from aleph_amd import workloads as W
from aleph_amd.circuit import GateInstance
from oracle import pipeline as P
from tests.helpers import max_abs
from tests.test_oracle import random_circuit


def run_both(orc, c, tile_bits, max_k=4):
    psi = W.gaussian_state(c.num_qubits, 3)
    o = orc.OracleSv()
    a = o.allocate(c.num_qubits)
    a.load(psi)
    o.run(c, a)
    b = o.allocate(c.num_qubits)
    b.load(psi)
    instrs, perm = P.cpu_pipeline(c, max_k=max_k, tile_bits=tile_bits)
    P.apply(o, b, instrs, tile_bits, perm)
    return a.amplitudes(), b.amplitudes(), instrs


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("tile_bits", [3, 5])
def test_pipeline_matches_gatewise_random(oracle_lib, seed, tile_bits):
    rng = np.random.default_rng(seed)
    c = random_circuit(int(rng.integers(5, 10)), 80, rng, kq=False)
    a, b, _ = run_both(oracle_lib, c, tile_bits)
    assert max_abs(a, b) < 1e-12


@pytest.mark.parametrize("name", ["brickwall", "qft", "hea", "qaoa"])
def test_pipeline_matches_gatewise_workloads(oracle_lib, name):
    n = 10
    c = {"brickwall": lambda: W.brickwall(n, 12, seed=5), "qft": lambda: W.qft(n),
         "hea": lambda: W.hea(n, 3, seed=1),
         "qaoa": lambda: W.qaoa_maxcut(n, W.random_3_regular(n, 2), [0.3, 0.9], [1.1, 0.4])}[name]()
    a, b, instrs = run_both(oracle_lib, c, 6)
    assert max_abs(a, b) < 1e-12
    assert len(instrs) < len(c.gates())


def test_pipeline_structure_on_brickwall():
    c = W.brickwall(20, 12, seed=1)
    instrs, perm = P.cpu_pipeline(c)
    assert perm is None          # every qubit carries the same traffic: no net win
    kq = [i for i in instrs if isinstance(i, GateInstance) and i.gate.name == "UnitaryKq"]
    tiled = [i for i in instrs if isinstance(i, P.Tiled)]
    assert kq and all(3 <= len(i.qubits) <= 4 for i in kq)
    assert tiled and all(all(q < 15 for g in t.gates for q in g.qubits) for t in tiled)
    assert len(instrs) < len(c.gates()) / 4


# ---------------- CancelInversePairs (crates/aleph-ir/src/passes/cancel.rs tests)

def _names(instrs):
    return [(i.gate.name, tuple(i.qubits)) for i in instrs]


@pytest.mark.parametrize("build,left", [
    (lambda c: c.h(0).h(0), []),
    (lambda c: c.s(0).sdg(0), []),
    (lambda c: c.tdg(0).t(0), []),
    (lambda c: c.rz(0.3, 0).rz(-0.3, 0), []),
    (lambda c: c.rz(0.3, 0).rz(0.3, 0), [("Rz", (0,)), ("Rz", (0,))]),
    (lambda c: c.rz(0.3, 0).rz(-0.3 + 1e-12, 0), [("Rz", (0,)), ("Rz", (0,))]),
    (lambda c: c.cx(0, 1).cx(0, 1), []),
    (lambda c: c.cx(0, 1).cx(1, 0), [("Cnot", (0, 1)), ("Cnot", (1, 0))]),
    (lambda c: c.h(0).x(1).h(0), [("X", (1,))]),            # X(1) is off the H pair's support
    (lambda c: c.h(0).cx(0, 1).h(0), [("H", (0,)), ("Cnot", (0, 1)), ("H", (0,))]),
    (lambda c: c.h(0).h(1).h(1).h(0), []),                  # nested pairs
    (lambda c: c.u3(0.1, 0.2, 0.3, 2).u3(-0.1, -0.3, -0.2, 2), []),
    (lambda c: c.h(0).h(0).h(0), [("H", (0,))]),
])
def test_cancel_inverse_pairs_cases(build, left):
    from aleph_amd.circuit import Circuit
    c = Circuit(3)
    build(c)
    assert _names(P.cancel_inverse_pairs(c.gates())) == left


def test_cancel_does_not_cross_a_barrier():
    from aleph_amd.circuit import Circuit
    c = Circuit(2).h(0).barrier([0]).h(0)
    out = P.cancel_inverse_pairs(list(c.instructions))
    assert len(out) == 3


# ---------------- RelabelQubits (crates/aleph-ir/src/passes/relabel.rs)

def test_relabel_moves_busy_high_qubits_low_and_unpermutes(oracle_lib):
    """Traffic on qubits 12..15 and little on 0..3 with a 4-bit tile: the
    relabelling is a net win, so it is committed; the final state after the
    un-permutation equals the gate-by-gate state."""
    from aleph_amd.circuit import Circuit
    rng = np.random.default_rng(9)
    c = Circuit(16)
    for _ in range(6):
        for q in range(12, 16):
            c.u3(*rng.uniform(0, 6, 3), q)
        c.cx(12, 13).cz(14, 15).cx(13, 14)
    c.h(0).cx(0, 12)
    instrs, perm = P.cpu_pipeline(c, tile_bits=4)
    assert perm is not None and sorted(perm) == list(range(16))
    assert all(perm[q] < 4 for q in (12, 13, 14, 15))
    a, b, _ = run_both(oracle_lib, c, 4)
    assert max_abs(a, b) < 1e-12


def test_relabel_keeps_identity_without_net_win():
    c = W.qft(12)
    _, perm = P.relabel_qubits(c.gates(), 12, 15)
    assert perm is None


@pytest.mark.parametrize("seed", range(4))
def test_pipeline_with_inverse_pairs_matches_gatewise(oracle_lib, seed):
    """A random circuit followed by its inverse collapses under cancellation;
    the pipeline output still equals the gate-by-gate run."""
    from tests.test_gpu_parity import mirror
    rng = np.random.default_rng(50 + seed)
    c = mirror(W.brickwall(int(rng.integers(5, 9)), 3, seed=seed))
    a, b, instrs = run_both(oracle_lib, c, 3)
    assert max_abs(a, b) < 1e-12
    assert len(instrs) < len(c.gates()) / 2
