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
    instrs = P.cpu_pipeline(c, max_k=max_k, tile_bits=tile_bits)
    P.apply(o, b, instrs, tile_bits)
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
    instrs = P.cpu_pipeline(c)
    kq = [i for i in instrs if isinstance(i, GateInstance) and i.gate.name == "UnitaryKq"]
    tiled = [i for i in instrs if isinstance(i, P.Tiled)]
    assert kq and all(3 <= len(i.qubits) <= 4 for i in kq)
    assert tiled and all(all(q < 15 for g in t.gates for q in g.qubits) for t in tiled)
    assert len(instrs) < len(c.gates()) / 4
