"""Mid-circuit measurement through the GPU paths: the surface-code cycle
fixture (a Z-parity and an X-parity stabilizer read out through ancillas)
gate by gate (run_with_outcomes, crates/aleph-backend/src/lib.rs:211-248)
and as planned measurement-free segments (run_circuit_with_outcomes), with
the measurement uniforms forced so the CPU oracle follows the same branch."""

import math

import numpy as np
import pytest

from aleph_amd import backend as B
from aleph_amd.circuit import Circuit
from tests.helpers import PARITY_TOL, fixture_amplitudes, fixture_circuit, load_fixture, max_abs

pytestmark = pytest.mark.gpu


def oracle_with_outcomes(orc, circuit, uniforms):
    o = orc.OracleSv()
    st = o.allocate(circuit.num_qubits)
    draws = list(uniforms)
    out = []
    for idx, inst in enumerate(circuit.instructions):
        if isinstance(inst, tuple) and inst[0] == "measure":
            bit, used = o.measure_with_uniform(st, inst[1], draws[0])
            if used:
                draws.pop(0)
            out.append((idx, inst[1], inst[2], bit))
        elif type(inst).__name__ == "GateInstance":
            o.apply_gate(st, inst)
    return st.amplitudes(), out


def gpu_gatewise_with_uniforms(hip, circuit, uniforms):
    st = hip.allocate(circuit.num_qubits)
    draws = list(uniforms)
    out = []
    for idx, inst in enumerate(circuit.instructions):
        if isinstance(inst, tuple) and inst[0] == "measure":
            bit, used = hip.measure_with_uniform(st, inst[1], draws[0])
            if used:
                draws.pop(0)
            out.append((idx, inst[1], inst[2], bit))
        else:
            hip.apply_gate(st, inst)
    return st, out


@pytest.mark.parametrize("u", [0.1, 0.4, 0.6, 0.95])
def test_surface_code_cycle_matches_oracle(hip, oracle_lib, u):
    c = fixture_circuit("surface_code_cycle")
    want, want_rec = oracle_with_outcomes(oracle_lib, c, [u, u, u])
    st, rec = gpu_gatewise_with_uniforms(hip, c, [u, u, u])
    assert rec == want_rec
    assert max_abs(st.amplitudes(), want) < PARITY_TOL
    st.free()
    st2, rec2 = B.run_circuit_with_outcomes(hip, c, uniforms=[u, u, u])
    assert [(r.instruction_index, r.qubit, r.clbit, r.outcome) for r in rec2] == want_rec
    assert max_abs(st2.amplitudes(), want) < PARITY_TOL
    st2.free()
    # Z parity of |0000> is even; the X-parity outcome leaves the data in
    # (|0000> + (-1)^c1 |1111>)/sqrt(2) with the ancilla in |c1>.
    c0, c1 = rec[0][3], rec[1][3]
    assert c0 is False or c0 == 0
    expect = np.zeros(64, dtype=np.complex128)
    base = 32 if c1 else 0
    expect[base] = 1 / math.sqrt(2)
    expect[base + 15] = (-1 if c1 else 1) / math.sqrt(2)
    assert max_abs(want, expect) < 1e-12


def test_surface_code_cycle_outcome_zero_state_is_the_fixture_state(hip):
    c = fixture_circuit("surface_code_cycle")
    st, rec = B.run_circuit_with_outcomes(hip, c, uniforms=[0.75])   # u >= P(1) = 1/2: X-parity outcome 0
    assert [r.outcome for r in rec] == [False, False]
    assert max_abs(st.amplitudes(), fixture_amplitudes(load_fixture("surface_code_cycle"))) < 1e-10
    st.free()


def test_surface_code_cycle_statistics(hip):
    """The device RNG path: the X-parity bit is a fair coin, the Z-parity
    bit never fires (Aer counts: four outcomes at 1/4 each)."""
    c = fixture_circuit("surface_code_cycle")
    ones = 0
    runs = 400
    for _ in range(runs):
        st, rec = B.run_circuit_with_outcomes(hip, c)
        assert rec[0].outcome is False or rec[0].outcome == 0
        ones += int(rec[1].outcome)
        st.free()
    assert abs(ones / runs - 0.5) <= 5 * math.sqrt(0.25 / runs)


def test_segments_with_measurements_match_gatewise(hip, oracle_lib):
    rng = np.random.default_rng(3)
    c = Circuit(8)
    for layer in range(6):
        for q in range(8):
            c.u3(*rng.uniform(0, 2 * np.pi, 3), q)
        for q in range(layer % 2, 7, 2):
            c.cx(q, q + 1)
        c.add(("measure", int(rng.integers(0, 8)), layer))
    us = rng.uniform(size=8)
    want, want_rec = oracle_with_outcomes(oracle_lib, c, us)
    st, rec = B.run_circuit_with_outcomes(hip, c, uniforms=us, jit=True)
    assert [(r.instruction_index, r.qubit, r.clbit, r.outcome) for r in rec] == want_rec
    assert max_abs(st.amplitudes(), want) < PARITY_TOL
    st.free()
