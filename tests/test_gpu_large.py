"""States of 2^32 amplitudes (FP32, 32 GiB): every launch stays under the
2^32 work-item limit of a dispatch and loops over the rest.  Checked with
size-independent properties: norm, Pauli expectations of product states,
a marginal, and the same values through the fused plan, the per-gate
kernels (1q, swap, diagonal) and the readout kernels."""

import math

import numpy as np
import pytest

from aleph_amd.backend import HipSvBackend
from aleph_amd.circuit import Circuit, Gate, GateInstance, PauliString

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    b = HipSvBackend(seed=3, precision=32)
    st = b.allocate(32)
    yield b, st
    st.free()


def test_fused_plan_and_readout_at_n32(big):
    b, st = big
    c = Circuit(32)
    for q in range(32):
        c.h(q)
    c.ry(0.7, 31)
    plan = b.compile(c)
    plan.execute(st)
    plan.free()
    assert abs(b.norm_sqr(st) - 1.0) < 1e-5
    for q in (0, 17, 30):
        assert abs(b.expectation_value(st, PauliString.new(1.0, [(q, "X")])) - 1.0) < 1e-4
    # Ry(0.7) H|0> on qubit 31: <X> = cos(0.7)
    assert abs(b.expectation_value(st, PauliString.new(1.0, [(31, "X")])) - math.cos(0.7)) < 1e-4
    p = b.probabilities(st, [31])          # Ry(t)|+> = ((c - s)|0> + (c + s)|1>)/sqrt(2)
    assert abs(p[0] - (1 - math.sin(0.7)) / 2) < 1e-4


def test_per_gate_kernels_at_n32(big):
    b, st = big
    b.apply_gate(st, GateInstance(Gate("Ry", (-0.7,)), [31]))   # undo the Ry: qubit 31 back to |+>
    for q in (0, 31):
        b.apply_gate(st, GateInstance(Gate("H"), [q]))           # k_1q over 2^31 pairs
    assert abs(b.expectation_value(st, PauliString.new(1.0, [(31, "Z")])) - 1.0) < 1e-4
    b.apply_gate(st, GateInstance(Gate("X"), [0]))
    b.apply_gate(st, GateInstance(Gate("Swap"), [0, 31]))        # k_2q / swap over 2^30 quads
    assert abs(b.expectation_value(st, PauliString.new(1.0, [(31, "Z")])) + 1.0) < 1e-4
    assert abs(b.expectation_value(st, PauliString.new(1.0, [(0, "Z")])) - 1.0) < 1e-4
    b.apply_gate(st, GateInstance(Gate("Rz", (1.3,)), [5]))      # diagonal over 2^32 amplitudes
    assert abs(b.expectation_value(st, PauliString.new(1.0, [(5, "X")])) - math.cos(1.3)) < 1e-4
    assert abs(b.norm_sqr(st) - 1.0) < 1e-5
