"""Noise trajectories, CPU side: the oracle against the reference's Aer
fixtures, the trajectory restatement against the exact density matrix, the
random stream against splitmix64's known outputs, and the host layer's
channel constructors, attachment order and validation (the checks
crates/aleph-sv/src/noise/{error,model,execution}.rs make before any
trajectory runs)."""

import numpy as np
import pytest

from aleph_amd import noise as N
from aleph_amd.circuit import Circuit
from oracle import noise_oracle as NO
from tests.helpers import (NOISE_FIXTURES, assert_distribution_close, load_noise_fixture, mixed_noise_case,
                           noise_fixture_case)


def test_splitmix64_known_outputs():
    # First three outputs of splitmix64 seeded with 0 (Vigna's reference generator).
    assert [NO.mix64(k * NO.GOLDEN) for k in (1, 2, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                             0x06C45D188009454F]
    key = 12345
    assert NO.uniform(key, 0) == (NO.mix64(key + NO.GOLDEN) >> 11) / 2.0 ** 53


def test_shot_seed_is_deterministic_and_distinct():
    # apply.rs:187-192
    assert NO.shot_seed(7, 3) == NO.shot_seed(7, 3)
    assert NO.shot_seed(7, 3) != NO.shot_seed(7, 4)
    assert NO.shot_seed(7, 3) != NO.shot_seed(8, 3)
    u = [NO.uniform(NO.shot_seed(1, s), c) for s in range(200) for c in range(5)]
    assert 0.0 <= min(u) and max(u) < 1.0 and abs(np.mean(u) - 0.5) < 0.05


@pytest.mark.parametrize("name", NOISE_FIXTURES)
def test_density_matrix_oracle_matches_aer_fixture(oracle_lib, name):
    fx = load_noise_fixture(name)
    c, m = noise_fixture_case(name)
    prog, att = N.resolve(c, m)
    got = NO.exact_distribution(prog, att, m.readout_map(), c.num_qubits)
    assert fx["num_qubits"] == c.num_qubits
    assert np.max(np.abs(got - np.asarray(fx["exact_probs"]))) < 1e-12


def test_density_matrix_oracle_catches_a_wrong_weight(oracle_lib):
    # Moving 0.02 of the two-qubit depolarizing weight from II to IX must
    # miss the depol_cx fixture (a dropped or misplaced term shows).
    c, _ = noise_fixture_case("depol_cx")
    w = N.depolarizing_error(0.1, 2)
    terms = list(w.terms)
    terms[0] = (terms[0][0] - 0.02, terms[0][1])
    terms[1] = (terms[1][0] + 0.02, terms[1][1])
    assert terms[1][1] == ("I", "X")
    bad = N.NoiseModel()
    bad.add_quantum_error(N.PauliChannel(2, tuple(terms)), ["Cnot"], [0, 1])
    prog, att = N.resolve(c, bad)
    got = NO.exact_distribution(prog, att, {}, 2)
    assert np.max(np.abs(got - np.asarray(load_noise_fixture("depol_cx")["exact_probs"]))) > 1e-2


@pytest.mark.parametrize("name", ["depol_cx", "amp_damp_h", "combined_ghz3"])
def test_trajectory_oracle_matches_exact_distribution(oracle_lib, name):
    c, m = noise_fixture_case(name)
    prog, att = N.resolve(c, m)
    shots = 3000
    out = NO.trajectory_outcomes(prog, att, m.readout_map(), c.num_qubits, 11, shots)
    counts = np.bincount(out.astype(np.int64), minlength=1 << c.num_qubits)
    assert_distribution_close(counts, np.asarray(load_noise_fixture(name)["exact_probs"]), shots, name)


def test_trajectory_oracle_matches_exact_on_a_mixed_model(oracle_lib):
    c, m = mixed_noise_case(4, 3, seed=5)
    prog, att = N.resolve(c, m)
    exact = NO.exact_distribution(prog, att, m.readout_map(), 4)
    assert abs(exact.sum() - 1.0) < 1e-12
    shots = 2000
    out = NO.trajectory_outcomes(prog, att, m.readout_map(), 4, 3, shots)
    assert_distribution_close(np.bincount(out.astype(np.int64), minlength=16), exact, shots, "mixed")


def test_trajectory_special_cases(oracle_lib):
    # apply.rs:206-330: certain X flips |0>, identity is a no-op, amplitude
    # damping fixes |0>, certain readout flips every bit.
    c = Circuit(2).h(0).h(0)
    m = N.NoiseModel()
    m.add_all_qubit_quantum_error(N.bit_flip_error(1.0), ["H"])
    prog, att = N.resolve(c, m)
    assert set(NO.trajectory_outcomes(prog, att, {}, 2, 1, 20)) == {1}     # X H X H |0> = |1>
    m = N.NoiseModel()
    m.add_quantum_error(N.amplitude_damping_error(0.7), ["H"], [1])
    prog, att = N.resolve(Circuit(2).h(1).h(1), m)
    assert set(NO.trajectory_outcomes(prog, att, {}, 2, 2, 30)) <= {0, 2}
    m = N.NoiseModel()
    for q in range(2):
        m.add_readout_error(N.ReadoutError([[0.0, 1.0], [1.0, 0.0]]), q)
    prog, att = N.resolve(Circuit(2).x(0), m)
    assert set(NO.trajectory_outcomes(prog, att, m.readout_map(), 2, 3, 10)) == {0b10}


def test_constructors_follow_aer_weights():
    d1 = N.depolarizing_error(0.2, 1)
    assert d1.terms[0] == (1 - 0.15, ("I",)) and all(w == 0.05 for w, _ in d1.terms[1:])
    d2 = N.depolarizing_error(0.16, 2)
    assert len(d2.terms) == 16 and d2.terms[0] == (1 - 0.15, ("I", "I"))
    assert all(abs(w - 0.01) < 1e-15 for w, _ in d2.terms[1:])
    assert abs(sum(w for w, _ in d2.terms) - 1.0) < 1e-15
    pe = N.pauli_error([("X", 1.0), ("Z", 3.0)])
    assert pe.terms == ((0.25, ("X",)), (0.75, ("Z",)))
    for err in (N.amplitude_damping_error(0.3), N.phase_damping_error(0.6)):
        comp = sum(np.asarray(k).conj().T @ np.asarray(k) for k in err.kraus)
        assert np.allclose(comp, np.eye(2), atol=1e-15)
    assert N.bit_flip_error(0.1).terms == ((0.9, ("I",)), (0.1, ("X",)))
    assert N.phase_flip_error(0.1).terms == ((0.9, ("I",)), (0.1, ("Z",)))
    with pytest.raises(ValueError):
        N.depolarizing_error(1.5, 1)
    with pytest.raises(ValueError):
        N.depolarizing_error(0.1, 3)
    with pytest.raises(ValueError):
        N.pauli_error([("XX", 1.0)])
    with pytest.raises(ValueError):
        N.ReadoutError([[1.0, 0.0]])


def test_model_attachment_order_and_names():
    # model.rs:100-122: all-qubit errors first, then the qubit-specific ones.
    m = N.NoiseModel()
    a, b, c = N.bit_flip_error(0.1), N.phase_flip_error(0.2), N.depolarizing_error(0.3, 1)
    m.add_quantum_error(a, ["H"], [0])
    m.add_all_qubit_quantum_error(b, "H")
    m.add_all_qubit_quantum_error(c, ["H", "X"])
    assert m.errors_for("H", [0]) == [b, c, a]
    assert m.errors_for("H", [1]) == [b, c]
    assert m.errors_for("X", [0]) == [c] and m.errors_for("h", [0]) == []
    ro = N.ReadoutError([[0.9, 0.1], [0.2, 0.8]])
    m.add_readout_error(ro, 2)
    assert m.readout_error(2) == ro and m.readout_map() == {2: ro}
    assert N.aer_gate_names(["h", "CX", "ccx", "p"]) == ["H", "Cnot", "Toffoli", "Phase"]
    with pytest.raises(ValueError):
        N.aer_gate_names(["id"])


def test_resolve_validates_before_running():
    c = Circuit(2).h(0).cx(0, 1)
    m = N.NoiseModel()
    m.add_all_qubit_quantum_error(N.depolarizing_error(0.1, 1), ["Cnot"])    # arity 1 on a 2q gate
    with pytest.raises(N.InvalidChannel, match="arity"):
        N.resolve(c, m)
    m = N.NoiseModel()
    m.add_all_qubit_quantum_error(N.PauliChannel(1, ((0.5, ("I",)), (0.4, ("X",)))), ["H"])
    with pytest.raises(N.InvalidChannel, match="sum to one"):
        N.resolve(c, m)
    m = N.NoiseModel()
    m.add_all_qubit_quantum_error(N.KrausChannel((np.eye(2), np.eye(2))), ["H"])
    with pytest.raises(N.InvalidChannel, match="trace preserving"):
        N.resolve(c, m)
    m = N.NoiseModel()
    m.add_readout_error(N.ReadoutError([[0.9, 0.2], [0.0, 1.0]]), 0)
    with pytest.raises(N.InvalidChannel, match="readout"):
        N.resolve(c, m)
    m = N.NoiseModel()
    m.add_readout_error(N.ReadoutError([[1.0, 0.0], [0.0, 1.0]]), 5)
    with pytest.raises(N.InvalidChannel, match="readout"):
        N.resolve(c, m)
    with pytest.raises(N.MidCircuit):
        N.resolve(Circuit(1).h(0).add(("measure", 0, 0)).h(0), N.NoiseModel())
    prog, att = N.resolve(c, N.NoiseModel())
    assert len(prog) == 2 and att == []


def test_buffer_plan_checks_the_budget():
    # A dense 2^20 histogram exceeds 1 MiB; sparse counts of 10 shots do not.
    with pytest.raises(N.MemoryBudget):
        N._plan(20, 10, N.NoiseOptions(memory_budget_bytes=1 << 20), sparse=False)
    with pytest.raises(N.MemoryBudget):     # ten 16 MiB device slices of a 20-qubit state
        N._plan(20, 10, N.NoiseOptions(memory_budget_bytes=1 << 20), sparse=True)
    assert N._plan(12, 10, N.NoiseOptions(memory_budget_bytes=1 << 20), sparse=True) == 0
    assert N._plan(16, 10, N.NoiseOptions(memory_budget_bytes=1 << 30), sparse=True) == 10 * (1 << 16) * 16
    with pytest.raises(N.Configuration):
        N._plan(4, 10, N.NoiseOptions(memory_budget_bytes=0), sparse=False)
