"""Shared test helpers: fixture loading, brute-force operators, statistics."""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"

STATE_TOLERANCE = 1e-10        # crates/aleph-oracle/src/harness.rs:18
FP32_STATE_TOLERANCE = 1e-5    # harness.rs:27
DISTRIBUTION_SHOTS = 100_000   # harness.rs:136
DISTRIBUTION_FLOOR = 1e-6      # harness.rs:144
PARITY_TOL = 1e-12             # GPU FP64 amplitudes vs the CPU FP64 oracle


def load_fixture(name: str) -> dict:
    return json.loads((GOLDEN / "aer" / f"{name}.json").read_text())


def fixture_names() -> list[str]:
    return sorted(p.stem for p in (GOLDEN / "aer").glob("*.json"))


def fixture_amplitudes(fx: dict) -> np.ndarray:
    a = np.asarray(fx["statevector"]["amplitudes"], dtype=np.float64)
    return a[:, 0] + 1j * a[:, 1]


def fixture_circuit(name: str):
    from aleph_amd.workloads import load_qasm
    return load_qasm(GOLDEN / "circuits" / f"{name}.qasm")


def has_measure(circuit) -> bool:
    return any(isinstance(i, tuple) and i[0] == "measure" for i in circuit.instructions)


def embed(matrix: np.ndarray, qubits, controls, n: int) -> np.ndarray:
    """Full 2^n x 2^n operator of a gate by enumerating basis states:
    column x maps to sum_s' M[s', s(x)] |x with target bits replaced by s'>,
    where s(x) reads qubits[0] as the most significant bit; identity where a
    control bit is 0."""
    dim = 1 << n
    k = len(qubits)
    U = np.zeros((dim, dim), dtype=np.complex128)
    cmask = sum(1 << c for c in controls)
    tmask = sum(1 << q for q in qubits)
    for x in range(dim):
        if (x & cmask) != cmask:
            U[x, x] = 1.0
            continue
        s = 0
        for j, q in enumerate(qubits):
            s |= ((x >> q) & 1) << (k - 1 - j)
        base = x & ~tmask
        for s2 in range(1 << k):
            y = base
            for j, q in enumerate(qubits):
                if (s2 >> (k - 1 - j)) & 1:
                    y |= 1 << q
            U[y, x] += matrix[s2, s]
    return U


def dense_unitary(circuit, gate_matrix) -> np.ndarray:
    """Product of the embedded operators of every gate of a circuit."""
    from aleph_amd.circuit import GateInstance
    n = circuit.num_qubits
    U = np.eye(1 << n, dtype=np.complex128)
    for inst in circuit.instructions:
        if isinstance(inst, GateInstance):
            U = embed(gate_matrix(inst.gate), inst.qubits, inst.controls, n) @ U
    return U


def assert_distribution_close(empirical_counts: np.ndarray, exact: np.ndarray, shots: int, name: str = "") -> None:
    """Per-outcome 5 sigma Bernoulli band + 1e-6 floor (harness.rs:225-270)."""
    emp = empirical_counts / shots
    for i, (e, p) in enumerate(zip(emp, exact)):
        sigma = math.sqrt(max(p * (1 - p), 0.0) / shots)
        band = 5 * sigma + DISTRIBUTION_FLOOR
        assert abs(e - p) <= band, f"{name}: outcome {i} empirical {e} exact {p} band {band}"


def random_unitary(dim: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))


def max_abs(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) if np.size(a) else 0.0


def qaoa_p1_edge_expectation(gamma: float, beta: float, d_u: int, d_v: int, triangles: int) -> float:
    """<(1 - Z_u Z_v)/2> after one QAOA layer on MaxCut, closed form of Wang,
    Hadfield, Jiang & Rieffel, PRA 97, 022304 (2018), Theorem 1, for the
    ansatz of aleph_amd.workloads.qaoa_maxcut (RZZ(gamma) = exp(-i gamma Z Z / 2),
    RX(beta) = exp(-i beta X / 2)).  In the paper's convention (e^{-i g C},
    e^{-i b B}) this circuit is g = gamma, b = -beta / 2 up to a global phase;
    d_u, d_v count the other neighbours of u and v, `triangles` the common ones."""
    g, b = gamma, -0.5 * beta
    return (0.5 + 0.25 * math.sin(4 * b) * math.sin(g) * (math.cos(g) ** d_u + math.cos(g) ** d_v)
            - 0.25 * math.sin(2 * b) ** 2 * math.cos(g) ** (d_u + d_v - 2 * triangles)
            * (1 - math.cos(2 * g) ** triangles))


def qaoa_p1_cut_expectation(n: int, edges, gamma: float, beta: float) -> float:
    nb = [set() for _ in range(n)]
    for u, v in edges:
        nb[u].add(v)
        nb[v].add(u)
    return sum(qaoa_p1_edge_expectation(gamma, beta, len(nb[u]) - 1, len(nb[v]) - 1, len(nb[u] & nb[v]))
               for u, v in edges)


# ---------------------------------------------------------------- noise

def load_noise_fixture(name: str) -> dict:
    """Exact Aer density-matrix distributions (reference oracle/noise/*.json)."""
    return json.loads((GOLDEN / "noise" / f"{name}.json").read_text())


def noise_fixture_case(name: str):
    """(circuit, model) of one Aer noise fixture, as the reference's
    crates/aleph-sv/tests/noise_oracle.rs:84-146 builds them: the damping
    channels ride on an identity Unitary1q carrier after H."""
    from aleph_amd import noise as N
    from aleph_amd.circuit import Circuit, Gate
    m = N.NoiseModel()
    if name == "depol_h":
        c = Circuit(1).h(0)
        m.add_all_qubit_quantum_error(N.depolarizing_error(0.05, 1), ["H"])
    elif name == "depol_cx":
        c = Circuit(2).h(0).cx(0, 1)
        m.add_quantum_error(N.depolarizing_error(0.1, 2), ["Cnot"], [0, 1])
    elif name in ("amp_damp_h", "phase_damp_h"):
        c = Circuit(1).h(0).gate(Gate.unitary(np.eye(2)), [0])
        err = N.amplitude_damping_error(0.2) if name == "amp_damp_h" else N.phase_damping_error(0.3)
        m.add_quantum_error(err, ["Unitary1q"], [0])
    elif name == "readout_x":
        c = Circuit(1).x(0)
        m.add_readout_error(N.ReadoutError([[0.98, 0.02], [0.05, 0.95]]), 0)
    elif name == "combined_ghz3":
        c = Circuit(3).h(0).cx(0, 1).cx(1, 2)
        m.add_all_qubit_quantum_error(N.depolarizing_error(0.02, 2), ["Cnot"])
        for q in range(3):
            m.add_readout_error(N.ReadoutError([[0.97, 0.03], [0.04, 0.96]]), q)
    else:
        raise KeyError(name)
    return c, m


NOISE_FIXTURES = ["depol_h", "depol_cx", "amp_damp_h", "phase_damp_h", "readout_x", "combined_ghz3"]


def mixed_noise_case(n: int, depth: int, seed: int):
    """A random circuit with every channel kind attached: depolarizing on
    1q and 2q gates, amplitude and phase damping, bit/phase flips, a
    weighted Pauli mixture, readout errors on some qubits."""
    from aleph_amd import noise as N
    from aleph_amd.circuit import Circuit, DiagonalPhase, Gate, PhaseTerm
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for layer in range(depth):
        for q in range(n):
            c.u3(*rng.uniform(0, 2 * np.pi, 3), q)
        for q in range(layer % 2, n - 1, 2):
            c.cx(q, q + 1)
        if layer == depth // 2:
            c.ccx(0, 1, 2)
            c.add(DiagonalPhase(n, [PhaseTerm((0b11,), 0.7), PhaseTerm((0b101, 0b110), -1.1)]))
            c.gate(Gate.Rz(0.3), [n - 1], [0])
    m = N.NoiseModel()
    m.add_all_qubit_quantum_error(N.depolarizing_error(0.08, 2), ["Cnot"])
    m.add_all_qubit_quantum_error(N.depolarizing_error(0.03, 1), ["U3"])
    m.add_quantum_error(N.amplitude_damping_error(0.15), ["U3"], [1])
    m.add_quantum_error(N.phase_damping_error(0.25), ["U3"], [2])
    m.add_quantum_error(N.bit_flip_error(0.1), ["U3"], [0])
    m.add_quantum_error(N.phase_flip_error(0.2), ["U3"], [n - 1])
    m.add_quantum_error(N.pauli_error([("X", 1.0), ("Y", 2.0), ("I", 5.0)]), ["Rz"], [n - 1])
    m.add_readout_error(N.ReadoutError([[0.9, 0.1], [0.2, 0.8]]), 0)
    m.add_readout_error(N.ReadoutError([[0.97, 0.03], [0.06, 0.94]]), n - 1)
    return c, m
