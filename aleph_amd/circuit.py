"""Circuit-side types: Gate, GateInstance, PauliString, DiagonalPhase, Circuit.

Python mirrors of crates/aleph-core/src/gate/kinds.rs (Gate),
gate/instance.rs (GateInstance), pauli.rs (PauliString),
crates/aleph-ir/src/diagonal_phase.rs (DiagonalPhase) and circuit.rs /
instruction.rs (Circuit, Instruction).  Gate matrices are computed by the
native library, not here; a Circuit encodes itself into the flat program
format of `aleph_hip_plan_create` (include/aleph_hip.h).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

# kind codes: order of the reference Gate enum (kinds.rs:13-115)
KINDS = ["H", "X", "Y", "Z", "S", "Sdg", "T", "Tdg", "Rx", "Ry", "Rz", "Phase", "U3",
         "Cnot", "Cz", "Swap", "Iswap", "IswapDg", "CRx", "CRy", "CRz", "Toffoli", "Ccz",
         "Unitary1q", "Unitary2q", "UnitaryKq", "Unitary1qDiag", "Unitary2qDiag"]
KIND = {name: i for i, name in enumerate(KINDS)}
ARITY = {**{k: 1 for k in KINDS[:13]}, **{k: 2 for k in KINDS[13:21]}, "Toffoli": 3, "Ccz": 3,
         "Unitary1q": 1, "Unitary2q": 2, "Unitary1qDiag": 1, "Unitary2qDiag": 2}
N_PARAMS = {"Rx": 1, "Ry": 1, "Rz": 1, "Phase": 1, "CRx": 1, "CRy": 1, "CRz": 1, "U3": 3}
DIAGONAL = {"Z", "S", "Sdg", "T", "Tdg", "Rz", "Phase", "CRz", "Cz", "Ccz", "Unitary1qDiag", "Unitary2qDiag"}


@dataclass(frozen=True)
class Gate:
    """One gate kind with its parameters / owned matrix (Gate, kinds.rs)."""

    name: str
    params: tuple = ()
    matrix: np.ndarray | None = None  # complex; (dim, dim) or (dim,) for the *Diag kinds
    k: int = 0                        # UnitaryKq width

    @property
    def kind(self) -> int:
        return KIND[self.name]

    def arity(self) -> int:
        return self.k if self.name == "UnitaryKq" else ARITY[self.name]

    def is_diagonal(self) -> bool:
        return self.name in DIAGONAL

    # constructors
    @staticmethod
    def unitary(m) -> "Gate":
        m = np.asarray(m, dtype=np.complex128)
        dim = m.shape[0]
        if dim == 2:
            return Gate("Unitary1q", (), m)
        if dim == 4:
            return Gate("Unitary2q", (), m)
        return Gate("UnitaryKq", (), m, k=int(dim).bit_length() - 1)

    @staticmethod
    def unitary_kq(m) -> "Gate":
        m = np.asarray(m, dtype=np.complex128)
        return Gate("UnitaryKq", (), m, k=int(m.shape[0]).bit_length() - 1)

    @staticmethod
    def diag(d) -> "Gate":
        d = np.asarray(d, dtype=np.complex128)
        return Gate("Unitary1qDiag" if d.shape[0] == 2 else "Unitary2qDiag", (), d)

    def payload(self) -> np.ndarray:
        if self.matrix is None:
            return np.zeros(0)
        m = np.ascontiguousarray(self.matrix, dtype=np.complex128)
        return m.reshape(-1).view(np.float64)


def _g(name):
    def make(*params):
        return Gate(name, tuple(float(p) for p in params))
    make.__name__ = name
    return make


for _name in KINDS[:23]:
    setattr(Gate, _name, staticmethod(_g(_name)))


@dataclass
class GateInstance:
    """A gate on concrete qubits (qubits[0] = matrix MSB) plus external controls."""

    gate: Gate
    qubits: tuple
    controls: tuple = ()

    def __init__(self, gate: Gate, qubits: Iterable[int], controls: Iterable[int] = ()):
        self.gate = gate
        self.qubits = tuple(int(q) for q in qubits)
        self.controls = tuple(int(c) for c in controls)

    @staticmethod
    def controlled(gate: Gate, qubits, controls) -> "GateInstance":
        return GateInstance(gate, qubits, controls)


PAULI_CODE = {"I": 0, "X": 1, "Y": 2, "Z": 3}


@dataclass
class PauliString:
    """coefficient * tensor product of Paulis (pauli.rs:34-60)."""

    coefficient: float
    terms: list  # [(qubit, 'X'|'Y'|'Z'|'I')]

    @staticmethod
    def new(coefficient: float, terms) -> "PauliString":
        if not np.isfinite(coefficient):
            raise ValueError("non-finite coefficient")
        terms = sorted(((int(q), p) for q, p in terms), key=lambda t: t[0])
        for a, b in zip(terms, terms[1:]):
            if a[0] == b[0]:
                raise ValueError(f"duplicate qubit {a[0]} in Pauli string")
        return PauliString(float(coefficient), [(q, p) for q, p in terms if p != "I"])

    @staticmethod
    def identity(coefficient: float) -> "PauliString":
        return PauliString(float(coefficient), [])


@dataclass
class PhaseTerm:
    conds: tuple   # parity masks (u64)
    angle: float


@dataclass
class DiagonalPhase:
    """psi[x] *= exp(i * sum_t angle_t [all conds of t have odd parity with x])."""

    n_qubits: int
    terms: list = field(default_factory=list)

    def phase_at(self, x: int) -> float:
        phi = 0.0
        for t in self.terms:
            if all(bin(m & x).count("1") & 1 for m in t.conds):
                phi += t.angle
        return phi


@dataclass
class Barrier:
    qubits: tuple


class Circuit:
    """Ordered instruction list over n qubits (circuit.rs, instruction.rs)."""

    def __init__(self, num_qubits: int):
        self.num_qubits = int(num_qubits)
        self.instructions: list = []

    def __len__(self) -> int:
        return len(self.instructions)

    def add(self, inst) -> "Circuit":
        self.instructions.append(inst)
        return self

    def gate(self, gate: Gate, qubits: Sequence[int], controls: Sequence[int] = ()) -> "Circuit":
        return self.add(GateInstance(gate, qubits, controls))

    # builder shorthands (aleph-py circuit.rs method names)
    def h(self, q): return self.gate(Gate.H(), [q])
    def x(self, q): return self.gate(Gate.X(), [q])
    def y(self, q): return self.gate(Gate.Y(), [q])
    def z(self, q): return self.gate(Gate.Z(), [q])
    def s(self, q): return self.gate(Gate.S(), [q])
    def sdg(self, q): return self.gate(Gate.Sdg(), [q])
    def t(self, q): return self.gate(Gate.T(), [q])
    def tdg(self, q): return self.gate(Gate.Tdg(), [q])
    def rx(self, theta, q): return self.gate(Gate.Rx(theta), [q])
    def ry(self, theta, q): return self.gate(Gate.Ry(theta), [q])
    def rz(self, theta, q): return self.gate(Gate.Rz(theta), [q])
    def p(self, theta, q): return self.gate(Gate.Phase(theta), [q])
    def u3(self, theta, phi, lam, q): return self.gate(Gate.U3(theta, phi, lam), [q])
    def cx(self, c, t): return self.gate(Gate.Cnot(), [c, t])
    cnot = cx
    def cz(self, a, b): return self.gate(Gate.Cz(), [a, b])
    def swap(self, a, b): return self.gate(Gate.Swap(), [a, b])
    def ccx(self, a, b, t): return self.gate(Gate.Toffoli(), [a, b, t])
    def ccz(self, a, b, c): return self.gate(Gate.Ccz(), [a, b, c])

    def cp(self, theta, control, target):
        """Controlled phase as GateInstance::controlled(Phase(theta), [target], [control])."""
        return self.gate(Gate.Phase(theta), [target], [control])

    def barrier(self, qubits): return self.add(Barrier(tuple(int(q) for q in qubits)))

    def gates(self):
        return [i for i in self.instructions if isinstance(i, GateInstance)]

    def encode(self) -> tuple[np.ndarray, np.ndarray]:
        """Flat (words, reals) program of include/aleph_hip.h."""
        words: list = []
        reals: list = []
        for inst in self.instructions:
            if isinstance(inst, GateInstance):
                g = inst.gate
                words += [g.kind, len(inst.qubits), len(inst.controls), g.k, *inst.qubits, *inst.controls]
                reals += list(g.params[:N_PARAMS.get(g.name, 0)])
                if g.matrix is not None:
                    reals += list(g.payload())
            elif isinstance(inst, DiagonalPhase):
                words += [64, len(inst.terms)]
                for t in inst.terms:
                    words.append(len(t.conds))
                    for m in t.conds:
                        words += [int(m) & 0xFFFFFFFF, (int(m) >> 32) & 0xFFFFFFFF]
                    reals.append(float(t.angle))
            elif isinstance(inst, Barrier):
                words += [65, len(inst.qubits), *inst.qubits]
            else:
                raise TypeError(f"cannot encode {inst!r}")
        return np.asarray(words, dtype=np.uint32), np.asarray(reals, dtype=np.float64)
