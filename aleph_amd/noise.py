"""Noisy execution: Monte-Carlo quantum-jump trajectories on the GPU.

Python mirror of crates/aleph-sv/src/noise (error.rs, model.rs,
execution.rs) and of the Aer-name front end of crates/aleph-py/src/noise.rs:
channel constructors with Aer's names and weights, an Aer-style
`NoiseModel` that attaches channels to (gate name, qubits), and
`run_noisy` / `run_noisy_sparse`, which run one trajectory per shot.

The trajectories run in the library (`aleph_hip_run_noisy`): one workgroup
per trajectory, the state in LDS up to 13 qubits.  Each shot draws from a
counter-based stream keyed by the reference's splitmix64 shot seed
(apply.rs:160-180), so a shot's outcome depends only on (seed, shot index).
Channels fire after their gate in attachment order (all-qubit first, then
qubit-specific: model.rs:60-80); a Pauli channel draws one uniform and
applies the chosen Pauli string, a Kraus channel draws one uniform, computes
p_i = |K_i psi|^2 and applies K_i / sqrt(p_i); the terminal outcome is one
draw, then one draw per qubit with a readout error.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .backend import HipSvBackend, _ptr
from .circuit import KINDS, Barrier, Circuit, DiagonalPhase, GateInstance

PAULIS = ("I", "X", "Y", "Z")
_CODE = {p: i for i, p in enumerate(PAULIS)}


class NoiseError(Exception):
    """Errors of the noisy driver (NoiseError, noise/mod.rs:49-72)."""


class MemoryBudget(NoiseError):
    def __init__(self, required: int, budget: int):
        super().__init__(f"noisy buffer plan reserves {required} bytes, budget is {budget}; use sparse counts "
                         f"or increase NoiseOptions.memory_budget_bytes")
        self.required, self.budget = required, budget


class Configuration(NoiseError):
    def __init__(self, reason: str):
        super().__init__(f"invalid noisy execution configuration: {reason}")
        self.reason = reason


class InvalidChannel(NoiseError):
    def __init__(self, reason: str):
        super().__init__(f"invalid noise channel: {reason}")
        self.reason = reason


class MidCircuit(NoiseError):
    def __init__(self, kind: str):
        super().__init__(f"mid-circuit {kind} is not supported under noise (terminal measurement only)")
        self.kind = kind


class Unsupported(NoiseError):
    def __init__(self, kind: str):
        super().__init__(f"instruction {kind} is not supported under noise")
        self.kind = kind


# ---------------------------------------------------------------- channels

@dataclass(frozen=True)
class PauliChannel:
    """Mixture of Pauli strings (error.rs:11-20): terms[i] = (prob, labels),
    labels[j] acting on the channel's local qubit j."""

    arity: int
    terms: tuple

    def num_qubits(self) -> int:
        return self.arity


@dataclass(frozen=True)
class KrausChannel:
    """One-qubit CPTP map by its Kraus operators (error.rs:22-28)."""

    kraus: tuple  # of 2x2 complex arrays

    def num_qubits(self) -> int:
        return 1


QuantumError = PauliChannel | KrausChannel


@dataclass(frozen=True)
class ReadoutError:
    """m[t][o] = P(read o | true t), rows summing to one (error.rs:48-68)."""

    m: tuple

    def __init__(self, m):
        rows = tuple(tuple(float(v) for v in r) for r in m)
        if len(rows) != 2 or any(len(r) != 2 for r in rows):
            raise ValueError("readout error is a 2x2 matrix")
        object.__setattr__(self, "m", rows)


def depolarizing_error(p: float, num_qubits: int) -> PauliChannel:
    """Aer's depolarizing_error (error.rs:77-119): identity weight
    1 - p (d^2 - 1) / d^2, every other Pauli p / d^2, d = 2^num_qubits."""
    if not 0.0 <= p <= 1.0:
        raise ValueError("depolarizing p must be in [0, 1]")
    if num_qubits not in (1, 2):
        raise ValueError("depolarizing is 1q or 2q")
    d2 = 1 << (2 * num_qubits)
    off = p / d2
    if num_qubits == 1:
        terms = tuple((1.0 - p * 3.0 / 4.0 if a == "I" else off, (a,)) for a in PAULIS)
    else:
        terms = tuple((1.0 - p * 15.0 / 16.0 if (a, b) == ("I", "I") else off, (a, b))
                      for a in PAULIS for b in PAULIS)
    return PauliChannel(num_qubits, terms)


def amplitude_damping_error(gamma: float) -> KrausChannel:
    """K0 = diag(1, sqrt(1-gamma)), K1 = sqrt(gamma) |0><1| (error.rs:121-140)."""
    if not 0.0 <= gamma <= 1.0:
        raise ValueError("gamma must be in [0, 1]")
    k0 = np.array([[1.0, 0.0], [0.0, np.sqrt(1.0 - gamma)]], dtype=np.complex128)
    k1 = np.array([[0.0, np.sqrt(gamma)], [0.0, 0.0]], dtype=np.complex128)
    return KrausChannel((k0, k1))


def phase_damping_error(lam: float) -> KrausChannel:
    """K0 = diag(1, sqrt(1-lam)), K1 = diag(0, sqrt(lam)) (error.rs:142-161)."""
    if not 0.0 <= lam <= 1.0:
        raise ValueError("lambda must be in [0, 1]")
    k0 = np.array([[1.0, 0.0], [0.0, np.sqrt(1.0 - lam)]], dtype=np.complex128)
    k1 = np.array([[0.0, 0.0], [0.0, np.sqrt(lam)]], dtype=np.complex128)
    return KrausChannel((k0, k1))


def _flip(label: str, p: float) -> PauliChannel:
    if not 0.0 <= p <= 1.0:
        raise ValueError("flip p must be in [0, 1]")
    return PauliChannel(1, ((1.0 - p, ("I",)), (p, (label,))))


def bit_flip_error(p: float) -> PauliChannel:
    return _flip("X", p)


def phase_flip_error(p: float) -> PauliChannel:
    return _flip("Z", p)


def pauli_error(terms) -> PauliChannel:
    """One-qubit Pauli mixture from (label, weight) pairs, weights
    renormalised to sum to one (error.rs:183-210)."""
    terms = [(str(s), float(w)) for s, w in terms]
    total = sum(w for _, w in terms)
    if not total > 0.0:
        raise ValueError("pauli_error weights must sum to > 0")
    for s, _ in terms:
        if s not in _CODE:
            raise ValueError(f"pauli_error label {s!r} is not a single-qubit Pauli (I/X/Y/Z); "
                             f"use depolarizing_error(p, 2) for 2-qubit noise")
    return PauliChannel(1, tuple((w / total, (s,)) for s, w in terms))


# ---------------------------------------------------------------- model

class NoiseModel:
    """(gate name, qubits) -> channels, Aer-style (model.rs:8-96).  Keys are
    the library's gate names ("H", "Cnot", "Unitary1q", ...); see
    `aer_gate_names` for Aer's mnemonics."""

    def __init__(self):
        self._specific: dict = {}
        self._all: dict = {}
        self._readout: dict = {}

    def add_quantum_error(self, err, gate_names, qubits) -> None:
        key = tuple(int(q) for q in qubits)
        for name in _names(gate_names):
            self._specific.setdefault(name, {}).setdefault(key, []).append(err)

    def add_all_qubit_quantum_error(self, err, gate_names) -> None:
        for name in _names(gate_names):
            self._all.setdefault(name, []).append(err)

    def add_readout_error(self, err: ReadoutError, qubit: int) -> None:
        self._readout[int(qubit)] = err

    def errors_for(self, gate_name: str, qubits) -> list:
        """All-qubit errors first, then the qubit-specific ones (model.rs:60-80)."""
        spec = self._specific.get(gate_name, {}).get(tuple(int(q) for q in qubits), [])
        return list(self._all.get(gate_name, [])) + list(spec)

    def readout_error(self, qubit: int):
        return self._readout.get(int(qubit))

    def readout_map(self) -> dict:
        return dict(self._readout)

    def is_empty(self) -> bool:
        return not (self._specific or self._all or self._readout)


def _names(gate_names):
    return [gate_names] if isinstance(gate_names, str) else list(gate_names)


_AER = {"h": "H", "x": "X", "y": "Y", "z": "Z", "s": "S", "sdg": "Sdg", "t": "T", "tdg": "Tdg", "rx": "Rx",
        "ry": "Ry", "rz": "Rz", "p": "Phase", "phase": "Phase", "u3": "U3", "u": "U3", "cx": "Cnot",
        "cnot": "Cnot", "cz": "Cz", "swap": "Swap", "iswap": "Iswap", "crx": "CRx", "cry": "CRy", "crz": "CRz",
        "ccx": "Toffoli", "toffoli": "Toffoli", "ccz": "Ccz"}


def aer_gate_names(names) -> list:
    """Aer / QASM mnemonics (case-insensitive) to the library's gate names,
    as the Python front end does (crates/aleph-py/src/noise.rs:18-60);
    unknown names (Aer's idle "id" among them) raise ValueError."""
    out = []
    for n in _names(names):
        k = _AER.get(n.lower())
        if k is None:
            raise ValueError(f"unknown gate name {n!r}; supported: {', '.join(sorted(_AER))} (no 'id'/idle gate)")
        out.append(k)
    return out


# ---------------------------------------------------------------- execution

@dataclass
class NoiseOptions:
    """Buffer budget of one noisy call (execution.rs:15-37): host outcome
    and count buffers plus the device slices of states wider than 13 qubits."""

    memory_budget_bytes: int = 1 << 30
    max_workers: int | None = None


def _validate_channel(err, arity: int) -> None:
    """Python side of validate_channel (execution.rs:220-260); the library
    checks the same conditions again."""
    if err.num_qubits() != arity:
        raise InvalidChannel("attachment arity does not match gate")
    if isinstance(err, PauliChannel):
        total = 0.0
        for p, labels in err.terms:
            if not np.isfinite(p) or p < 0.0 or len(labels) != arity:
                raise InvalidChannel("invalid Pauli probability or term arity")
            total += p
        if not np.isfinite(total) or abs(total - 1.0) > 1e-9:
            raise InvalidChannel("Pauli probabilities must sum to one")
    elif isinstance(err, KrausChannel):
        comp = np.zeros((2, 2), dtype=np.complex128)
        for k in err.kraus:
            k = np.asarray(k, dtype=np.complex128)
            if k.shape != (2, 2) or not np.all(np.isfinite(k)):
                raise InvalidChannel("non-finite Kraus matrix")
            comp += k.conj().T @ k
        if np.max(np.abs(comp - np.eye(2))) > 1e-9:
            raise InvalidChannel("Kraus operators must be trace preserving")
    else:
        raise InvalidChannel("unknown channel type")


def resolve(circuit: Circuit, model: NoiseModel):
    """The program circuit and, per channel, (record index, channel, qubits),
    after the checks the reference makes before any trajectory runs
    (execution.rs:261-300)."""
    program = Circuit(circuit.num_qubits)
    attached = []
    for inst in circuit.instructions:
        if isinstance(inst, tuple) and inst and inst[0] in ("measure", "reset"):
            raise MidCircuit(inst[0])
        if isinstance(inst, GateInstance):
            for err in model.errors_for(inst.gate.name, inst.qubits):
                _validate_channel(err, len(inst.qubits))
                attached.append((len(program), err, inst.qubits))
        elif not isinstance(inst, (DiagonalPhase, Barrier)):
            raise Unsupported(type(inst).__name__)
        program.add(inst)
    for q, ro in model.readout_map().items():
        m = np.asarray(ro.m, dtype=np.float64)
        if q >= circuit.num_qubits or not np.all(np.isfinite(m)) or np.any(m < 0.0) \
                or np.any(np.abs(m.sum(axis=1) - 1.0) > 1e-9):
            raise InvalidChannel("invalid readout probability or qubit")
    return program, attached


class _Channels:
    """ctypes array of aleph_hip_channel plus the buffers it points into."""

    def __init__(self, attached):
        self.keep = []
        arr = (L.Channel * max(1, len(attached)))()
        for i, (after, err, qubits) in enumerate(attached):
            ch = arr[i]
            ch.after = after
            ch.n_qubits = len(qubits)
            for j, q in enumerate(qubits[:2]):
                ch.qubits[j] = int(q)
            if isinstance(err, PauliChannel):
                w = np.ascontiguousarray([float(p) for p, _ in err.terms], dtype=np.float64)
                codes = np.ascontiguousarray([_CODE[s] for _, labels in err.terms for s in labels], dtype=np.uint8)
                ch.kind, ch.n_branches = L.CHANNEL_PAULI, len(err.terms)
                ch.weights, ch.paulis = _ptr(w, C.c_double), _ptr(codes, C.c_uint8)
                self.keep += [w, codes]
            else:
                k = np.ascontiguousarray(np.stack([np.asarray(m, dtype=np.complex128) for m in err.kraus]))
                flat = k.reshape(-1).view(np.float64)
                ch.kind, ch.n_branches = L.CHANNEL_KRAUS, len(err.kraus)
                ch.kraus = _ptr(flat, C.c_double)
                self.keep += [k, flat]
        self.array, self.count = arr, len(attached)


def outcomes(circuit: Circuit, model: NoiseModel, shots: int, seed: int, backend: HipSvBackend | None = None,
             shot_offset: int = 0, device_budget: int = 0) -> np.ndarray:
    """Per-shot basis-index outcomes of `shots` trajectories (shot i is
    absolute shot shot_offset + i)."""
    backend = backend or HipSvBackend(seed=0)
    program, attached = resolve(circuit, model)
    n = circuit.num_qubits
    words, reals = program.encode()
    nw, nr = int(words.size), int(reals.size)
    words = np.ascontiguousarray(words if nw else np.zeros(1, np.uint32))
    reals = np.ascontiguousarray(reals if nr else np.zeros(1, np.float64))
    chans = _Channels(attached)
    ro = np.zeros(4 * max(n, 1), dtype=np.float64)
    mask = 0
    for q, err in model.readout_map().items():
        ro[4 * q:4 * q + 4] = np.asarray(err.m, dtype=np.float64).reshape(-1)
        mask |= 1 << q
    out = np.zeros(max(shots, 1), dtype=np.uint64)
    lib = L.load()
    code = lib.aleph_hip_run_noisy(backend._ctx, n, _ptr(words, C.c_uint32), nw, _ptr(reals, C.c_double), nr,
                                   chans.array, chans.count, _ptr(ro, C.c_double), C.c_uint64(mask), shots,
                                   C.c_uint64(seed & (2**64 - 1)), C.c_uint64(shot_offset), device_budget,
                                   _ptr(out, C.c_uint64))
    if code != L.OK:
        err = L.Error()
        lib.aleph_hip_last_error(backend._ctx, C.byref(err))
        kind = err.kind.decode(errors="replace")
        if kind == "invalid-channel":
            raise InvalidChannel(err.message.decode(errors="replace"))
        if kind == "unsupported":
            raise Unsupported(err.message.decode(errors="replace"))
        backend._check(code)
    return out[:shots]


def _plan(n: int, shots: int, options: NoiseOptions, sparse: bool) -> int:
    """Buffer bytes of one call, checked against the budget before anything
    is allocated (plan_noisy, execution.rs:62-190): the outcomes, the
    dense or sparse result and the device slices of wide states."""
    if options.memory_budget_bytes <= 0 or (options.max_workers is not None and options.max_workers <= 0):
        raise Configuration("budget and max_workers must be positive")
    result = min(shots, 1 << n) * 16 if sparse else (1 << n) * 8
    slices = 0
    if n > 13 and shots:
        slices = min(shots, 1024, options.max_workers or 1024) * (1 << n) * 16
    required = shots * 8 + result + slices
    if required > options.memory_budget_bytes:
        raise MemoryBudget(required, options.memory_budget_bytes)
    return slices


def run_noisy(circuit: Circuit, model: NoiseModel, shots: int, seed: int, options: NoiseOptions | None = None,
              backend: HipSvBackend | None = None) -> np.ndarray:
    """Dense histogram of length 2^n (run_noisy_with_options, execution.rs:360-376)."""
    options = options or NoiseOptions()
    slices = _plan(circuit.num_qubits, shots, options, sparse=False)
    hist = np.zeros(1 << circuit.num_qubits, dtype=np.uint64)
    if shots:
        idx = outcomes(circuit, model, shots, seed, backend, device_budget=slices)
        np.add.at(hist, idx.astype(np.int64), 1)
    return hist


def run_noisy_sparse(circuit: Circuit, model: NoiseModel, shots: int, seed: int,
                     options: NoiseOptions | None = None, backend: HipSvBackend | None = None) -> list:
    """Ascending (basis index, count) pairs of the observed outcomes
    (run_noisy_sparse, execution.rs:378-400)."""
    options = options or NoiseOptions()
    slices = _plan(circuit.num_qubits, shots, options, sparse=True)
    if not shots:
        return []
    idx = outcomes(circuit, model, shots, seed, backend, device_budget=slices)
    vals, counts = np.unique(idx, return_counts=True)
    return [(int(v), int(c)) for v, c in zip(vals, counts)]


__all__ = ["NoiseError", "MemoryBudget", "Configuration", "InvalidChannel", "MidCircuit", "Unsupported",
           "PauliChannel", "KrausChannel", "ReadoutError", "depolarizing_error", "amplitude_damping_error",
           "phase_damping_error", "bit_flip_error", "phase_flip_error", "pauli_error", "NoiseModel",
           "aer_gate_names", "NoiseOptions", "resolve", "outcomes", "run_noisy", "run_noisy_sparse", "KINDS"]
