"""`HipSvBackend`: the Backend trait (crates/aleph-backend/src/lib.rs:86-163)
over the C ABI of `libaleph_hip.so`.

Method names, argument meaning and error variants follow the reference:
`allocate`, `apply_gate`, `measure`, `sample`, `expectation_value`,
`probabilities`, `apply_diagonal_phase`, `apply_tiled_block`,
`unpermute_state`, plus the drivers `run`, `run_with_outcomes`
(lib.rs:182-248) and the plan hook `run_program` (the GPU planner path,
crates/aleph-cuda/src/execution.rs:183-194).  Failures raise `BackendError`
subclasses named after the `BackendError` variants (lib.rs:22-81).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .circuit import Barrier, Circuit, DiagonalPhase, GateInstance, PauliString, PAULI_CODE


class BackendError(Exception):
    """Base of the BackendError variants."""

    code = -1

    def __init__(self, message: str = "", **fields):
        super().__init__(message)
        self.fields = fields
        for k, v in fields.items():
            setattr(self, k, v)


def _variant(name: str, code: int):
    return type(name, (BackendError,), {"code": code})


DeviceFailure = _variant("DeviceFailure", L.ERR_DEVICE_FAILURE)
QubitOutOfRange = _variant("QubitOutOfRange", L.ERR_QUBIT_OUT_OF_RANGE)
DuplicateQubit = _variant("DuplicateQubit", L.ERR_DUPLICATE_QUBIT)
ArityMismatch = _variant("ArityMismatch", L.ERR_ARITY_MISMATCH)
UnsupportedGate = _variant("UnsupportedGate", L.ERR_UNSUPPORTED_GATE)
UnsupportedInstruction = _variant("UnsupportedInstruction", L.ERR_UNSUPPORTED_INSTRUCTION)
SymbolicParam = _variant("SymbolicParam", L.ERR_SYMBOLIC_PARAM)
NonFiniteParam = _variant("NonFiniteParam", L.ERR_NON_FINITE_PARAM)
NonUnitaryMatrix = _variant("NonUnitaryMatrix", L.ERR_NON_UNITARY_MATRIX)
EmptyCircuit = _variant("EmptyCircuit", L.ERR_EMPTY_CIRCUIT)
DegenerateMeasurement = _variant("DegenerateMeasurement", L.ERR_DEGENERATE_MEASUREMENT)
TooManyQubits = _variant("TooManyQubits", L.ERR_TOO_MANY_QUBITS)
InvalidPauliString = _variant("InvalidPauliString", L.ERR_INVALID_PAULI_STRING)
InvalidState = _variant("InvalidState", L.ERR_INVALID_STATE)
Optimization = _variant("Optimization", L.ERR_OPTIMIZATION)
NoDevice = _variant("NoDevice", L.ERR_NO_DEVICE)
InvalidArgument = _variant("InvalidArgument", L.ERR_INVALID_ARGUMENT)
OutOfMemory = _variant("OutOfMemory", L.ERR_OUT_OF_MEMORY)

_BY_CODE = {cls.code: cls for cls in (
    DeviceFailure, QubitOutOfRange, DuplicateQubit, ArityMismatch, UnsupportedGate, UnsupportedInstruction,
    SymbolicParam, NonFiniteParam, NonUnitaryMatrix, EmptyCircuit, DegenerateMeasurement, TooManyQubits,
    InvalidPauliString, InvalidState, Optimization, NoDevice, InvalidArgument, OutOfMemory)}


def _raise(ctx, code: int):
    err = L.Error()
    if ctx:
        L.load().aleph_hip_last_error(ctx, C.byref(err))
    cls = _BY_CODE.get(code, BackendError)
    kind = err.kind.decode(errors="replace")
    fields = {}
    if cls is QubitOutOfRange:
        fields = {"qubit": err.qubit, "num_qubits": err.num_qubits}
    elif cls is DuplicateQubit:
        fields = {"qubit": err.qubit}
    elif cls is ArityMismatch:
        fields = {"kind": kind, "expected": err.expected, "got": err.got}
    elif cls in (UnsupportedGate, NonFiniteParam, UnsupportedInstruction):
        fields = {"kind": kind}
    elif cls is NonUnitaryMatrix:
        fields = {"deviation": err.value}
    elif cls is DegenerateMeasurement:
        fields = {"qubit": err.qubit, "probability": err.value}
    elif cls is TooManyQubits:
        fields = {"requested": err.got, "limit": err.expected}
    elif cls in (InvalidState, InvalidPauliString):
        fields = {"reason": kind}
    raise cls(err.message.decode(errors="replace") or cls.__name__, **fields)


def _u32(seq) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(seq, dtype=np.uint32).reshape(-1))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class HipState:
    """A device-resident state vector (CudaSvState analogue)."""

    def __init__(self, backend: "HipSvBackend", handle, num_qubits: int, precision: int):
        self._b = backend
        self._h = handle
        self.num_qubits = num_qubits
        self.precision = precision

    @property
    def handle(self):
        return self._h

    def amplitudes(self) -> np.ndarray:
        out = np.empty(1 << self.num_qubits, dtype=np.complex128)
        self._b._check(L.load().aleph_hip_state_read(self._b._ctx, self._h, _ptr(out.view(np.float64), C.c_double)))
        return out

    def load(self, amps) -> None:
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128).reshape(-1))
        if a.shape[0] != 1 << self.num_qubits:
            raise ValueError("amplitude count must be 2^num_qubits")
        self._b._check(L.load().aleph_hip_state_load(self._b._ctx, self._h, _ptr(a.view(np.float64), C.c_double)))

    def device_ptr(self) -> int:
        p = C.c_void_p()
        L.load().aleph_hip_state_info(self._h, None, None, C.byref(p))
        return int(p.value or 0)

    def free(self) -> None:
        if self._h:
            L.load().aleph_hip_state_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class MeasurementRecord:
    instruction_index: int
    qubit: int
    clbit: int
    outcome: bool


class HipPlan:
    """A compiled program (CudaExecutionPlan analogue)."""

    def __init__(self, backend: "HipSvBackend", handle, num_qubits: int, precision: int):
        self._b, self._h, self.num_qubits, self.precision = backend, handle, num_qubits, precision

    @property
    def handle(self):
        return self._h

    def execute(self, state: HipState) -> None:
        self._b._check(L.load().aleph_hip_plan_execute(self._b._ctx, self._h, state.handle))

    def execute_timed(self, state: HipState) -> tuple[np.ndarray, np.ndarray]:
        """Per-launch device milliseconds and launch kinds (event pairs on the context stream)."""
        n = self.stats()["n_launches"]
        ms = np.zeros(max(n, 1), dtype=np.float64)
        kinds = np.zeros(max(n, 1), dtype=np.uint32)
        self._b._check(L.load().aleph_hip_plan_execute_timed(self._b._ctx, self._h, state.handle,
                                                             _ptr(ms, C.c_double), _ptr(kinds, C.c_uint32)))
        return ms[:n], kinds[:n]

    def stats(self) -> dict:
        s = L.PlanStats()
        L.load().aleph_hip_plan_stats_get(self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in L.PlanStats._fields_ if f != "reserved"}

    def free(self) -> None:
        if self._h:
            L.load().aleph_hip_plan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipSvBackend:
    """MI355X state-vector backend (`--backend hip` / `hip-f32`)."""

    def __init__(self, seed: int = 0, precision: int = L.F64, ordinal: int = 0):
        lib = L.load()
        ctx = C.c_void_p()
        code = lib.aleph_hip_ctx_create(ordinal, C.c_uint64(seed & (2**64 - 1)), C.byref(ctx))
        if code != L.OK:
            _raise(None, code)
        self._ctx = ctx
        self.precision = precision

    @staticmethod
    def with_seed(seed: int, precision: int = L.F64) -> "HipSvBackend":
        return HipSvBackend(seed=seed, precision=precision)

    def with_qubit_cap(self, cap: int) -> "HipSvBackend":
        self._check(L.load().aleph_hip_ctx_set_qubit_cap(self._ctx, cap))
        return self

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            L.load().aleph_hip_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code: int) -> None:
        if code != L.OK:
            _raise(self._ctx, code)

    def stream(self) -> int:
        p = C.c_void_p()
        self._check(L.load().aleph_hip_ctx_stream(self._ctx, C.byref(p)))
        return int(p.value or 0)

    def synchronize(self) -> None:
        self._check(L.load().aleph_hip_synchronize(self._ctx))

    # ---- Backend trait ----
    def allocate(self, num_qubits: int, precision: int | None = None) -> HipState:
        """A |0...0> state in the backend's precision (or `precision`, 64 or 32)."""
        prec = self.precision if precision is None else precision
        h = C.c_void_p()
        self._check(L.load().aleph_hip_state_alloc(self._ctx, num_qubits, prec, C.byref(h)))
        return HipState(self, h, num_qubits, prec)

    def apply_gate(self, state: HipState, gate: GateInstance) -> None:
        g = L.Gate()
        g.kind = gate.gate.kind
        g.n_qubits = len(gate.qubits)
        g.n_controls = len(gate.controls)
        g.k = gate.gate.k
        qs, cs = _u32(gate.qubits or [0]), _u32(gate.controls or [0])
        g.qubits = _ptr(qs, C.c_uint32)
        g.controls = _ptr(cs, C.c_uint32)
        for i, p in enumerate(gate.gate.params[:3]):
            g.params[i] = p
        pay = np.ascontiguousarray(gate.gate.payload())
        g.matrix = _ptr(pay, C.c_double) if pay.size else None
        self._check(L.load().aleph_hip_apply_gate(self._ctx, state.handle, C.byref(g)))

    def measure(self, state: HipState, qubit: int) -> bool:
        out = C.c_int32()
        self._check(L.load().aleph_hip_measure(self._ctx, state.handle, qubit, C.byref(out)))
        return bool(out.value)

    def measure_with_uniform(self, state: HipState, qubit: int, u: float) -> tuple[bool, bool]:
        out, used = C.c_int32(), C.c_int32()
        self._check(L.load().aleph_hip_measure_with_uniform(self._ctx, state.handle, qubit, u, C.byref(out),
                                                            C.byref(used)))
        return bool(out.value), bool(used.value)

    def sample(self, state: HipState, shots: int, uniforms=None) -> np.ndarray:
        out = np.zeros(max(shots, 1), dtype=np.uint64)
        if uniforms is None:
            self._check(L.load().aleph_hip_sample(self._ctx, state.handle, shots, _ptr(out, C.c_uint64)))
        else:
            u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
            if u.shape[0] < shots:
                raise ValueError("need one uniform per shot")
            self._check(L.load().aleph_hip_sample_with_uniforms(self._ctx, state.handle, _ptr(u, C.c_double), shots,
                                                                _ptr(out, C.c_uint64)))
        return out[:shots]

    def expectation_value(self, state: HipState, pauli: PauliString) -> float:
        qs = _u32([q for q, _ in pauli.terms] or [0])
        ps = np.asarray([PAULI_CODE[p] for _, p in pauli.terms] or [0], dtype=np.uint8)
        out = C.c_double()
        self._check(L.load().aleph_hip_expectation(self._ctx, state.handle, float(pauli.coefficient),
                                                   _ptr(qs, C.c_uint32), _ptr(ps, C.c_uint8), len(pauli.terms),
                                                   C.byref(out)))
        return float(out.value)

    def expectation_pauli_sum(self, state: HipState, paulis) -> float:
        """sum of <P> over PauliStrings (each with its coefficient), one call
        (Backend::expectation_pauli_sum, crates/aleph-backend/src/lib.rs:422-432)."""
        paulis = list(paulis)
        coeffs = np.asarray([float(p.coefficient) for p in paulis] or [0.0], dtype=np.float64)
        offs = np.zeros(len(paulis) + 1, dtype=np.uint32)
        offs[1:] = np.cumsum([len(p.terms) for p in paulis]) if paulis else []
        qs = _u32([q for p in paulis for q, _ in p.terms] or [0])
        ps = np.asarray([PAULI_CODE[c] for p in paulis for _, c in p.terms] or [0], dtype=np.uint8)
        out = C.c_double()
        self._check(L.load().aleph_hip_expectation_sum(self._ctx, state.handle, len(paulis), _ptr(coeffs, C.c_double),
                                                       _ptr(offs, C.c_uint32), _ptr(qs, C.c_uint32),
                                                       _ptr(ps, C.c_uint8), C.byref(out)))
        return float(out.value)

    def probabilities(self, state: HipState, qubits) -> np.ndarray:
        qs = _u32(list(qubits) or [0])
        out = np.zeros(1 << len(qubits), dtype=np.float64)
        self._check(L.load().aleph_hip_probabilities(self._ctx, state.handle, _ptr(qs, C.c_uint32), len(qubits),
                                                     _ptr(out, C.c_double)))
        return out

    def apply_diagonal_phase(self, state: HipState, dp: DiagonalPhase) -> None:
        angles = np.asarray([t.angle for t in dp.terms] or [0.0], dtype=np.float64)
        conds = np.asarray([m for t in dp.terms for m in t.conds] or [0], dtype=np.uint64)
        offs = np.zeros(len(dp.terms) + 1, dtype=np.uint32)
        offs[1:] = np.cumsum([len(t.conds) for t in dp.terms]) if dp.terms else []
        self._check(L.load().aleph_hip_apply_diagonal_phase(self._ctx, state.handle, len(dp.terms),
                                                            _ptr(angles, C.c_double), _ptr(conds, C.c_uint64),
                                                            _ptr(offs, C.c_uint32)))

    def apply_tiled_block(self, state: HipState, gates) -> None:
        """Default replay (lib.rs:140-154): tile-major execution is the planner's job here."""
        for g in gates:
            self.apply_gate(state, g)

    def unpermute_state(self, state: HipState, perm) -> None:
        p = _u32(perm)
        if p.shape[0] != state.num_qubits:
            raise InvalidArgument("permutation length must equal num_qubits")
        self._check(L.load().aleph_hip_unpermute(self._ctx, state.handle, _ptr(p, C.c_uint32)))

    def linear_xeb(self, state: HipState) -> float:
        """2^n * sum_x p(x)^2 - 1 of the exact state (benches/src/lib.rs:29 linear_xeb)."""
        out = C.c_double()
        self._check(L.load().aleph_hip_linear_xeb(self._ctx, state.handle, C.byref(out)))
        return float(out.value)

    def norm_sqr(self, state: HipState) -> float:
        out = C.c_double()
        self._check(L.load().aleph_hip_norm_sqr(self._ctx, state.handle, C.byref(out)))
        return float(out.value)

    def normalize(self, state: HipState) -> float:
        """psi /= sqrt(sum |psi|^2), the sum in FP64 on the device; returns the norm before."""
        out = C.c_double()
        self._check(L.load().aleph_hip_state_normalize(self._ctx, state.handle, C.byref(out)))
        return float(out.value)

    def compare(self, a: HipState, b: HipState) -> dict:
        """<a|b>, fidelity |<a|b>|^2, sum |a-b|^2 and max |a-b| of two states of one size
        (any precisions), accumulated in FP64 on the device."""
        out = (C.c_double * 4)()
        self._check(L.load().aleph_hip_state_compare(self._ctx, a.handle, b.handle, out))
        ov = complex(out[0], out[1])
        return {"overlap": ov, "fidelity": abs(ov) ** 2, "l2_sq": out[2], "max_abs": out[3]}

    # ---- plans ----
    def compile(self, circuit: Circuit, planner: int = L.PLANNER_RULE, local_qubits: int = 0, max_ops: int = 0,
                coalesce_bits: int = 0, fuse_1q: bool = True, reg_bits: int = 0, nontemporal: bool = False,
                fuse_2q: bool = True, fuse_diag: bool = True, kq_width: int = 0, jit: bool = False,
                renormalize: bool = False, generic_phase: bool = False, precision: int | None = None) -> HipPlan:
        prec = self.precision if precision is None else precision
        words, reals = circuit.encode()
        nw, nr = int(words.size), int(reals.size)
        words = np.ascontiguousarray(words if nw else np.zeros(1, np.uint32))
        reals = np.ascontiguousarray(reals if nr else np.zeros(1, np.float64))
        opts = L.PlanOpts(planner, local_qubits, max_ops, coalesce_bits, int(fuse_1q), reg_bits,
                          (L.PLAN_NONTEMPORAL if nontemporal else 0) | (0 if fuse_2q else L.PLAN_NO_FUSE_2Q)
                          | (0 if fuse_diag else L.PLAN_NO_FUSE_DIAG) | (L.PLAN_JIT if jit else 0)
                          | (L.PLAN_RENORMALIZE if renormalize else 0)
                          | (L.PLAN_GENERIC_PHASE if generic_phase else 0), kq_width)
        h = C.c_void_p()
        self._check(L.load().aleph_hip_plan_create(self._ctx, circuit.num_qubits, prec,
                                                   _ptr(words, C.c_uint32), nw, _ptr(reals, C.c_double), nr,
                                                   C.byref(opts), C.byref(h)))
        return HipPlan(self, h, circuit.num_qubits, prec)

    def run_program(self, state: HipState, circuit: Circuit, **plan_kw) -> dict:
        plan = self.compile(circuit, **{"precision": state.precision, **plan_kw})
        try:
            plan.execute(state)
            self.synchronize()
            return plan.stats()
        finally:
            plan.free()


def run_with_outcomes(backend: HipSvBackend, circuit: Circuit):
    """Instruction loop of lib.rs:211-248 (gates one by one, measurements recorded)."""
    if circuit.num_qubits == 0 and len(circuit) == 0:
        raise EmptyCircuit("cannot run an empty circuit")
    state = backend.allocate(circuit.num_qubits)
    outcomes = []
    for idx, inst in enumerate(circuit.instructions):
        if isinstance(inst, GateInstance):
            backend.apply_gate(state, inst)
        elif isinstance(inst, DiagonalPhase):
            backend.apply_diagonal_phase(state, inst)
        elif isinstance(inst, Barrier):
            continue
        elif isinstance(inst, tuple) and inst and inst[0] == "measure":
            _, q, clbit = inst
            outcomes.append(MeasurementRecord(idx, q, clbit, backend.measure(state, q)))
        else:
            raise UnsupportedInstruction(f"unsupported instruction {inst!r}", kind=type(inst).__name__)
    return state, outcomes


def run(backend: HipSvBackend, circuit: Circuit) -> HipState:
    return run_with_outcomes(backend, circuit)[0]


def run_circuit_with_outcomes(backend: HipSvBackend, circuit: Circuit, uniforms=None, **plan_kw):
    """Whole-circuit path: every measurement-free segment is planned into
    fused passes and executed in one library call; measurements in between
    go through `measure` (or `measure_with_uniform` with the next of
    `uniforms`, consumed only when the outcome is not forced), with the
    records of run_with_outcomes (crates/aleph-backend/src/lib.rs:211-248)."""
    if circuit.num_qubits == 0 and len(circuit) == 0:
        raise EmptyCircuit("cannot run an empty circuit")
    state = backend.allocate(circuit.num_qubits)
    outcomes = []
    draws = iter(uniforms) if uniforms is not None else None
    segment = Circuit(circuit.num_qubits)

    def flush():
        nonlocal segment
        if len(segment):
            backend.run_program(state, segment, **plan_kw)
        segment = Circuit(circuit.num_qubits)

    for idx, inst in enumerate(circuit.instructions):
        if isinstance(inst, tuple) and inst and inst[0] == "measure":
            flush()
            _, q, clbit = inst
            if draws is None:
                bit = backend.measure(state, q)
            else:
                u = next(draws)
                bit, used = backend.measure_with_uniform(state, q, u)
                if not used:      # forced outcome: the uniform stays for the next measurement
                    draws = iter([u, *draws])
            outcomes.append(MeasurementRecord(idx, q, clbit, bit))
        elif isinstance(inst, (GateInstance, DiagonalPhase, Barrier)):
            segment.add(inst)
        else:
            raise UnsupportedInstruction(f"unsupported instruction {inst!r}", kind=type(inst).__name__)
    flush()
    return state, outcomes
