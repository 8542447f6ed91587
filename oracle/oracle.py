"""ctypes wrapper of the CPU oracle `liboracle_sv.so` — TEST INFRASTRUCTURE ONLY.

Used by tests/, `__graft_entry__.smoke()` (as the checker) and bench.py's
cpu_baseline leg (as the timed CPU baseline).  The HIP product never
imports this module.  The oracle restates the reference CPU FP64 `sv`
backend (`NaiveSvBackend`, crates/aleph-sv/src/backend.rs); see
sv_oracle.cpp for the function-by-function citations.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_sv.so"

u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
PU32, PU64, PF64, PI32, PU8 = (C.POINTER(u32), C.POINTER(u64), C.POINTER(f64), C.POINTER(i32),
                               C.POINTER(C.c_uint8))

_SIG = {
    "orc_set_threads": (None, [C.c_int]),
    "orc_get_threads": (C.c_int, []),
    "orc_last_error": (i32, [PU32, PU32, PU32, PU32, PF64, C.c_char_p, C.c_int]),
    "orc_state_new": (i32, [u32, u32, C.POINTER(C.c_void_p)]),
    "orc_state_free": (None, [C.c_void_p]),
    "orc_state_load": (None, [C.c_void_p, PF64]),
    "orc_state_read": (None, [C.c_void_p, PF64]),
    "orc_apply_gate": (i32, [C.c_void_p, u32, PF64, PU32, u32, PU32, u32, PF64, u32]),
    "orc_apply_diagonal_phase": (i32, [C.c_void_p, u32, PF64, PU64, PU32]),
    "orc_apply_tiled_block": (i32, [C.c_void_p, u32, u32, PU32, PF64, PU32, PU32, PU32, PF64]),
    "orc_unpermute": (None, [C.c_void_p, PU32]),
    "orc_validate_state": (i32, [C.c_void_p]),
    "orc_norm_sqr": (f64, [C.c_void_p]),
    "orc_measure": (i32, [C.c_void_p, u32, f64, PI32, PI32]),
    "orc_sample": (i32, [C.c_void_p, u32, u64, PU64]),
    "orc_expectation": (i32, [C.c_void_p, f64, PU32, PU8, u32, PF64]),
    "orc_probabilities": (i32, [C.c_void_p, PU32, u32, PF64]),
    "orc_gate_matrix": (i32, [u32, PF64, PF64, PF64, PI32]),
}

ERR_NAMES = {0: "OK", 1: "DeviceFailure", 2: "QubitOutOfRange", 3: "DuplicateQubit", 4: "ArityMismatch",
             5: "UnsupportedGate", 6: "UnsupportedInstruction", 7: "SymbolicParam", 8: "NonFiniteParam",
             9: "NonUnitaryMatrix", 10: "EmptyCircuit", 11: "DegenerateMeasurement", 12: "TooManyQubits",
             13: "InvalidPauliString", 14: "InvalidState", 17: "InvalidArgument"}
PAULI = {"I": 0, "X": 1, "Y": 2, "Z": 3}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing; run `python oracle/build.py`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class OracleError(Exception):
    def __init__(self, code: int):
        q, nq, e, g = u32(), u32(), u32(), u32()
        v = f64()
        kind = C.create_string_buffer(64)
        load().orc_last_error(C.byref(q), C.byref(nq), C.byref(e), C.byref(g), C.byref(v), kind, 64)
        self.code = code
        self.name = ERR_NAMES.get(code, str(code))
        self.qubit, self.num_qubits, self.expected, self.got = q.value, nq.value, e.value, g.value
        self.value = v.value
        self.kind = kind.value.decode()
        super().__init__(f"{self.name}(kind={self.kind}, qubit={self.qubit})")


def _chk(code: int) -> None:
    if code != 0:
        raise OracleError(code)


def _arr(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1))


def set_threads(n: int) -> None:
    load().orc_set_threads(int(n))


def gate_matrix(gate) -> np.ndarray:
    """The oracle's Gate::matrix for a Gate (kinds.rs:213-399)."""
    p = _arr(list(gate.params) + [0.0] * 3, np.float64)
    pay = _arr(gate.payload() if gate.matrix is not None else [0.0], np.float64)
    out = np.zeros(128, dtype=np.float64)
    dim = i32()
    _chk(load().orc_gate_matrix(gate.kind, p.ctypes.data_as(PF64), pay.ctypes.data_as(PF64),
                                out.ctypes.data_as(PF64), C.byref(dim)))
    d = dim.value
    return out[: 2 * d * d].view(np.complex128).reshape(d, d)


class OracleState:
    def __init__(self, n: int, cap: int = 28):
        h = C.c_void_p()
        _chk(load().orc_state_new(n, cap, C.byref(h)))
        self.h = h
        self.num_qubits = n

    def __del__(self):
        try:
            if self.h:
                load().orc_state_free(self.h)
                self.h = None
        except Exception:
            pass

    def amplitudes(self) -> np.ndarray:
        out = np.empty(1 << self.num_qubits, dtype=np.complex128)
        load().orc_state_read(self.h, out.view(np.float64).ctypes.data_as(PF64))
        return out

    def load(self, amps) -> None:
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128))
        load().orc_state_load(self.h, a.view(np.float64).ctypes.data_as(PF64))


class OracleSv:
    """Backend-shaped facade over the oracle (allocate/apply_gate/measure/...)."""

    def __init__(self, seed: int = 0, cap: int = 28):
        self.seed = seed
        self.cap = cap
        self._draws = 0

    def allocate(self, n: int) -> OracleState:
        return OracleState(n, self.cap)

    def apply_gate(self, st: OracleState, g) -> None:
        params = _arr(list(g.gate.params) + [0.0] * 3, np.float64)
        qs = _arr(list(g.qubits) or [0], np.uint32)
        cs = _arr(list(g.controls) or [0], np.uint32)
        pay = _arr(g.gate.payload() if g.gate.matrix is not None else [0.0], np.float64)
        _chk(load().orc_apply_gate(st.h, g.gate.kind, params.ctypes.data_as(PF64), qs.ctypes.data_as(PU32),
                                   len(g.qubits), cs.ctypes.data_as(PU32), len(g.controls),
                                   pay.ctypes.data_as(PF64), g.gate.k))

    def apply_diagonal_phase(self, st: OracleState, dp) -> None:
        ang = _arr([t.angle for t in dp.terms] or [0.0], np.float64)
        conds = _arr([m for t in dp.terms for m in t.conds] or [0], np.uint64)
        offs = np.zeros(len(dp.terms) + 1, dtype=np.uint32)
        if dp.terms:
            offs[1:] = np.cumsum([len(t.conds) for t in dp.terms])
        _chk(load().orc_apply_diagonal_phase(st.h, len(dp.terms), ang.ctypes.data_as(PF64),
                                             conds.ctypes.data_as(PU64), offs.ctypes.data_as(PU32)))

    def apply_tiled_block(self, st: OracleState, gates, tile_bits: int) -> None:
        ng = len(gates)
        kinds = _arr([g.gate.kind for g in gates], np.uint32)
        params = np.zeros(3 * ng)
        qubits = np.zeros(2 * ng, dtype=np.uint32)
        mats = np.zeros(32 * ng)
        ctrl = [c for g in gates for c in g.controls]
        off = np.zeros(ng + 1, dtype=np.uint32)
        for i, g in enumerate(gates):
            params[3 * i: 3 * i + len(g.gate.params)] = g.gate.params
            qubits[2 * i: 2 * i + len(g.qubits)] = g.qubits
            if g.gate.matrix is not None:
                pay = g.gate.payload()
                mats[32 * i: 32 * i + pay.size] = pay
            off[i + 1] = off[i] + len(g.controls)
        ctrl = _arr(ctrl or [0], np.uint32)
        _chk(load().orc_apply_tiled_block(st.h, tile_bits, ng, kinds.ctypes.data_as(PU32),
                                          params.ctypes.data_as(PF64), qubits.ctypes.data_as(PU32),
                                          off.ctypes.data_as(PU32), ctrl.ctypes.data_as(PU32),
                                          mats.ctypes.data_as(PF64)))

    def unpermute_state(self, st: OracleState, perm) -> None:
        p = _arr(perm, np.uint32)
        load().orc_unpermute(st.h, p.ctypes.data_as(PU32))

    def measure_with_uniform(self, st: OracleState, qubit: int, u: float):
        out, used = i32(), i32()
        _chk(load().orc_measure(st.h, qubit, u, C.byref(out), C.byref(used)))
        return bool(out.value), bool(used.value)

    def sample(self, st: OracleState, shots: int, seed: int | None = None) -> np.ndarray:
        out = np.zeros(max(shots, 1), dtype=np.uint64)
        s = self.seed + self._draws if seed is None else seed
        self._draws += 1
        _chk(load().orc_sample(st.h, shots, s, out.ctypes.data_as(PU64)))
        return out[:shots]

    def expectation_value(self, st: OracleState, pauli) -> float:
        qs = _arr([q for q, _ in pauli.terms] or [0], np.uint32)
        ps = _arr([PAULI[p] for _, p in pauli.terms] or [0], np.uint8)
        out = f64()
        _chk(load().orc_expectation(st.h, float(pauli.coefficient), qs.ctypes.data_as(PU32),
                                    ps.ctypes.data_as(PU8), len(pauli.terms), C.byref(out)))
        return out.value

    def probabilities(self, st: OracleState, qubits) -> np.ndarray:
        qs = _arr(list(qubits) or [0], np.uint32)
        out = np.zeros(1 << len(qubits))
        _chk(load().orc_probabilities(st.h, qs.ctypes.data_as(PU32), len(qubits), out.ctypes.data_as(PF64)))
        return out

    def validate(self, st: OracleState) -> None:
        _chk(load().orc_validate_state(st.h))

    def run(self, circuit, state: OracleState | None = None) -> OracleState:
        """Gate-by-gate replay of a Circuit (run_with_outcomes without measurements)."""
        from aleph_amd.circuit import Barrier, DiagonalPhase, GateInstance
        st = state if state is not None else self.allocate(circuit.num_qubits)
        for inst in circuit.instructions:
            if isinstance(inst, GateInstance):
                self.apply_gate(st, inst)
            elif isinstance(inst, DiagonalPhase):
                self.apply_diagonal_phase(st, inst)
            elif isinstance(inst, Barrier):
                continue
            else:
                raise TypeError(f"oracle cannot run {inst!r}")
        return st
