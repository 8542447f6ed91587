"""Distributed state vector, one-device part: rank slices, exchanges, plans.

Python mirror of crates/aleph-ir/src/dist (DistLayout, DistStep, DistPlan,
specialize) and of the executor crates/aleph-cuda/src/dist/mod.rs with the
single-card transport LocalExchange (exchange.rs).  A 2^n state is split
over R = 2^g ranks: physical index = (rank << m) | local, m = n - g;
physical qubits >= m are global and select the rank (Häner & Steiger,
SC'17, section 3).

* `plan` routes a circuit: it keeps a lazy logical -> physical map, absorbs
  uncontrolled SWAPs as relabels, and inserts an `Exchange` only when a
  non-diagonal target sits on a global qubit, evicting the local qubits
  whose next use is farthest away (Belady).  Everything else is rewritten
  per rank by `specialize`.
* `DistSvBackend` holds every rank slice on one device (the LocalExchange
  pattern: all ranks share one GPU), runs each rank's specialised local
  steps through the fused-pass engine and performs each exchange with one
  in-place swap kernel over all moved chunk pairs
  (`aleph_hip_dist_exchange`).

Over several GPUs the same plan runs one rank (or a block of ranks) per
device and the exchange becomes RCCL send/recv of the same chunk pairs.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .backend import HipState, HipSvBackend, _ptr
from .circuit import Barrier, Circuit, DiagonalPhase, Gate, GateInstance, PhaseTerm

MAX_GLOBAL_QUBITS = 31


class DistError(Exception):
    """DistError (crates/aleph-ir/src/dist/mod.rs:146-165)."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind


@dataclass(frozen=True)
class DistLayout:
    """Global/local split of an n-qubit state over 2^g ranks."""

    n: int
    g: int

    def __post_init__(self):
        if self.n < 1 or self.n > 64 or self.g >= self.n or self.g > MAX_GLOBAL_QUBITS or self.g < 0:
            raise DistError("BadLayout", f"bad layout: n={self.n}, g={self.g} (need 1 <= n <= 64, g < n, g <= 31)")

    def m(self) -> int:
        return self.n - self.g

    def ranks(self) -> int:
        return 1 << self.g

    def is_global(self, p: int) -> bool:
        return p >= self.m()

    def rank_bit(self, rank: int, p: int) -> int:
        return (rank >> (p - self.m())) & 1


@dataclass(frozen=True)
class Exchange:
    """Swap global physical bit global_bits[j] with local bit m - k + j."""

    global_bits: tuple


@dataclass
class CommStats:
    exchanges: int = 0
    amps_moved_per_rank: int = 0
    local_swaps: int = 0
    relabels: int = 0


@dataclass
class DistPlan:
    layout: DistLayout
    steps: list                 # list[list[instruction]] (physical indices) | Exchange
    final_map: list             # final_map[logical] = physical
    stats: CommStats = field(default_factory=CommStats)


# ---------------------------------------------------------------- gates

_SPLIT_CONTROL = {"Cnot": "X", "CRx": "Rx", "CRy": "Ry"}


def gate_matrix(gate: Gate) -> np.ndarray:
    """The library's Gate::matrix of `gate` (aleph_hip_gate_matrix)."""
    k = gate.arity()
    g = L.Gate()
    g.kind, g.n_qubits, g.n_controls, g.k = gate.kind, k, 0, gate.k
    qs = np.ascontiguousarray(np.arange(k, dtype=np.uint32))
    g.qubits = _ptr(qs, C.c_uint32)
    g.controls = _ptr(qs, C.c_uint32)
    for i, p in enumerate(gate.params[:3]):
        g.params[i] = p
    pay = np.ascontiguousarray(gate.payload())
    g.matrix = _ptr(pay, C.c_double) if pay.size else None
    out = np.zeros(2 * 4 ** 6)
    dim = C.c_uint32()
    code = L.load().aleph_hip_gate_matrix(C.byref(g), k, _ptr(out, C.c_double), C.byref(dim))
    if code != L.OK:
        raise DistError("Gate", f"invalid gate {gate.name}")
    d = dim.value
    return out[:2 * d * d].view(np.complex128).reshape(d, d)


def _required_local(inst: GateInstance) -> tuple:
    """Qubits a gate acts on non-diagonally: they must be local when it runs.
    A control of Cnot/CRx/CRy/Toffoli and every qubit of a diagonal gate may
    stay global (specialize folds them per rank)."""
    g = inst.gate
    if g.is_diagonal():
        return ()
    if g.name in _SPLIT_CONTROL:
        return (inst.qubits[1],)
    if g.name == "Toffoli":
        return (inst.qubits[2],)
    return tuple(inst.qubits)


def specialize(inst, layout: DistLayout, rank: int):
    """Rewrite a physical instruction for `rank` (specialize.rs:14-33):
    None when it is a no-op there; every output qubit is < m."""
    if isinstance(inst, Barrier):
        return None
    if isinstance(inst, DiagonalPhase):
        return _specialize_phase(inst, layout, rank)
    if not isinstance(inst, GateInstance):
        raise DistError("Unsupported", f"instruction {inst!r} is not supported by the distributed planner")
    controls = []
    for c in inst.controls:
        if layout.is_global(c):
            if layout.rank_bit(rank, c) == 0:
                return None
        else:
            controls.append(c)
    gate, qubits = inst.gate, list(inst.qubits)
    if gate.name in _SPLIT_CONTROL and layout.is_global(qubits[0]):
        if layout.rank_bit(rank, qubits[0]) == 0:
            return None
        gate = Gate(_SPLIT_CONTROL[gate.name], gate.params)
        qubits = qubits[1:]
    elif gate.name == "Toffoli" and (layout.is_global(qubits[0]) or layout.is_global(qubits[1])):
        kept = []
        for c in qubits[:2]:
            if layout.is_global(c):
                if layout.rank_bit(rank, c) == 0:
                    return None
            else:
                kept.append(c)
        gate, qubits = (Gate.X(), [qubits[2]]) if not kept else (Gate.Cnot(), [kept[0], qubits[2]])
    if gate.is_diagonal() and any(layout.is_global(p) for p in qubits):
        return _reduce_diagonal(gate, qubits, controls, layout, rank)
    for p in qubits:
        if layout.is_global(p):
            raise DistError("GlobalTarget", f"non-diagonal target on global qubit {p} reached specialize")
    return GateInstance(gate, qubits, controls)


def _reduce_diagonal(gate: Gate, qubits, controls, layout: DistLayout, rank: int):
    """A diagonal gate restricted to this rank's values of its global qubits."""
    d = np.diag(gate_matrix(gate)) if gate.matrix is None or gate.matrix.ndim == 2 else np.asarray(gate.matrix)
    a = len(qubits)
    local = [p for p in qubits if not layout.is_global(p)]
    nl = len(local)
    dl = []
    for ip in range(1 << nl):
        i, t = 0, 0
        for j, p in enumerate(qubits):
            if layout.is_global(p):
                bit = layout.rank_bit(rank, p)
            else:
                bit = (ip >> (nl - 1 - t)) & 1
                t += 1
            i |= bit << (a - 1 - j)
        dl.append(complex(d[i]))
    if nl == 0:
        return DiagonalPhase(layout.m(), [PhaseTerm(tuple(1 << c for c in controls), float(np.angle(dl[0])))])
    if nl > 2:
        raise DistError("Unsupported", "diagonal gate on >2 local qubits with a global qubit")
    return GateInstance(Gate.diag(np.array(dl)), local, controls)


def _specialize_phase(dp: DiagonalPhase, layout: DistLayout, rank: int):
    """Fold the rank's global parities (specialize.rs:190-236): a condition
    with a known odd global parity and a local part B fires iff B's parity
    is even, [A and not B1 .. not Bk] = sum over S of (-1)^|S| [A and S]."""
    m = layout.m()
    lo_mask = (1 << m) - 1
    rank_mask = rank << m
    out = []
    for t in dp.terms:
        pos, neg = [], []
        never = False
        for mask in t.conds:
            lo = mask & lo_mask
            hi = bin(mask & rank_mask).count("1") & 1
            if hi == 0 and lo == 0:
                never = True
                break
            if hi == 0:
                pos.append(lo)
            elif lo != 0:
                neg.append(lo)
        if never:
            continue
        for s in range(1 << len(neg)):
            conds = list(pos) + [b for j, b in enumerate(neg) if (s >> j) & 1]
            sign = -1.0 if bin(s).count("1") & 1 else 1.0
            out.append(PhaseTerm(tuple(conds), sign * t.angle))
    if not out:
        return None
    return DiagonalPhase(m, out)


# ---------------------------------------------------------------- planner

def plan(circuit: Circuit, layout: DistLayout) -> DistPlan:
    """Route `circuit` onto `layout` (exchanges on demand, Belady victims)."""
    n, m = layout.n, layout.m()
    if circuit.num_qubits != n:
        raise DistError("QubitCountMismatch", f"circuit has {circuit.num_qubits} qubits but layout has n={n}")
    insts = list(circuit.instructions)
    uses = [[] for _ in range(n)]
    for i, inst in enumerate(insts):
        if isinstance(inst, tuple):
            raise DistError("Unsupported", f"instruction `{inst[0]}` is not supported by the distributed planner")
        if isinstance(inst, GateInstance):
            for q in (*inst.qubits, *inst.controls):
                uses[q].append(i)
        elif isinstance(inst, DiagonalPhase):
            for t in inst.terms:
                for mask in t.conds:
                    for q in range(n):
                        if (mask >> q) & 1:
                            uses[q].append(i)
    for u in uses:
        u.sort()
    cursor = [0] * n
    phys = list(range(n))
    inv = list(range(n))
    stats = CommStats()
    steps: list = []
    local: list = []

    def next_use(q: int, after: int) -> int:
        u = uses[q]
        while cursor[q] < len(u) and u[cursor[q]] <= after:
            cursor[q] += 1
        return u[cursor[q]] if cursor[q] < len(u) else 1 << 62

    def relabel(a: int, b: int) -> None:
        pa, pb = phys[a], phys[b]
        phys[a], phys[b] = pb, pa
        inv[pa], inv[pb] = b, a

    for i, inst in enumerate(insts):
        if isinstance(inst, Barrier):
            continue
        if isinstance(inst, DiagonalPhase):
            terms = [PhaseTerm(tuple(_map_mask(mask, phys) for mask in t.conds), t.angle) for t in inst.terms]
            local.append(DiagonalPhase(n, terms))
            continue
        if inst.gate.name == "Swap" and not inst.controls:
            relabel(*inst.qubits)
            stats.relabels += 1
            continue
        need = [q for q in _required_local(inst) if layout.is_global(phys[q])]
        if need:
            k = len(need)
            if len(_required_local(inst)) > m:
                raise DistError("TooFewLocalQubits", f"gate needs {len(_required_local(inst))} local qubits but "
                                                     f"only m={m} are local")
            busy = set(inst.qubits) | set(inst.controls)
            cands = [inv[p] for p in range(m) if inv[p] not in busy]
            if len(cands) < k:
                raise DistError("TooFewLocalQubits", f"gate needs {k} free local qubits but only m={m} are local")
            cands.sort(key=lambda q: (-next_use(q, i), -phys[q]))
            victims = cands[:k]
            # Victims onto the top k local slots (local SWAPs move data inside every rank).
            top = set(range(m - k, m))
            movers = [v for v in victims if phys[v] not in top]
            free = [p for p in range(m - k, m) if inv[p] not in victims]
            for v, slot in zip(movers, free):
                local.append(GateInstance(Gate.Swap(), [phys[v], slot]))
                relabel(v, inv[slot])
                stats.local_swaps += 1
            gbits = tuple(phys[q] for q in need)
            if local:
                steps.append(local)
                local = []
            steps.append(Exchange(gbits))
            for j, gp in enumerate(gbits):
                relabel(inv[gp], inv[m - k + j])
            stats.exchanges += 1
            stats.amps_moved_per_rank += ((1 << k) - 1) << (m - k)
        local.append(GateInstance(inst.gate, [phys[q] for q in inst.qubits], [phys[c] for c in inst.controls]))
    if local:
        steps.append(local)
    return DistPlan(layout, steps, list(phys), stats)


def _map_mask(mask: int, phys) -> int:
    out = 0
    q = 0
    while mask:
        if mask & 1:
            out |= 1 << phys[q]
        mask >>= 1
        q += 1
    return out


# ---------------------------------------------------------------- executor

class DistSvState:
    """Rank slices plus the plan's final logical -> physical map."""

    def __init__(self, layout: DistLayout, final_map, ranks):
        self.layout = layout
        self.final_map = list(final_map)
        self.ranks = ranks

    def physical(self) -> np.ndarray:
        return np.concatenate([r.amplitudes() for r in self.ranks])

    def amplitudes(self) -> np.ndarray:
        """The 2^n state in logical qubit order (dist_ref.rs:63-72)."""
        phys = self.physical()
        n = self.layout.n
        x = np.arange(1 << n, dtype=np.int64)
        p = np.zeros_like(x)
        for lq, pq in enumerate(self.final_map):
            p |= ((x >> lq) & 1) << pq
        return phys[p]

    def free(self) -> None:
        for r in self.ranks:
            r.free()
        self.ranks = []


class DistSvBackend:
    """Runs DistPlans with every rank slice on one backend's device."""

    def __init__(self, backend: HipSvBackend, fuse: bool = True, **plan_kw):
        self.backend = backend
        self.fuse = fuse
        self.plan_kw = plan_kw

    def alloc_ranks(self, layout: DistLayout) -> list:
        out = []
        for r in range(layout.ranks()):
            h = C.c_void_p()
            self.backend._check(L.load().aleph_hip_dist_alloc_rank(self.backend._ctx, layout.m(),
                                                                   self.backend.precision, r, C.byref(h)))
            out.append(HipState(self.backend, h, layout.m(), self.backend.precision))
        return out

    def exchange(self, ranks, layout: DistLayout, global_bits) -> None:
        handles = (C.c_void_p * len(ranks))(*[r.handle for r in ranks])
        bits = np.ascontiguousarray(np.asarray(global_bits, dtype=np.uint32))
        self.backend._check(L.load().aleph_hip_dist_exchange(self.backend._ctx, handles, len(ranks),
                                                             _ptr(bits, C.c_uint32), len(bits)))

    def run(self, plan: DistPlan) -> DistSvState:
        layout = plan.layout
        ranks = self.alloc_ranks(layout)
        kw = dict(self.plan_kw)
        if not self.fuse:
            kw["planner"] = L.PLANNER_UNFUSED
        for step in plan.steps:
            if isinstance(step, Exchange):
                self.exchange(ranks, layout, step.global_bits)
                continue
            for r, st in enumerate(ranks):
                c = Circuit(layout.m())
                for inst in step:
                    s = specialize(inst, layout, r)
                    if s is not None:
                        c.add(s)
                if len(c):
                    self.backend.run_program(st, c, **kw)
        return DistSvState(layout, plan.final_map, ranks)

    def run_circuit(self, circuit: Circuit, g: int) -> DistSvState:
        return self.run(plan(circuit, DistLayout(circuit.num_qubits, g)))
