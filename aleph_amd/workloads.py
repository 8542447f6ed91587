"""Seeded workload builders and an OpenQASM 3 subset reader.

The benchmark workloads are defined by their parameters (qubits, depth,
seed), not by stored files:

* `brickwall(n, depth, seed)`: every layer applies U3(θ, φ, λ) to every
  qubit with θ, φ, λ uniform in [0, 2π) (SplitMix64 stream, three draws per
  gate in layer-major, qubit-major order), then CZ on (0,1),(2,3),... in
  even layers and (1,2),(3,4),... in odd layers.
* `qft(n)`: H on qubit j, then controlled-phase 2π/2^(k−j+1) with control k
  onto target j for every k > j, then the qubit-reversal swaps — the
  reference's `qft_circuit` (benches/src/lib.rs:66-86) plus the swaps.
* `gaussian_state(n, seed)`: a normalised complex Gaussian vector.
* `hea(n, layers, seed)`: RY and RZ on every qubit, then a CX ladder.
* `qaoa_maxcut(n, edges, gammas, betas)`: H^n, then per layer RZZ(γ) on
  every edge (as CNOT·RZ(γ)·CNOT = exp(−iγ Z⊗Z/2)) and RX(β) on every qubit.
* `qaoa_batch(count, seed)`: seeded QAOA MaxCut instances with n even in
  [12, 24], random 3-regular graphs and p = 4 layers of angles in [0, π).
"""

from __future__ import annotations

import ast
import math
import operator
import re
from pathlib import Path

import numpy as np

from .circuit import Circuit, Gate

MASK64 = (1 << 64) - 1
TWO_PI = 2.0 * math.pi


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014) with 53-bit uniform doubles."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def brickwall(n: int, depth: int, seed: int = 0) -> Circuit:
    rng = SplitMix64(seed)
    c = Circuit(n)
    for layer in range(depth):
        for q in range(n):
            t, p, l = TWO_PI * rng.uniform(), TWO_PI * rng.uniform(), TWO_PI * rng.uniform()
            c.u3(t, p, l, q)
        q = layer & 1
        while q + 1 < n:
            c.cz(q, q + 1)
            q += 2
    return c


def qft(n: int, swaps: bool = True) -> Circuit:
    c = Circuit(n)
    for j in range(n):
        c.h(j)
        for k in range(j + 1, n):
            c.cp(TWO_PI / float(1 << (k - j + 1)), k, j)
    if swaps:
        for j in range(n // 2):
            c.swap(j, n - 1 - j)
    return c


def gaussian_state(n: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def hea(n: int, layers: int, seed: int = 0) -> Circuit:
    rng = SplitMix64(seed)
    c = Circuit(n)
    for _ in range(layers):
        for q in range(n):
            c.ry(TWO_PI * rng.uniform(), q)
            c.rz(TWO_PI * rng.uniform(), q)
        for q in range(n - 1):
            c.cx(q, q + 1)
    return c


def random_3_regular(n: int, seed: int) -> list[tuple[int, int]]:
    """Seeded random simple 3-regular graph by the pairing model with rejection."""
    if n % 2 or n < 4:
        raise ValueError("3-regular graphs need an even n >= 4")
    rng = SplitMix64(seed)
    while True:
        stubs = [v for v in range(n) for _ in range(3)]
        for i in range(len(stubs) - 1, 0, -1):
            j = rng.next_u64() % (i + 1)
            stubs[i], stubs[j] = stubs[j], stubs[i]
        edges = set()
        ok = True
        for a, b in zip(stubs[0::2], stubs[1::2]):
            e = (min(a, b), max(a, b))
            if a == b or e in edges:
                ok = False
                break
            edges.add(e)
        if ok:
            return sorted(edges)


def qaoa_maxcut(n: int, edges, gammas, betas) -> Circuit:
    """QAOA MaxCut ansatz: H^n, then per layer RZZ(γ) = exp(−iγ Z⊗Z/2) on every
    edge (as CNOT · Rz(γ) · CNOT) and RX(β) on every qubit.  The reference's
    `build_qaoa` (crates/aleph-ir/src/ansatz.rs) uses Rz(2γ), i.e. its γ is
    half of the one here."""
    c = Circuit(n)
    for q in range(n):
        c.h(q)
    for g, b in zip(gammas, betas):
        for (u, v) in edges:
            c.cx(u, v)
            c.rz(g, v)
            c.cx(u, v)
        for q in range(n):
            c.rx(b, q)
    return c


def qaoa_batch(count: int, seed: int, n_min: int = 12, n_max: int = 24, p: int = 4):
    """Seeded batch of QAOA MaxCut instances: n uniform over the even values in
    [n_min, n_max], a random 3-regular graph per instance, and gamma, beta
    uniform in [0, pi) per layer.  Returns (graphs, gammas, betas) with graphs a
    list of (n, edges) and the angle arrays of shape (count, p)."""
    rng = SplitMix64(seed)
    evens = list(range(n_min + (n_min & 1), n_max + 1, 2))
    graphs, gammas, betas = [], np.zeros((count, p)), np.zeros((count, p))
    for i in range(count):
        n = evens[rng.next_u64() % len(evens)]
        graphs.append((n, random_3_regular(n, rng.next_u64())))
        for l in range(p):
            gammas[i, l] = math.pi * rng.uniform()
            betas[i, l] = math.pi * rng.uniform()
    return graphs, gammas, betas


def planner_corpus(count: int, seed: int, n_min: int = 24, n_max: int = 30, d_min: int = 20, d_max: int = 200):
    """Seeded corpus for the learned planner: n uniform in [n_min, n_max],
    depth uniform in [d_min, d_max], family cycling brickwall / QFT / HEA
    (QFT ignores the depth).  Returns a list of (family, n, depth, seed)
    specs; `corpus_circuit(spec)` builds one."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        n = n_min + int(rng.next_u64() % (n_max - n_min + 1))
        d = d_min + int(rng.next_u64() % (d_max - d_min + 1))
        fam = ("brickwall", "qft", "hea")[int(rng.next_u64() % 3)]
        out.append((fam, n, d, int(rng.next_u64() & 0xFFFFFFFF)))
    return out


def corpus_circuit(spec) -> Circuit:
    fam, n, d, s = spec
    if fam == "brickwall":
        return brickwall(n, d, seed=s)
    if fam == "qft":
        return qft(n)
    return hea(n, max(1, d // 4), seed=s)


# ---------------- OpenQASM 3 subset ----------------

_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
        ast.Pow: operator.pow, ast.USub: operator.neg, ast.UAdd: operator.pos}


def _eval(expr: str) -> float:
    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name) and node.id in ("pi", "π"):
            return math.pi
        if isinstance(node, ast.Name) and node.id in ("tau", "τ"):
            return TWO_PI
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.UnaryOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.operand))
        raise ValueError(f"unsupported expression {expr!r}")
    return ev(ast.parse(expr.replace("π", "pi"), mode="eval"))


_QASM_GATES = {
    "h": ("H", 0), "x": ("X", 0), "y": ("Y", 0), "z": ("Z", 0), "s": ("S", 0), "sdg": ("Sdg", 0),
    "t": ("T", 0), "tdg": ("Tdg", 0), "rx": ("Rx", 1), "ry": ("Ry", 1), "rz": ("Rz", 1), "p": ("Phase", 1),
    "phase": ("Phase", 1), "u3": ("U3", 3), "u": ("U3", 3), "cx": ("Cnot", 0), "cnot": ("Cnot", 0),
    "cz": ("Cz", 0), "swap": ("Swap", 0), "ccx": ("Toffoli", 0), "ccz": ("Ccz", 0),
    "crx": ("CRx", 1), "cry": ("CRy", 1), "crz": ("CRz", 1), "id": (None, 0),
}


def parse_qasm(src: str) -> Circuit:
    """Read the gate subset used by the reference's fixture and corpus files.
    `measure` statements are recorded as ("measure", qubit, clbit) tuples."""
    n = None
    circuit = None
    text = re.sub(r"//[^\n]*", "", src)
    for stmt in text.split(";"):
        s = stmt.strip()
        if not s or s.startswith("OPENQASM") or s.startswith("include"):
            continue
        m = re.match(r"qubit\s*\[\s*(\d+)\s*\]\s*(\w+)$", s) or re.match(r"qreg\s+(\w+)\s*\[\s*(\d+)\s*\]$", s)
        if m:
            n = int(m.group(1)) if s.startswith("qubit") else int(m.group(2))
            circuit = Circuit(n)
            continue
        if s.startswith("bit") or s.startswith("creg"):
            continue
        if circuit is None:
            raise ValueError("gate before qubit declaration")
        if s.startswith("barrier"):
            qs = [int(x) for x in re.findall(r"\[(\d+)\]", s)] or list(range(n))
            circuit.barrier(qs)
            continue
        mm = re.match(r"(?:\w+\s*\[\s*(\d+)\s*\]\s*=\s*)?measure\s+\w+\s*\[\s*(\d+)\s*\]"
                      r"(?:\s*->\s*\w+\s*\[\s*(\d+)\s*\])?$", s)
        if mm:
            q = int(mm.group(2))
            clbit = int(mm.group(1) or mm.group(3) or q)
            circuit.add(("measure", q, clbit))
            continue
        mg = re.match(r"(\w+)\s*(?:\((.*)\))?\s+(.*)$", s, re.S)
        if not mg:
            raise ValueError(f"cannot parse statement {s!r}")
        name, args, operands = mg.group(1).lower(), mg.group(2), mg.group(3)
        if name not in _QASM_GATES:
            raise ValueError(f"unsupported gate {name!r}")
        kind, npar = _QASM_GATES[name]
        params = [_eval(a.strip()) for a in args.split(",")] if args else []
        if len(params) != npar:
            raise ValueError(f"gate {name} expects {npar} parameters")
        qubits = [int(x) for x in re.findall(r"\[\s*(\d+)\s*\]", operands)]
        if kind is None:
            continue
        circuit.gate(Gate(kind, tuple(params)), qubits)
    if circuit is None:
        raise ValueError("no qubit declaration")
    return circuit


def load_qasm(path) -> Circuit:
    return parse_qasm(Path(path).read_text())


def noisy_brickwall(n: int, depth: int, seed: int = 0):
    """The seeded brickwall with a device-style noise model: depolarizing
    1e-3 after every U3 and 1e-2 after every CZ, amplitude damping 2e-3
    after every U3 (a Kraus channel, state-dependent), 2 % / 3 % readout
    error on every qubit.  Returns (circuit, NoiseModel)."""
    from . import noise as N
    c = brickwall(n, depth, seed=seed)
    m = N.NoiseModel()
    m.add_all_qubit_quantum_error(N.depolarizing_error(1e-3, 1), ["U3"])
    m.add_all_qubit_quantum_error(N.depolarizing_error(1e-2, 2), ["Cz"])
    m.add_all_qubit_quantum_error(N.amplitude_damping_error(2e-3), ["U3"])
    for q in range(n):
        m.add_readout_error(N.ReadoutError([[0.98, 0.02], [0.03, 0.97]]), q)
    return c, m
