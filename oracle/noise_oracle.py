"""CPU restatement of the reference's noise trajectories and an exact
density-matrix evolution of the same noisy circuit — TEST INFRASTRUCTURE
ONLY (used by tests/; the HIP product never imports it).

* `trajectory_outcome` follows crates/aleph-sv/src/noise/execution.rs:102-135
  (run_trajectory) and apply.rs (apply_pauli_channel :46-66, apply_kraus_1q
  :68-133, apply_readout :135-158), with the terminal draw of
  sampling.rs:61-91.  Gates and diagonal phases go through the pinned C
  oracle (sv_oracle.cpp); channels act on the amplitude vector here.  The
  random stream is the one the library's kernel uses: draw c of shot s is
  the c-th output of splitmix64 seeded with the reference's shot seed
  (apply.rs:160-180), as a 53-bit uniform.
* `exact_distribution` evolves the density matrix: rho -> U rho U^dagger
  per gate, rho -> sum_b p_b P_b rho P_b per Pauli channel,
  rho -> sum_i K_i rho K_i^dagger per Kraus channel, then folds per-qubit
  readout confusion matrices into diag(rho).  It is pinned against the
  reference's Aer fixtures (oracle/noise/*.json, copied as data under
  tests/golden/noise/).
"""

from __future__ import annotations

import numpy as np

from . import oracle as orc

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
PAULI_M = {"I": np.eye(2, dtype=np.complex128),
           "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
           "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
           "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128)}


def mix64(z: int) -> int:
    z &= MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def shot_seed(seed: int, shot: int) -> int:
    """apply.rs:169-180."""
    return mix64((seed * GOLDEN + shot + GOLDEN) & MASK)


def uniform(key: int, c: int) -> float:
    return (mix64((key + (c + 1) * GOLDEN) & MASK) >> 11) * (1.0 / 9007199254740992.0)


def _apply_local(amps: np.ndarray, n: int, m: np.ndarray, q: int) -> np.ndarray:
    """2x2 matrix on qubit q of a 2^n vector (bit q of the index)."""
    t = amps.reshape(1 << (n - q - 1), 2, 1 << q)
    return np.einsum("ab,ibj->iaj", m, t).reshape(-1)


def _pick(weights, u: float) -> int:
    acc = 0.0
    for i, w in enumerate(weights):
        acc += w
        if u < acc:
            return i
    return len(weights) - 1


def trajectory_outcome(program, attached, readout: dict, n: int, seed: int, shot: int) -> int:
    """Outcome of absolute shot `shot` (attached: (record, channel, qubits))."""
    o = orc.OracleSv()
    st = o.allocate(n)
    key = shot_seed(seed, shot)
    draw = 0
    by_record: dict = {}
    for rec, err, qubits in attached:
        by_record.setdefault(rec, []).append((err, qubits))
    for i, inst in enumerate(program.instructions):
        kind = type(inst).__name__
        if kind == "GateInstance":
            o.apply_gate(st, inst)
        elif kind == "DiagonalPhase":
            o.apply_diagonal_phase(st, inst)
        for err, qubits in by_record.get(i, []):
            amps = st.amplitudes()
            u = uniform(key, draw)
            draw += 1
            if type(err).__name__ == "PauliChannel":
                b = _pick([p for p, _ in err.terms], u)
                for label, q in zip(err.terms[b][1], qubits):
                    if label != "I":
                        amps = _apply_local(amps, n, PAULI_M[label], q)
            else:
                probs = []
                for k in err.kraus:
                    out = _apply_local(amps, n, np.asarray(k, dtype=np.complex128), qubits[0])
                    probs.append(float(np.sum(np.abs(out) ** 2)))
                b = _pick(probs, u)
                if probs[b] >= 1e-300:
                    amps = _apply_local(amps, n, np.asarray(err.kraus[b], dtype=np.complex128), qubits[0]) \
                        / np.sqrt(probs[b])
            st.load(amps)
    p = np.abs(st.amplitudes()) ** 2
    target = uniform(key, draw) * float(np.sum(p))
    draw += 1
    idx = None
    cum = 0.0
    pos = np.flatnonzero(p > 0.0)
    for j in pos:
        cum += p[j]
        if target < cum:
            idx = int(j)
            break
    if idx is None:
        idx = int(pos[-1])
    for q in range(n):
        ro = readout.get(q)
        if ro is None:
            continue
        bit = (idx >> q) & 1
        if uniform(key, draw) < ro.m[bit][1 - bit]:
            idx ^= 1 << q
        draw += 1
    return idx


def trajectory_outcomes(program, attached, readout: dict, n: int, seed: int, shots: int, shot0: int = 0):
    return np.array([trajectory_outcome(program, attached, readout, n, seed, shot0 + s) for s in range(shots)],
                    dtype=np.uint64)


def _full(m: np.ndarray, qubits, controls, n: int) -> np.ndarray:
    """Embed a 2^k x 2^k matrix (qubits[0] = MSB of its index) with controls
    into the 2^n space by its action on basis vectors."""
    dim = 1 << n
    out = np.zeros((dim, dim), dtype=np.complex128)
    k = len(qubits)
    cmask = sum(1 << c for c in controls)
    for x in range(dim):
        if (x & cmask) != cmask:
            out[x, x] = 1.0
            continue
        j = 0
        for q in qubits:
            j = (j << 1) | ((x >> q) & 1)
        base = x
        for q in qubits:
            base &= ~(1 << q)
        for r in range(1 << k):
            y = base
            for t, q in enumerate(qubits):
                if (r >> (k - 1 - t)) & 1:
                    y |= 1 << q
            out[y, x] += m[r, j]
    return out


def exact_distribution(program, attached, readout: dict, n: int) -> np.ndarray:
    """Outcome distribution of the noisy circuit, readout errors included."""
    dim = 1 << n
    rho = np.zeros((dim, dim), dtype=np.complex128)
    rho[0, 0] = 1.0
    by_record: dict = {}
    for rec, err, qubits in attached:
        by_record.setdefault(rec, []).append((err, qubits))
    for i, inst in enumerate(program.instructions):
        kind = type(inst).__name__
        if kind == "GateInstance":
            U = _full(orc.gate_matrix(inst.gate), inst.qubits, inst.controls, n)
            rho = U @ rho @ U.conj().T
        elif kind == "DiagonalPhase":
            d = np.exp(1j * np.array([inst.phase_at(x) for x in range(dim)]))
            rho = d[:, None] * rho * d.conj()[None, :]
        for err, qubits in by_record.get(i, []):
            new = np.zeros_like(rho)
            if type(err).__name__ == "PauliChannel":
                for p, labels in err.terms:
                    P = np.eye(dim, dtype=np.complex128)
                    for label, q in zip(labels, qubits):
                        P = _full(PAULI_M[label], (q,), (), n) @ P
                    new += p * (P @ rho @ P.conj().T)
            else:
                for k in err.kraus:
                    K = _full(np.asarray(k, dtype=np.complex128), (qubits[0],), (), n)
                    new += K @ rho @ K.conj().T
            rho = new
    probs = np.real(np.diag(rho)).copy()
    for q, ro in readout.items():
        m = np.asarray(ro.m, dtype=np.float64)
        t = probs.reshape(1 << (n - q - 1), 2, 1 << q)
        probs = np.einsum("to,itj->ioj", m, t).reshape(-1)
    return probs
