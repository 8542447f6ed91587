"""Batched execution: many small circuits of uneven qubit count run as one
persistent launch (or one launch per pass round), captured once into a
hipGraph and replayed (include/aleph_hip.h, `aleph_hip_batch_*`).

The reference runs a variational sweep as one `run_with_outcomes` call per
circuit (crates/aleph-backend/src/lib.rs:211-248) and scores MaxCut with
`expectation_pauli_sum` over `maxcut_pauli_sum` (lib.rs:422-432,
crates/aleph-ir/src/ansatz.rs); `HipBatch.expect_maxcut` returns the same
<C> = sum_edges (1 - <Z_u Z_v>)/2 per instance, accumulated in FP64.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .backend import HipSvBackend, _ptr, _raise


def _opts(mode, local_qubits, chunks_per_item, groups_per_cu, max_ops, spin_limit, nontemporal, flags=0,
          solo_qubits=0):
    return L.BatchOpts(mode, local_qubits, chunks_per_item, groups_per_cu, max_ops, spin_limit,
                       (L.BATCH_NONTEMPORAL if nontemporal else 0) | flags, solo_qubits)


def _edge_arrays(graphs):
    offs = np.zeros(len(graphs) + 1, dtype=np.uint64)
    flat = []
    for i, (_, edges) in enumerate(graphs):
        offs[i + 1] = offs[i] + len(edges)
        for u, v in edges:
            flat += [u, v]
    return offs, np.ascontiguousarray(np.array(flat if flat else [0], dtype=np.uint32))


class HipBatch:
    """A packed batch of instances on one backend's device."""

    def __init__(self, backend: HipSvBackend, handle, num_qubits, precision: int, graphs=None):
        self.backend = backend
        self.handle = handle
        self.num_qubits = list(num_qubits)
        self.precision = precision
        self.graphs = graphs

    # ---- construction ----
    @classmethod
    def qaoa(cls, backend: HipSvBackend, graphs, gammas, betas, precision: int = L.F32, mode: int = L.BATCH_PERSISTENT,
             local_qubits: int = 0, chunks_per_item: int = 0, groups_per_cu: int = 0, max_ops: int = 0,
             spin_limit: int = 0, nontemporal: bool = False, flags: int = 0, solo_qubits: int = 0) -> "HipBatch":
        """graphs: list of (n, edges); gammas/betas: (instances, p) arrays.  Instances of at least
        `solo_qubits` qubits (0: 21, above 40: none) run as their own plans with generated kernels."""
        nq = np.ascontiguousarray(np.array([g[0] for g in graphs], dtype=np.uint32))
        offs, flat = _edge_arrays(graphs)
        ga = np.ascontiguousarray(np.asarray(gammas, dtype=np.float64).reshape(len(graphs), -1))
        be = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(len(graphs), -1))
        if ga.shape != be.shape:
            raise ValueError("gammas and betas must have the same shape")
        o = _opts(mode, local_qubits, chunks_per_item, groups_per_cu, max_ops, spin_limit, nontemporal, flags,
                  solo_qubits)
        h = C.c_void_p()
        backend._check(L.load().aleph_hip_batch_create_qaoa(
            backend._ctx, len(graphs), _ptr(nq, C.c_uint32), _ptr(offs, C.c_uint64), _ptr(flat, C.c_uint32),
            ga.shape[1], _ptr(ga, C.c_double), _ptr(be, C.c_double), precision, C.byref(o), C.byref(h)))
        return cls(backend, h, nq.tolist(), precision, graphs=list(graphs))

    @classmethod
    def circuits(cls, backend: HipSvBackend, circuits, precision: int = L.F32, mode: int = L.BATCH_PERSISTENT,
                 local_qubits: int = 0, chunks_per_item: int = 0, groups_per_cu: int = 0, max_ops: int = 0,
                 spin_limit: int = 0, nontemporal: bool = False, solo_qubits: int = 0) -> "HipBatch":
        nq = np.ascontiguousarray(np.array([c.num_qubits for c in circuits], dtype=np.uint32))
        words, reals = [], []
        woff = np.zeros(len(circuits) + 1, dtype=np.uint64)
        roff = np.zeros(len(circuits) + 1, dtype=np.uint64)
        for i, c in enumerate(circuits):
            w, r = c.encode()
            words.append(w)
            reals.append(r)
            woff[i + 1] = woff[i] + w.size
            roff[i + 1] = roff[i] + r.size
        W = np.ascontiguousarray(np.concatenate(words + [np.zeros(1, np.uint32)]).astype(np.uint32))
        R = np.ascontiguousarray(np.concatenate(reals + [np.zeros(1)]).astype(np.float64))
        o = _opts(mode, local_qubits, chunks_per_item, groups_per_cu, max_ops, spin_limit, nontemporal,
                  solo_qubits=solo_qubits)
        h = C.c_void_p()
        backend._check(L.load().aleph_hip_batch_create(
            backend._ctx, len(circuits), _ptr(nq, C.c_uint32), precision, _ptr(W, C.c_uint32),
            _ptr(woff, C.c_uint64), _ptr(R, C.c_double), _ptr(roff, C.c_uint64), C.byref(o), C.byref(h)))
        return cls(backend, h, nq.tolist(), precision)

    # ---- execution ----
    def _check(self, code):
        if code != L.OK:
            _raise(self.backend._ctx, code)

    def run(self) -> None:
        self._check(L.load().aleph_hip_batch_run(self.backend._ctx, self.handle))

    def run_timed(self) -> float:
        ms = C.c_double()
        self._check(L.load().aleph_hip_batch_run_timed(self.backend._ctx, self.handle, C.byref(ms)))
        return float(ms.value)

    def set_qaoa_angles(self, gammas, betas) -> None:
        ga = np.ascontiguousarray(np.asarray(gammas, dtype=np.float64))
        be = np.ascontiguousarray(np.asarray(betas, dtype=np.float64))
        self._check(L.load().aleph_hip_batch_set_qaoa_angles(self.backend._ctx, self.handle, _ptr(ga, C.c_double),
                                                             _ptr(be, C.c_double)))

    def expect_maxcut(self, graphs=None) -> np.ndarray:
        graphs = graphs if graphs is not None else self.graphs
        if graphs is None or len(graphs) != len(self.num_qubits):
            raise ValueError("one graph per instance is required")
        offs, flat = _edge_arrays(graphs)
        out = np.zeros(len(graphs))
        self._check(L.load().aleph_hip_batch_expect_maxcut(self.backend._ctx, self.handle, _ptr(offs, C.c_uint64),
                                                           _ptr(flat, C.c_uint32), _ptr(out, C.c_double)))
        return out

    def state(self, i: int) -> np.ndarray:
        out = np.zeros(2 << self.num_qubits[i])
        self._check(L.load().aleph_hip_batch_read_state(self.backend._ctx, self.handle, i, _ptr(out, C.c_double)))
        return out.view(np.complex128)

    def stats(self) -> dict:
        s = L.BatchStats()
        self._check(L.load().aleph_hip_batch_stats_get(self.handle, C.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}

    def autotune(self, reps: int = 2) -> list[dict]:
        buf = (L.BatchTune * 64)()
        n = C.c_uint32()
        self._check(L.load().aleph_hip_batch_autotune(self.backend._ctx, self.handle, reps, buf, 64, C.byref(n)))
        return [{name: getattr(buf[i], name) for name, _ in buf[i]._fields_} for i in range(n.value)]

    def free(self) -> None:
        if self.handle:
            L.load().aleph_hip_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def maxcut_value(n: int, edges, probs: np.ndarray) -> float:
    """sum_x p(x) cut(x) for a dense probability vector (test helper)."""
    x = np.arange(1 << n, dtype=np.int64)
    cut = np.zeros(1 << n)
    for u, v in edges:
        cut += ((x >> u) ^ (x >> v)) & 1
    return float(np.dot(probs, cut))
