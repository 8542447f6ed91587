"""The learned planner's cost model (include/aleph_hip.h, `aleph_hip_model_*`).

A perceptron (16 pass features -> 3 x 64 ReLU -> 1) predicts ln(µs) of
one fused-pass launch, scaled to a 2^24-amplitude state (the planner
multiplies by 2^(n-24), so the model carries to qubit counts it was not
trained at).  Training data are the launch times the engine
records for candidate plans (`aleph_hip_plan_execute_timed`) paired with
the features of the same launches (`aleph_hip_plan_features`); training,
inference and the gradient run as the library's own HIP kernels.  With a
model set on the backend, `compile(..., planner=PLANNER_LEARNED)` plans
every candidate (chunk qubits x pass budget x 4x4 fusion) and keeps the
one with the smallest predicted time; without a model the rule planner
plans.  `python -m aleph_amd.cost_model --out PATH` collects timings on
the device, trains, evaluates on held-out circuits and writes the model.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib as L
from . import workloads as W
from .backend import HipSvBackend, _ptr

DEFAULT_PATH = Path(__file__).resolve().parent / "cost_model.json"
REF_QUBITS = 24   # launch times are learned per 2^24 amplitudes (kCostRefQubits in the library)
CANDIDATES = [dict(local_qubits=m, max_ops=b, fuse_2q=f) for m in (10, 11, 12) for b in (12, 24, 48)
              for f in (True, False)]


class CostModel:
    def __init__(self, backend: HipSvBackend, params=None, seed: int = 0):
        self.backend = backend
        h = C.c_void_p()
        p = None
        if params is not None:
            p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
            if p.size != L.MODEL_PARAMS:
                raise ValueError(f"expected {L.MODEL_PARAMS} parameters")
        backend._check(L.load().aleph_hip_model_create(backend._ctx, _ptr(p, C.c_float) if p is not None else None,
                                                       seed, C.byref(h)))
        self.handle = h
        self.meta = {}

    def params(self) -> np.ndarray:
        out = np.zeros(L.MODEL_PARAMS, dtype=np.float32)
        self.backend._check(L.load().aleph_hip_model_get_params(self.backend._ctx, self.handle, _ptr(out, C.c_float)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32))
        self.backend._check(L.load().aleph_hip_model_set_params(self.backend._ctx, self.handle, _ptr(p, C.c_float)))

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float32).reshape(-1, L.MODEL_INPUTS))
        out = np.zeros(X.shape[0], dtype=np.float32)
        self.backend._check(L.load().aleph_hip_model_predict(self.backend._ctx, self.handle, _ptr(X, C.c_float),
                                                             X.shape[0], _ptr(out, C.c_float)))
        return out

    def gradient(self, X, y):
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float32).reshape(-1, L.MODEL_INPUTS))
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float32))
        g = np.zeros(L.MODEL_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self.backend._check(L.load().aleph_hip_model_gradient(self.backend._ctx, self.handle, _ptr(X, C.c_float),
                                                              _ptr(y, C.c_float), X.shape[0], _ptr(g, C.c_float),
                                                              C.byref(loss)))
        return float(loss.value), g

    def train(self, X, y, epochs: int = 200, batch: int = 64, lr: float = 1e-3, seed: int = 0) -> float:
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float32).reshape(-1, L.MODEL_INPUTS))
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float32))
        loss = C.c_float()
        self.backend._check(L.load().aleph_hip_model_train(self.backend._ctx, self.handle, _ptr(X, C.c_float),
                                                           _ptr(y, C.c_float), X.shape[0], epochs, batch, lr, seed,
                                                           C.byref(loss)))
        return float(loss.value)

    def install(self) -> None:
        """Make this model the one LEARNED plans of the backend use."""
        self.backend._check(L.load().aleph_hip_ctx_set_cost_model(self.backend._ctx, self.handle))

    def save(self, path) -> None:
        Path(path).write_text(json.dumps({"format": "aleph-hip cost model v1", "inputs": L.MODEL_INPUTS,
                                          "hidden": [64, 64, 64], "target": "ln(microseconds per fused launch on 2^24 amplitudes)",
                                          "meta": self.meta, "params": [float(v) for v in self.params()]}))

    @classmethod
    def load(cls, backend: HipSvBackend, path=DEFAULT_PATH) -> "CostModel":
        d = json.loads(Path(path).read_text())
        m = cls(backend, params=np.array(d["params"], dtype=np.float32))
        m.meta = d.get("meta", {})
        return m

    def free(self) -> None:
        if self.handle:
            L.load().aleph_hip_model_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def plan_features(plan) -> tuple[np.ndarray, np.ndarray]:
    """(rows x 16 features, launch index of each row) of a compiled plan."""
    lib = L.load()
    n = C.c_uint32()
    lib.aleph_hip_plan_features(plan.handle, None, None, 0, C.byref(n))
    X = np.zeros((n.value, L.MODEL_INPUTS), dtype=np.float32)
    idx = np.zeros(n.value, dtype=np.uint32)
    lib.aleph_hip_plan_features(plan.handle, _ptr(X, C.c_float), _ptr(idx, C.c_uint32), n.value, C.byref(n))
    return X, idx


def record(backend: HipSvBackend, circuit, state, reps: int = 2, **plan_kw):
    """Features and measured ln(µs) of every fused launch of one plan, plus
    the plan's total measured ms (best of `reps` event-timed executions)."""
    plan = backend.compile(circuit, **plan_kw)
    X, idx = plan_features(plan)
    best = None
    for _ in range(reps):
        ms, _ = plan.execute_timed(state)
        best = ms if best is None else np.minimum(best, ms)
    plan.free()
    # Target: ln(µs) scaled to 2^REF_QUBITS amplitudes (the planner scales back by 2^(n - REF_QUBITS)).
    y = (np.log(np.maximum(best[idx], 1e-6) * 1e3) - (circuit.num_qubits - REF_QUBITS) * math.log(2.0))
    y = y.astype(np.float32)
    return X, y, float(best.sum())


def collect(backends, specs, candidates=CANDIDATES, reps: int = 2, log=print, jit: bool = False):
    """Timings of every candidate plan of every circuit; circuit k runs on
    backends[k % len(backends)] (e.g. alternating FP64 and FP32).  With
    `jit`, fused passes run as per-pass generated kernels."""
    Xs, ys = [], []
    states = {}
    for k, spec in enumerate(specs):
        backend = backends[k % len(backends)]
        c = W.corpus_circuit(spec)
        key = (id(backend), c.num_qubits)
        if key not in states:
            states[key] = backend.allocate(c.num_qubits)
        for cand in candidates:
            X, y, _ = record(backend, c, states[key], reps=reps, jit=jit, **cand)
            Xs.append(X)
            ys.append(y)
        if log:
            log(f"collect: {k + 1}/{len(specs)} circuits, {sum(len(v) for v in ys)} launches")
    for st in states.values():
        st.free()
    return np.concatenate(Xs), np.concatenate(ys)


def evaluate(backends, models, specs, reps: int = 2, log=print, jit: bool = False):
    """Measured ms of the rule plan and of the learned plan per held-out
    circuit (circuit k on backends[k % len(backends)] with models[...])."""
    for m in models:
        m.install()
    rows = []
    states = {}
    for k, spec in enumerate(specs):
        backend = backends[k % len(backends)]
        c = W.corpus_circuit(spec)
        key = (id(backend), c.num_qubits)
        if key not in states:
            states[key] = backend.allocate(c.num_qubits)
        _, _, rule_ms = record(backend, c, states[key], reps=reps, jit=jit)
        plan = backend.compile(c, planner=L.PLANNER_LEARNED, jit=jit)
        st = plan.stats()
        best = None
        for _ in range(reps):
            ms, _ = plan.execute_timed(states[key])
            best = ms.sum() if best is None else min(best, ms.sum())
        plan.free()
        rows.append({"spec": list(spec), "precision": backend.precision, "rule_ms": rule_ms,
                     "learned_ms": float(best),
                     "chosen": [st["chosen_local_qubits"], st["chosen_max_ops"], st["chosen_flags"]],
                     "predicted_ms": st["predicted_us"] / 1e3, "plan_ms": st["plan_ms"]})
        if log:
            log(f"eval {spec} p{backend.precision}: rule {rule_ms:.2f} ms, learned {best:.2f} ms, "
                f"chose {rows[-1]['chosen']}")
    for s in states.values():
        s.free()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned planner's cost model on this device")
    ap.add_argument("--out", default=str(DEFAULT_PATH))
    ap.add_argument("--train", type=int, default=80, help="training circuits (alternating FP64 / FP32)")
    ap.add_argument("--held-out", type=int, default=16, help="evaluation circuits")
    ap.add_argument("--n-min", type=int, default=22)
    ap.add_argument("--n-max", type=int, default=25)
    ap.add_argument("--d-max", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--jit", action="store_true", help="time per-pass generated kernels (the default code path)")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    backends = [HipSvBackend(seed=1, precision=64), HipSvBackend(seed=1, precision=32)]
    specs = W.planner_corpus(args.train + args.held_out, args.seed, n_min=args.n_min, n_max=args.n_max,
                             d_max=args.d_max)
    train, held = specs[:args.train], specs[args.train:]
    t0 = time.time()
    X, y = collect(backends, train, log=log, jit=args.jit)
    log(f"collected {len(y)} launches in {time.time() - t0:.1f} s")
    model = CostModel(backends[0], seed=args.seed)
    t0 = time.time()
    loss = float("nan")
    for e in range(0, args.epochs, 10):
        loss = model.train(X, y, epochs=min(10, args.epochs - e), batch=args.batch, lr=args.lr, seed=args.seed + e)
        log(f"train: epoch {e + 10}/{args.epochs}, mse(ln us) {loss:.4f}")
    log(f"trained on the device in {time.time() - t0:.1f} s, mse(ln us) {loss:.4f}")
    twin = CostModel(backends[1], params=model.params())
    rows = evaluate(backends, [model, twin], held, log=log, jit=args.jit)
    gain = float(np.exp(np.mean([math.log(r["rule_ms"] / r["learned_ms"]) for r in rows])))
    model.meta = {"train_circuits": len(train), "launches": int(len(y)), "train_mse": loss,
                  "held_out": rows, "geomean_speedup_vs_rule": gain,
                  "corpus": {"seed": args.seed, "n": [args.n_min, args.n_max], "depth": [20, args.d_max],
                             "precisions": [64, 32]},
                  "codegen": "generated" if args.jit else "compiled"}
    model.save(args.out)
    print(json.dumps({"launches": int(len(y)), "train_mse": loss, "geomean_speedup_vs_rule": gain}))


if __name__ == "__main__":
    main()
