"""The learned planner's cost model (include/aleph_hip.h, `aleph_hip_model_*`).

A perceptron (16 pass features -> 3 x 64 ReLU -> 1) predicts ln(µs) of
one fused-pass launch, scaled to a 2^24-amplitude state (the planner
multiplies by 2^(n-24)); one feature is the state's size, since a state
that stays in the 256 MiB Infinity Cache between passes streams faster.
Training data are the launch times the engine records for candidate plans
(`aleph_hip_plan_execute_timed`) paired with the features of the same
launches (`aleph_hip_plan_features`); training, inference and the gradient
run as the library's own HIP kernels.  With a model set on the backend,
`compile(..., planner=PLANNER_LEARNED)` plans every candidate (chunk qubits
x pass budget x 4x4 fusion x non-temporal streaming, dense blocks of width
3..6, generated or compiled pass kernels) and keeps the one with the
smallest predicted time when it beats the rule planner's plan by at least
1 %; without a model the rule planner plans.

    python -m aleph_amd.cost_model collect --slice 0:75 --data a.npz
    python -m aleph_amd.cost_model collect --slice 75:150 --data b.npz
    python -m aleph_amd.cost_model train --data a.npz b.npz --out cost_model.json --minutes 10
    python -m aleph_amd.cost_model evaluate --model cost_model.json --slice 20:40 --out rows.json
    python -m aleph_amd.cost_model merge --model cost_model.json --data rows.json --out cost_model.json

times a seeded sample of candidates of each training circuit on the
device, trains on the device, times the rule and the learned plan of each
held-out circuit and writes the model with those measurements.
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
CANDIDATES = [dict(local_qubits=m, max_ops=b, fuse_2q=f, nontemporal=nt) for m in (10, 11, 12)
              for b in (12, 24, 48) for f in (True, False) for nt in (False, True)]


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


RULE = dict(local_qubits=12, max_ops=24, fuse_2q=True, nontemporal=False)


def sample_candidates(rng: np.random.Generator, generated: int = 7, compiled: int = 3):
    """The candidates timed for one corpus circuit: the rule planner's own
    options plus a seeded sample of the others, as generated kernels, and a
    few (the rule's among them) as the compiled op loop, so the model sees
    both code variants the learned planner chooses between."""
    others = [c for c in CANDIDATES if c != RULE]
    pick = [others[i] for i in rng.choice(len(others), size=generated - 1, replace=False)]
    out = [dict(RULE, jit=True)] + [dict(c, jit=True) for c in pick]
    comp = [RULE] + [others[i] for i in rng.choice(len(others), size=compiled - 1, replace=False)]
    return out + [dict(c, jit=False) for c in comp]


def collect(backends, specs, reps: int = 2, log=print, seed: int = 0, checkpoint=None, deadline=None,
            generated: int = 7, compiled: int = 3):
    """Timings of a seeded sample of candidate plans (sample_candidates) of
    every circuit; circuit k runs on backends[k % len(backends)] (e.g.
    alternating FP64 and FP32).  With `checkpoint`, the rows so far are
    saved after every circuit; collection stops before `deadline`
    (time.time() value) and returns what it has."""
    Xs, ys, done = [], [], 0
    states = {}
    rng = np.random.default_rng(seed)
    for k, spec in enumerate(specs):
        if deadline is not None and time.time() > deadline:
            log(f"collect: deadline reached after {done} circuits")
            break
        backend = backends[k % len(backends)]
        c = W.corpus_circuit(spec)
        key = (id(backend), c.num_qubits)
        if key not in states:
            for s in states.values():
                s.free()
            states = {key: backend.allocate(c.num_qubits)}
        t0 = time.time()
        for cand in sample_candidates(rng, generated, compiled):
            X, y, _ = record(backend, c, states[key], reps=reps, **cand)
            Xs.append(X)
            ys.append(y)
        done += 1
        if checkpoint:
            np.savez(checkpoint, X=np.concatenate(Xs), y=np.concatenate(ys), circuits=done)
        if log:
            log(f"collect: {k + 1}/{len(specs)} circuits ({spec[0]} n={spec[1]} p{backend.precision}, "
                f"{time.time() - t0:.1f} s), {sum(len(v) for v in ys)} launches")
    for st in states.values():
        st.free()
    if not ys:
        return np.zeros((0, L.MODEL_INPUTS), np.float32), np.zeros(0, np.float32)
    return np.concatenate(Xs), np.concatenate(ys)


def evaluate(backends, models, specs, reps: int = 2, log=print, jit: bool = True):
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
            for s in states.values():
                s.free()
            states = {key: backend.allocate(c.num_qubits)}
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
                     "chosen": [st["chosen_local_qubits"], st["chosen_max_ops"], st["chosen_flags"],
                                st["chosen_kq_width"]],
                     "kept_rule": bool(st["chosen_rule"]),
                     "predicted_ms": st["predicted_us"] / 1e3, "rule_predicted_ms": st["rule_predicted_us"] / 1e3,
                     "plan_ms": st["plan_ms"]})
        if log:
            log(f"eval {k + 1}/{len(specs)} {spec} p{backend.precision}: rule {rule_ms:.2f} ms, "
                f"learned {best:.2f} ms, chose {rows[-1]['chosen']}")
    for s in states.values():
        s.free()
    return rows


def corpus_split(train: int, held_out: int, seed: int, n_min: int, n_max: int, d_max: int):
    specs = W.planner_corpus(train + held_out, seed, n_min=n_min, n_max=n_max, d_max=d_max)
    return specs[:train], specs[train:]


def _rows_gain(rows) -> float:
    return float(np.exp(np.mean([math.log(r["rule_ms"] / r["learned_ms"]) for r in rows]))) if rows else float("nan")


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned planner's cost model on this device")
    ap.add_argument("command", choices=["collect", "train", "evaluate", "merge"],
                    help="collect: time candidate plans of a slice of the training corpus into --data; "
                         "train: train on --data files, write --out, evaluate held-out circuits until --minutes; "
                         "evaluate: time held-out circuits [--slice] with the model --model into --out (JSON rows); "
                         "merge: add evaluated rows files (--data) to the model --model (no device needed)")
    ap.add_argument("--out", default=str(DEFAULT_PATH))
    ap.add_argument("--model", default=str(DEFAULT_PATH))
    ap.add_argument("--data", nargs="*", default=[], help="collected .npz files (collect writes the first), "
                                                         "or evaluated rows files for merge")
    ap.add_argument("--train", type=int, default=150, help="training circuits (alternating FP64 / FP32)")
    ap.add_argument("--held-out", type=int, default=40, help="evaluation circuits")
    ap.add_argument("--slice", default="0:", help="the circuits [a:b] of the training (collect) or held-out "
                                                  "(evaluate) corpus")
    ap.add_argument("--n-min", type=int, default=24)
    ap.add_argument("--n-max", type=int, default=30)
    ap.add_argument("--d-max", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--minutes", type=float, default=0.0, help="stop collecting / evaluating after this long")
    ap.add_argument("--generated", type=int, default=3, help="candidates timed per circuit as generated kernels")
    ap.add_argument("--compiled", type=int, default=3, help="candidates timed per circuit as the compiled op loop")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    train, held = corpus_split(args.train, args.held_out, args.seed, args.n_min, args.n_max, args.d_max)
    a, b = (args.slice.split(":") + [""])[:2]
    deadline = time.time() + args.minutes * 60 if args.minutes > 0 else None

    if args.command == "merge":
        d = json.loads(Path(args.model).read_text())
        rows = list(d["meta"].get("held_out", []))
        for f in args.data:
            rows += json.loads(Path(f).read_text())
        seen, uniq = set(), []
        for r in rows:
            key = json.dumps(r["spec"])
            if key not in seen:
                seen.add(key)
                uniq.append(r)
        d["meta"]["held_out"] = uniq
        d["meta"]["held_out_evaluated"] = len(uniq)
        d["meta"]["geomean_speedup_vs_rule"] = _rows_gain(uniq)
        Path(args.out).write_text(json.dumps(d))
        print(json.dumps({"held_out_evaluated": len(uniq), "geomean_speedup_vs_rule": _rows_gain(uniq)}))
        return

    backends = [HipSvBackend(seed=1, precision=64), HipSvBackend(seed=1, precision=32)]
    if args.command == "collect":
        lo, hi = int(a or 0), int(b or len(train))
        out = args.data[0] if args.data else None
        t0 = time.time()
        # Backend alternation keeps the corpus index: circuit k is FP64 when k is even.
        shifted = backends if lo % 2 == 0 else backends[::-1]
        X, y = collect(shifted, train[lo:hi], log=log, seed=args.seed + lo, checkpoint=out, deadline=deadline,
                       generated=args.generated, compiled=args.compiled)
        log(f"collected {len(y)} launches in {time.time() - t0:.1f} s")
        print(json.dumps({"launches": int(len(y)), "slice": [lo, hi], "data": out}))
        return

    if args.command == "evaluate":
        lo, hi = int(a or 0), int(b or len(held))
        model = CostModel.load(backends[0], args.model)
        twin = CostModel(backends[1], params=model.params())
        shifted = backends if lo % 2 == 0 else backends[::-1]
        rows = []
        for k in range(lo, hi):
            if deadline is not None and time.time() > deadline:
                log(f"evaluate: deadline reached after {len(rows)} circuits")
                break
            rows += evaluate(shifted[(k - lo) % 2:] + shifted[:(k - lo) % 2], [model, twin], [held[k]], log=log)
            Path(args.out).write_text(json.dumps(rows))
        print(json.dumps({"evaluated": len(rows), "geomean_speedup_vs_rule": _rows_gain(rows)}))
        return

    Xs, ys, circuits = [], [], 0
    for f in args.data:
        d = np.load(f)
        Xs.append(d["X"])
        ys.append(d["y"])
        circuits += int(d["circuits"]) if "circuits" in d else 0
    X, y = np.concatenate(Xs), np.concatenate(ys)
    model = CostModel(backends[0], seed=args.seed)
    t0 = time.time()
    loss = float("nan")
    for e in range(0, args.epochs, 10):
        loss = model.train(X, y, epochs=min(10, args.epochs - e), batch=args.batch, lr=args.lr, seed=args.seed + e)
        log(f"train: epoch {min(e + 10, args.epochs)}/{args.epochs}, mse(ln us) {loss:.4f}")
    log(f"trained on the device in {time.time() - t0:.1f} s, mse(ln us) {loss:.4f}")
    model.meta = {"train_circuits": circuits, "launches": int(len(y)), "train_mse": loss,
                  "held_out": [], "held_out_evaluated": 0, "geomean_speedup_vs_rule": float("nan"),
                  "corpus": {"seed": args.seed, "n": [args.n_min, args.n_max], "depth": [20, args.d_max],
                             "precisions": [64, 32], "held_out_circuits": args.held_out,
                             "candidates_per_circuit": "the rule options plus a seeded sample, as generated kernels "
                                                       "and as the compiled op loop"},
                  "codegen": "generated"}
    model.save(args.out)
    twin = CostModel(backends[1], params=model.params())
    rows = []
    for k in range(len(held)):
        if deadline is not None and time.time() > deadline:
            log(f"evaluate: deadline reached after {len(rows)} circuits")
            break
        rows += evaluate(backends[k % 2:] + backends[:k % 2], [model, twin], [held[k]], log=log)
        model.meta.update({"held_out": rows, "held_out_evaluated": len(rows),
                           "geomean_speedup_vs_rule": _rows_gain(rows)})
        model.save(args.out)
    print(json.dumps({"launches": int(len(y)), "train_mse": loss, "held_out_evaluated": len(rows),
                      "geomean_speedup_vs_rule": _rows_gain(rows)}))


if __name__ == "__main__":
    main()
