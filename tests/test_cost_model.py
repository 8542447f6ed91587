"""Cost-model checks without a device: the CPU restatement of the planner's
perceptron is pinned by central finite differences and by an Adam step
computed by hand; the parameter layout matches the library's; plan
features are consistent with the plan they describe; the planner corpus
is seeded and within its ranges."""

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd import workloads as W
from aleph_amd.plan_tools import HostPlan
from oracle import mlp_oracle as M


def test_layout_matches_library():
    assert M.N_PARAMS == L.MODEL_PARAMS == 9473
    assert M.IN == L.MODEL_INPUTS


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, M.N_PARAMS)
    X = rng.normal(0, 1, (23, M.IN))
    y = rng.normal(0, 1, 23)
    _, g = M.loss_and_grad(p, X, y)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [M.N_PARAMS - 1, 0, M.IN * M.H]])   # b4, W1[0][0], b1[0]
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        num = (M.loss_and_grad(p + e, X, y)[0] - M.loss_and_grad(p - e, X, y)[0]) / (2 * h)
        assert abs(num - g[k]) < 1e-6 * max(1.0, abs(num)), k


def test_oracle_forward_is_the_layered_formula():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.2, M.N_PARAMS)
    X = rng.normal(0, 1, (5, M.IN))
    w = M.unpack(p)
    relu = lambda v: np.maximum(v, 0)
    want = relu(relu(relu(X @ w["W1"] + w["b1"]) @ w["W2"] + w["b2"]) @ w["W3"] + w["b3"]) @ w["w4"] + w["b4"][0]
    assert np.allclose(M.forward(p, X), want)
    assert np.allclose(M.pack(w), p)


def test_adam_step_by_hand():
    p, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    p1, m1, v1 = M.adam_step(p, g, np.zeros(2), np.zeros(2), 1, 0.1)
    # first step: m/(1-b1) = g, v/(1-b2) = g^2, so the update is lr * sign(g) (eps aside)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)


def test_plan_features_describe_the_plan():
    c = W.brickwall(20, 8, seed=1)
    hp = HostPlan(c)
    fused = [i for i, s in enumerate(hp.steps) if s.kind == 0]
    assert list(hp.feature_launch) == fused
    assert hp.features.shape == (len(fused), L.MODEL_INPUTS)
    n2q = sum(1 for op in hp.ops if op.type == 2)
    assert round(float(hp.features[:, 5].sum() * 32)) == n2q
    assert np.all(hp.features[:, 0] == 0.0) and np.allclose(hp.features[:, 1], 1.0)   # slot 0 unused
    q = HostPlan(W.qft(12), precision=32)
    assert np.all(q.features[:, 1] == 0.0) and q.features[:, 12].max() == 1.0


def test_planner_corpus_is_seeded_and_in_range():
    a = W.planner_corpus(300, 7)
    assert a == W.planner_corpus(300, 7) and a != W.planner_corpus(300, 8)
    assert {f for f, *_ in a} == {"brickwall", "qft", "hea"}
    assert all(24 <= n <= 30 and 20 <= d <= 200 for _, n, d, _ in a)
    c = W.corpus_circuit(("hea", 10, 40, 3))
    assert c.num_qubits == 10
