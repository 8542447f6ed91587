"""GPU parity of batched execution (aleph_hip_batch_*) through the C ABI.

Every instance of a batch must equal the CPU oracle run of the same circuit
(FP64 within 1e-12, FP32 within 1e-5), and <C> must equal sum_x p(x) cut(x)
from the oracle's amplitudes.  Both executors (persistent, rounds) are
covered, graph replays are checked against the eager oracle result after
angle updates, and the bounded dependency wait is forced to fail from the
host with a spin limit of one poll.
"""

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd import workloads as W
from aleph_amd.backend import DeviceFailure
from aleph_amd.batch import HipBatch, maxcut_value
from tests.helpers import GOLDEN, PARITY_TOL, max_abs, qaoa_p1_cut_expectation
from tests.test_oracle import random_circuit

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def oracle_amps(orc, circuit):
    o = orc.OracleSv()
    st = o.allocate(circuit.num_qubits)
    o.run(circuit, st)
    return st.amplitudes()


def small_qaoa(count, seed, n_min=6, n_max=16, p=3):
    return W.qaoa_batch(count, seed, n_min=n_min, n_max=n_max, p=p)


def check_qaoa_batch(orc, batch, graphs, gammas, betas, tol):
    cvals = batch.expect_maxcut(graphs)
    for i, (n, edges) in enumerate(graphs):
        want = oracle_amps(orc, W.qaoa_maxcut(n, edges, gammas[i], betas[i]))
        got = batch.state(i)
        assert max_abs(got, want) < tol, (i, n)
        c_want = maxcut_value(n, edges, np.abs(want) ** 2)
        assert abs(cvals[i] - c_want) < tol * max(1.0, len(edges)) * 10, (i, cvals[i], c_want)


@pytest.mark.parametrize("mode", [L.BATCH_PERSISTENT, L.BATCH_ROUNDS])
@pytest.mark.parametrize("precision", [64, 32])
def test_qaoa_batch_matches_oracle(hip, orc, mode, precision):
    graphs, gammas, betas = small_qaoa(14, seed=3 + mode + precision)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=precision, mode=mode, local_qubits=9,
                      chunks_per_item=2)
    b.run()
    check_qaoa_batch(orc, b, graphs, gammas, betas, PARITY_TOL if precision == 64 else FP32_TOL)
    b.free()


def test_executors_agree_bitwise(hip):
    graphs, gammas, betas = small_qaoa(20, seed=11, n_min=10, n_max=18)
    states = []
    for mode in (L.BATCH_PERSISTENT, L.BATCH_ROUNDS):
        b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=32, mode=mode, local_qubits=10, chunks_per_item=1)
        b.run()
        states.append([b.state(i) for i in range(len(graphs))])
        b.free()
    for a, c in zip(*states):
        assert np.array_equal(a, c)


def test_generic_circuit_batch(hip, orc):
    rng = np.random.default_rng(5)
    circuits = [random_circuit(int(n), 40, rng, kq=False) for n in rng.integers(5, 12, size=9)]
    circuits.append(W.brickwall(13, 6, seed=2))
    circuits.append(W.hea(11, 3, seed=9))
    b = HipBatch.circuits(hip, circuits, precision=64, local_qubits=8)
    b.run()
    for i, c in enumerate(circuits):
        assert max_abs(b.state(i), oracle_amps(orc, c)) < PARITY_TOL, i
    b.free()


def leading_hadamard_mask(c):
    """Qubits whose first instruction is an uncontrolled H."""
    from aleph_amd.circuit import DiagonalPhase, GateInstance
    touched = plus = 0
    for inst in c.instructions:
        if isinstance(inst, GateInstance):
            if inst.gate.name == "H" and not inst.controls and not (touched >> inst.qubits[0]) & 1:
                plus |= 1 << inst.qubits[0]
            for q in inst.qubits + inst.controls:
                touched |= 1 << q
        elif isinstance(inst, DiagonalPhase):
            for t in inst.terms:
                for m in t.conds:
                    touched |= m
    return plus


def test_leading_hadamards_prepared_by_the_reset(hip, orc):
    """An uncontrolled H that is the first instruction on its qubit is left
    out of the plan and the reset writes |+> there instead.  Kept as gates:
    an H after another gate on its qubit, a controlled H, and H on a qubit
    a diagonal phase touched first.  Every instance equals the oracle on
    two runs (the second a graph replay), packed and solo alike."""
    from aleph_amd.circuit import Circuit, DiagonalPhase, Gate, GateInstance, PhaseTerm
    rng = np.random.default_rng(8)
    circuits, expect = [], []
    for k in range(8):
        n = 6 + k
        c = Circuit(n)
        if k % 2:
            c.cx(0, 1)                                   # 0 and 1 touched before their H
        if k % 3 == 0:
            c.add(GateInstance.controlled(Gate.H(), [2], [3]))
        if k == 4:
            c.add(DiagonalPhase(n, [PhaseTerm((1 << 5,), 0.7)]))
        for q in range(n):
            if rng.random() < 0.7:
                c.h(q)
        body = random_circuit(n, 30, rng, kq=False)
        for inst in body.instructions:
            c.add(inst)
        expect.append(leading_hadamard_mask(c))
        circuits.append(c)
    assert all(expect) and any(m != (1 << c.num_qubits) - 1 for m, c in zip(expect, circuits))
    for solo in (0, 10):
        b = HipBatch.circuits(hip, circuits, precision=64, local_qubits=6, solo_qubits=solo)
        st = b.stats()
        assert st["prepared_instances"] == sum(1 for m in expect if m)
        assert st["prepared_qubits"] == sum(bin(m).count("1") for m in expect)
        for rep in range(2):
            b.run()
            for i, c in enumerate(circuits):
                assert max_abs(b.state(i), oracle_amps(orc, c)) < PARITY_TOL, (solo, rep, i)
        b.free()


def test_graph_replay_with_angle_updates(hip, orc):
    """The captured graph is replayed after every angle update, with eager
    device work and synchronisations in between; each replay is checked."""
    graphs, gammas, betas = small_qaoa(8, seed=21, n_min=8, n_max=14, p=2)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=64, local_qubits=8)
    rng = np.random.default_rng(0)
    for rep in range(3):
        if rep:
            gammas = rng.uniform(0, np.pi, size=gammas.shape)
            betas = rng.uniform(0, np.pi, size=betas.shape)
            b.set_qaoa_angles(gammas, betas)
        b.run()
        st = hip.allocate(6)          # unrelated eager work between replays
        hip.norm_sqr(st)
        st.free()
        hip.synchronize()
        check_qaoa_batch(orc, b, graphs, gammas, betas, PARITY_TOL)
        b.run()                       # a second replay of the same graph
        check_qaoa_batch(orc, b, graphs, gammas, betas, PARITY_TOL)
    b.free()


def test_bounded_wait_failure_is_reported(hip, orc):
    """With every dependency made unsatisfiable from the host and a spin
    limit of 16 polls, the waits give up: every workgroup still exits, the
    run reports a device failure, and the device keeps working."""
    graphs, gammas, betas = W.qaoa_batch(2, seed=5, n_min=16, n_max=16, p=2)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=32, mode=L.BATCH_PERSISTENT, local_qubits=10,
                      chunks_per_item=1, spin_limit=16, flags=L.BATCH_UNREACHABLE_DEPS)
    with pytest.raises(DeviceFailure):
        b.run()
    b.free()
    good = HipBatch.qaoa(hip, graphs, gammas, betas, precision=32, mode=L.BATCH_PERSISTENT, local_qubits=10)
    good.run()
    check_qaoa_batch(orc, good, graphs, gammas, betas, FP32_TOL)
    good.free()


def test_batch_stats_and_autotune(hip, orc):
    graphs, gammas, betas = small_qaoa(10, seed=31, n_min=10, n_max=16, p=2)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=32)
    st = b.stats()
    assert st["n_instances"] == 10 and st["n_passes"] >= 10 and st["n_items"] >= st["n_passes"]
    assert st["sweep_bytes"] > 0 and st["pool_bytes"] == 8 * sum(1 << n for n, _ in graphs)
    res = b.autotune(reps=1)
    assert len(res) == 18 and all(r["ms"] > 0 for r in res)
    best = min(res, key=lambda r: r["ms"])
    st = b.stats()
    assert (st["mode"], st["local_qubits"], st["chunks_per_item"]) == (best["mode"], best["local_qubits"],
                                                                       best["chunks_per_item"])
    b.run()
    check_qaoa_batch(orc, b, graphs, gammas, betas, FP32_TOL)
    b.free()


def test_batch_rejects_bad_input(hip):
    with pytest.raises(Exception):
        HipBatch.qaoa(hip, [(4, [(0, 1)])], [[0.1]], [[0.2]])          # n < 5
    with pytest.raises(Exception):
        HipBatch.qaoa(hip, [(6, [(0, 9)])], [[0.1]], [[0.2]])          # edge out of range


# ---------------------------------------------------------------- the QAOA workload at its own sizes


@pytest.mark.parametrize("mode", [L.BATCH_PERSISTENT, L.BATCH_ROUNDS])
def test_qaoa_p4_up_to_n20_matches_oracle(hip, orc, mode):
    """p = 4 with n even in [12, 20] (the workload's depth, and sizes the
    oracle finishes in seconds), FP32 amplitudes and FP64 <C>."""
    graphs, gammas, betas = W.qaoa_batch(10, seed=404 + mode, n_min=12, n_max=20, p=4)
    assert max(n for n, _ in graphs) == 20
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=32, mode=mode)
    b.run()
    check_qaoa_batch(orc, b, graphs, gammas, betas, FP32_TOL)
    b.free()


def test_qaoa_4096_instances_p1_closed_form(hip):
    """The whole workload shape (4096 seeded 3-regular instances, n even in
    [12, 24], FP32, one batch) with p = 1 angles: every instance's <C> equals
    the analytic p = 1 expectation (tests/helpers.qaoa_p1_cut_expectation,
    pinned against brute force in tests/test_qaoa_closed_form.py)."""
    graphs, gammas, betas = W.qaoa_batch(4096, seed=2024, p=1)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=32, mode=L.BATCH_ROUNDS)
    b.run()
    got = b.expect_maxcut(graphs)
    b.free()
    want = np.array([qaoa_p1_cut_expectation(n, e, gammas[i, 0], betas[i, 0]) for i, (n, e) in enumerate(graphs)])
    err = np.abs(got - want) / np.array([len(e) for _, e in graphs])
    print(f"4096 instances p=1: max |<C> - closed form| per edge {err.max():.2e}, mean <C> {got.mean():.4f}",
          flush=True)
    assert err.max() < 1e-5


@pytest.mark.parametrize("name", ["qaoa_n6", "qaoa_n10", "qaoa_n14"])
def test_qaoa_reference_graphs(hip, orc, name):
    """The reference's committed graphs (scripts/qaoa/graphs): p = 4 batch vs the
    oracle, and <C> never above the recorded maximum cut."""
    edges = [tuple(int(t) for t in line.split()) for line in (GOLDEN / "qaoa" / f"{name}.edges").read_text().split("\n")
             if line.strip() and not line.startswith("#")]
    n = 1 + max(max(e) for e in edges)
    maxcut = float((GOLDEN / "qaoa" / f"{name}.maxcut").read_text().split()[0])
    rng = np.random.default_rng(len(edges))
    gammas, betas = rng.uniform(0, np.pi, (1, 4)), rng.uniform(0, np.pi, (1, 4))
    if n < 5:
        pytest.skip("batches take n >= 5")
    b = HipBatch.qaoa(hip, [(n, edges)], gammas, betas, precision=64)
    b.run()
    check_qaoa_batch(orc, b, [(n, edges)], gammas, betas, PARITY_TOL)
    assert b.expect_maxcut([(n, edges)])[0] <= maxcut + 1e-9
    b.free()


def test_rounds_replay_after_a_zero_angle_changes_the_plan(hip, orc):
    """A zero gamma drops an instance's cost layer, so its pass count and the
    per-round item ranges change while the totals may not: the captured graph
    must be re-captured, and every replay must still equal the oracle."""
    graphs, gammas, betas = small_qaoa(3, seed=77, n_min=10, n_max=12, p=2)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=64, mode=L.BATCH_ROUNDS, local_qubits=8)
    b.run()
    check_qaoa_batch(orc, b, graphs, gammas, betas, PARITY_TOL)
    gammas = gammas.copy()
    gammas[0, :] = 0.0
    gammas[1, 0] = 0.0
    b.set_qaoa_angles(gammas, betas)
    for _ in range(2):
        b.run()
        check_qaoa_batch(orc, b, graphs, gammas, betas, PARITY_TOL)
    b.free()


@pytest.mark.parametrize("mode", [L.BATCH_PERSISTENT, L.BATCH_ROUNDS])
@pytest.mark.parametrize("precision", [64, 32])
def test_solo_instances_match_oracle(hip, orc, mode, precision):
    """Instances at or above solo_qubits run as their own plans (generated
    kernels) after the packed ones; every instance must still equal the
    oracle, with graph replays across angle updates."""
    graphs, gammas, betas = small_qaoa(12, seed=31 + mode + precision, n_min=8, n_max=14, p=2)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=precision, mode=mode, local_qubits=8,
                      solo_qubits=12)
    st = b.stats()
    assert st["solo_instances"] == sum(1 for n, _ in graphs if n >= 12) > 0
    tol = PARITY_TOL if precision == 64 else FP32_TOL
    rng = np.random.default_rng(1)
    for rep in range(2):
        if rep:
            gammas = rng.uniform(0, np.pi, size=gammas.shape)
            betas = rng.uniform(0, np.pi, size=betas.shape)
            b.set_qaoa_angles(gammas, betas)
        b.run()
        check_qaoa_batch(orc, b, graphs, gammas, betas, tol)
        b.run()
        check_qaoa_batch(orc, b, graphs, gammas, betas, tol)
    b.free()


def test_all_solo_batch_and_default_threshold(hip, orc):
    graphs, gammas, betas = small_qaoa(4, seed=41, n_min=10, n_max=12, p=2)
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=64, solo_qubits=5)
    assert b.stats()["solo_instances"] == 4 and b.stats()["n_items"] == 0
    b.run()
    check_qaoa_batch(orc, b, graphs, gammas, betas, PARITY_TOL)
    b.free()
    b = HipBatch.qaoa(hip, graphs, gammas, betas, precision=64)
    assert b.stats()["solo_instances"] == 0      # default threshold: 21 qubits
    b.free()
