"""Host-side planner checks (no device): plans built on the host are
emulated op by op (tests/plan_emulator.py) and compared with the CPU oracle
at 1e-12; structural properties of the plans are checked directly."""

import math

import numpy as np
import pytest

from aleph_amd import workloads as W
from aleph_amd.circuit import Circuit, DiagonalPhase, Gate, PhaseTerm
from aleph_amd.plan_tools import HostPlan
from tests.helpers import max_abs
from tests.plan_emulator import emulate
from tests.test_oracle import random_circuit


def oracle_state(orc, c, psi):
    o = orc.OracleSv()
    st = o.allocate(c.num_qubits)
    st.load(psi)
    o.run(c, st)
    return st.amplitudes()


OPTION_GRID = [
    {},
    {"local_qubits": 4}, {"local_qubits": 6, "coalesce_bits": 1}, {"local_qubits": 9, "max_ops": 2},
    {"fuse_1q": False, "fuse_2q": False}, {"reg_bits": 5, "local_qubits": 8}, {"max_ops": 500},
    {"planner": 0},
]


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("opt", range(len(OPTION_GRID)))
def test_emulated_plan_matches_oracle(oracle_lib, seed, opt):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 10))
    c = random_circuit(n, 60, rng)
    psi = W.gaussian_state(n, seed)
    kw = OPTION_GRID[opt]
    if kw.get("local_qubits", 0) > n:
        kw = {**kw, "local_qubits": n}
    plan = HostPlan(c, **kw)
    assert max_abs(emulate(plan, psi), oracle_state(oracle_lib, c, psi)) < 1e-12


@pytest.mark.parametrize("name", ["brickwall", "qft", "hea", "qaoa", "phase"])
def test_emulated_workloads(oracle_lib, name):
    n = 10
    if name == "brickwall":
        c = W.brickwall(n, 12, seed=3)
    elif name == "qft":
        c = W.qft(n)
    elif name == "hea":
        c = W.hea(n, 4, seed=2)
    elif name == "qaoa":
        c = W.qaoa_maxcut(n, W.random_3_regular(n, 1), [0.4, 1.3], [0.9, 0.2])
    else:
        rng = np.random.default_rng(5)
        c = Circuit(n).h(0).h(5)
        c.add(DiagonalPhase(n, [PhaseTerm(tuple(1 << int(q) for q in rng.choice(n, 3, replace=False)), 0.7),
                                PhaseTerm((0b101,), -0.4), PhaseTerm((), 0.3)]))
        c.cx(5, 2).h(2)
    psi = W.gaussian_state(n, 9)
    for kw in ({}, {"local_qubits": 6}, {"reg_bits": 5, "local_qubits": 9}):
        plan = HostPlan(c, **kw)
        assert max_abs(emulate(plan, psi), oracle_state(oracle_lib, c, psi)) < 1e-12, kw


def test_brickwall_pair_fusion_structure():
    """U3xU3·CZ·U3xU3 blocks on even pairs, odd-layer CZs left as sign flips."""
    n, depth = 30, 20
    plan = HostPlan(W.brickwall(n, depth, seed=2024))
    kinds = [op.type for op in plan.ops]
    assert kinds.count(2) == (n // 2) * (depth // 2)            # 4x4 blocks
    assert kinds.count(4) == (n // 2 - 1) * (depth // 2)        # bare odd-layer CZs
    assert kinds.count(1) == 0
    unfused = HostPlan(W.brickwall(n, depth, seed=2024), fuse_2q=False)
    assert plan.stats["n_passes"] < unfused.stats["n_passes"]


def test_plan_stages_respect_register_width():
    plan = HostPlan(W.brickwall(20, 8, seed=1), local_qubits=11)
    for p in plan.passes:
        assert p.m == 11
        assert list(p.local_pos[:p.m]) == sorted(p.local_pos[:p.m])
        for sd in plan.stages[p.stage_begin:p.stage_begin + p.n_stages]:
            regs = [int(sd.rg[g]) for g in range(p.reg_bits)]
            assert len(set(regs)) == p.reg_bits
            assert sd.n_thread_bits == p.m - p.reg_bits
            chunk = sum(1 << q for q in p.local_pos[:p.m])
            assert all(r & chunk for r in regs)


def _random_sandwich(n, rng):
    """CNOT networks that compose to the identity around diagonal gates and
    DiagonalPhase terms, between dense layers."""
    c = Circuit(n)
    for q in range(n):
        c.h(q)
    for _ in range(3):
        cx = [tuple(int(v) for v in rng.choice(n, 2, replace=False)) for _ in range(int(rng.integers(1, 4)))]
        for a, b in cx:
            c.cx(a, b)
        c.rz(float(rng.uniform(0, 6.3)), int(rng.integers(n)))
        c.cz(*[int(v) for v in rng.choice(n, 2, replace=False)])
        c.cp(float(rng.uniform(0, 6.3)), *[int(v) for v in rng.choice(n, 2, replace=False)])
        c.ccz(*[int(v) for v in rng.choice(n, 3, replace=False)])
        terms = [PhaseTerm(tuple(int(m) for m in rng.integers(1, 1 << n, size=int(rng.integers(0, 3)))),
                           float(rng.uniform(-3, 3))) for _ in range(3)]
        c.add(DiagonalPhase(n, terms))
        c.t(int(rng.integers(n)))
        for a, b in reversed(cx):
            c.cx(a, b)
        for q in range(n):
            c.ry(float(rng.uniform(0, 6.3)), q)
    return c


@pytest.mark.parametrize("seed", range(6))
def test_diagonal_run_fusion_matches_oracle(oracle_lib, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(6, 10))
    c = _random_sandwich(n, rng)
    psi = W.gaussian_state(n, seed)
    want = oracle_state(oracle_lib, c, psi)
    for kw in ({}, {"local_qubits": 5}, {"fuse_diag": False}):
        plan = HostPlan(c, **kw)
        assert max_abs(emulate(plan, psi), want) < 1e-12, kw


def test_qaoa_cost_layer_becomes_phase_terms(oracle_lib):
    """CNOT·Rz·CNOT per edge compose to one DiagonalPhase per layer, which the
    fused passes apply as AND-mask phase terms: no controlled-X ops remain."""
    n = 12
    c = W.qaoa_maxcut(n, W.random_3_regular(n, 4), [0.3, 1.1, 2.0], [0.7, 0.1, 1.9])
    fused = HostPlan(c)
    plain = HostPlan(c, fuse_diag=False)
    assert not any(op.type == 1 and op.cmask for op in fused.ops)
    assert any(op.type == 1 and op.cmask for op in plain.ops)
    assert sum(op.type == 5 for op in fused.ops) >= 3
    assert fused.stats["n_passes"] <= plain.stats["n_passes"]
    psi = W.gaussian_state(n, 2)
    want = oracle_state(oracle_lib, c, psi)
    for plan in (fused, plain):
        assert max_abs(emulate(plan, psi), want) < 1e-12


def test_non_identity_cnot_runs_are_kept():
    """A CNOT ladder (HEA) does not compose to the identity: the run stays
    gate by gate and the plan is the same with and without run fusion."""
    c = W.hea(10, 3, seed=1)
    a, b = HostPlan(c), HostPlan(c, fuse_diag=False)
    assert [op.type for op in a.ops] == [op.type for op in b.ops]


@pytest.mark.parametrize("name", ["brickwall", "qaoa", "random"])
def test_emulated_fp32_plans(oracle_lib, name):
    """FP32 plans carry their op matrices as floats; emulating them in
    double precision from those floats stays within the FP32 budget."""
    n = 9
    if name == "brickwall":
        c = W.brickwall(n, 8, seed=4)
    elif name == "qaoa":
        c = W.qaoa_maxcut(8, W.random_3_regular(8, 2), [0.5, 1.7], [0.3, 2.2])
    else:
        c = random_circuit(n, 50, np.random.default_rng(8))
    psi = W.gaussian_state(c.num_qubits, 1)
    plan = HostPlan(c, precision=32)
    assert max_abs(emulate(plan, psi, precision=32), oracle_state(oracle_lib, c, psi)) < 1e-5


@pytest.mark.parametrize("width", [3, 4, 5, 6])
@pytest.mark.parametrize("name", ["brickwall", "random", "qft"])
def test_dense_block_fusion_matches_oracle(oracle_lib, width, name):
    """Greedy dense blocks of up to `width` qubits (standalone sweeps): the
    emulated plan equals the oracle's gate-by-gate run at 1e-12."""
    n = 9
    if name == "brickwall":
        c = W.brickwall(n, 10, seed=6)
    elif name == "qft":
        c = W.qft(n)
    else:
        c = random_circuit(n, 70, np.random.default_rng(width), kq=False)
    psi = W.gaussian_state(n, width)
    plan = HostPlan(c, kq_width=width)
    ks = [s.k for s in plan.steps if s.kind == 1]
    assert ks and max(ks) <= width
    assert max_abs(emulate(plan, psi), oracle_state(oracle_lib, c, psi)) < 1e-12
    if width >= 4:
        assert max(ks) >= 4


def test_dense_block_count_falls_with_width():
    c = W.brickwall(16, 20, seed=1)
    counts = [sum(1 for s in HostPlan(c, kq_width=w).steps if s.kind == 1) for w in (3, 4, 5, 6)]
    assert counts[0] > counts[1] > counts[2] >= counts[3], counts


@pytest.mark.parametrize("seed", range(6))
def test_trailing_swap_involution_is_one_permutation_step(oracle_lib, seed):
    """A program ending in disjoint SWAPs plans them as one in-place qubit
    permutation (step kind 7); anything else keeps the swap passes."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(13, 16))
    c = W.brickwall(n, 3, seed=seed)
    perm = rng.permutation(n)
    pairs = [(int(perm[2 * i]), int(perm[2 * i + 1])) for i in range(int(rng.integers(2, n // 2 + 1)))]
    for a, b in pairs:
        c.swap(a, b)
    c.barrier(list(range(n)))
    psi = W.gaussian_state(n, 10 + seed)
    plan = HostPlan(c)
    assert [s.kind for s in plan.steps].count(7) == 1 and plan.steps[-1].kind == 7
    assert max_abs(emulate(plan, psi), oracle_state(oracle_lib, c, psi)) < 1e-12
    # a 3-cycle of swaps is no involution: no permutation step
    c3 = W.brickwall(n, 2, seed=seed).swap(0, 1).swap(1, 2)
    plan3 = HostPlan(c3)
    assert 7 not in [s.kind for s in plan3.steps]
    assert max_abs(emulate(plan3, psi), oracle_state(oracle_lib, c3, psi)) < 1e-12


def test_qft_final_reversal_is_one_permutation_step():
    for n in (16, 30):
        kinds = [s.kind for s in HostPlan(W.qft(n)).steps]
        assert kinds[-1] == 7 and kinds.count(7) == 1


def test_tile_gates_fold_into_the_permutation(oracle_lib):
    """One-qubit gates on tile qubits just before a final SWAP involution run
    inside the permutation sweep: the QFT's last Hadamard no longer costs a
    pass of its own."""
    hp = HostPlan(W.qft(30))
    assert hp.steps[-1].kind == 7 and hp.steps[-1].k >= 1
    body = Circuit(30)
    for inst in W.qft(30, swaps=False).instructions[:-1]:      # all but the last H
        body.add(inst)
    assert len(hp.steps) == len(HostPlan(body).steps) + 1
    assert 7 not in [s.kind for s in HostPlan(W.qft(12)).steps]     # one chunk: swaps are register renames
    n = 16
    hp = HostPlan(W.qft(n))
    assert hp.steps[-1].kind == 7 and hp.steps[-1].k >= 1
    psi = W.gaussian_state(n, 3)
    assert max_abs(emulate(hp, psi), oracle_state(oracle_lib, W.qft(n), psi)) < 1e-12
    c = W.brickwall(n, 2, seed=1).h(0).rx(0.3, n - 1).u3(0.1, 0.2, 0.3, 1).swap(0, n - 1).swap(1, n - 2)
    hp = HostPlan(c)
    assert hp.steps[-1].kind == 7 and hp.steps[-1].k == 3
    assert max_abs(emulate(hp, psi), oracle_state(oracle_lib, c, psi)) < 1e-12


def _cancel_case(build, n=6):
    c = Circuit(n)
    build(c)
    return c


@pytest.mark.parametrize("name,build,cancelled", [
    ("h_h", lambda c: c.h(0).h(0), 2),
    ("s_sdg", lambda c: c.s(1).sdg(1), 2),
    ("t_tdg_nested", lambda c: c.t(2).h(2).h(2).tdg(2), 4),
    ("rz_neg", lambda c: c.rz(0.7, 3).rz(-0.7, 3), 2),
    ("cnot_cnot", lambda c: c.cx(0, 1).cx(0, 1), 2),
    ("cnot_reversed", lambda c: c.cx(0, 1).cx(1, 0), 0),
    ("blocked_by_gate", lambda c: c.h(0).cx(0, 1).h(0), 0),
    ("blocked_by_barrier", lambda c: c.h(0).barrier([0]).h(0), 0),
    ("blocked_by_phase", lambda c: c.h(0).add(DiagonalPhase(6, [PhaseTerm((1,), 0.3)])).h(0), 0),
    ("rx_not_inverse", lambda c: c.rx(0.7, 3).rx(0.7, 3), 0),
    ("swap_swap_other_qubit_between", lambda c: c.swap(0, 5).h(2).swap(0, 5), 2),
])
def test_cancel_inverse_pairs(oracle_lib, name, build, cancelled):
    """CancelInversePairs (crates/aleph-ir/src/passes/cancel.rs) in the GPU
    planner: adjacent exact inverses on the same qubits cancel, nested pairs
    cancel in one pass, and nothing cancels across a gate, a barrier or a
    diagonal phase on those qubits.  The plan still equals the oracle's
    gate-by-gate run."""
    c = _cancel_case(build).u3(0.3, 0.2, 0.1, 4)       # keep the program non-empty after cancelling
    hp = HostPlan(c)
    assert hp.stats["cancelled_gates"] == cancelled, name
    psi = W.gaussian_state(6, 5)
    assert max_abs(emulate(hp, psi), oracle_state(oracle_lib, c, psi)) < 1e-12


def test_cancel_inverse_pairs_random_mirrors(oracle_lib):
    """U followed by its gate-by-gate inverse in reverse order cancels to
    nothing for the closed-form kinds (each gate meets its inverse once the
    inner pairs are gone)."""
    rng = np.random.default_rng(3)
    n = 7
    c = Circuit(n)
    inv = {"H": "H", "X": "X", "Y": "Y", "Z": "Z", "S": "Sdg", "Sdg": "S", "T": "Tdg", "Tdg": "T"}
    seq = []
    for _ in range(40):
        if rng.random() < 0.6:
            name = list(inv)[int(rng.integers(len(inv)))]
            q = int(rng.integers(n))
            seq.append(("1", name, q))
        else:
            a, b = (int(v) for v in rng.choice(n, 2, replace=False))
            seq.append(("2", "Cnot" if rng.random() < 0.5 else "Cz", (a, b)))
    for kind, name, q in seq:
        c.gate(Gate(name), [q] if kind == "1" else list(q))
    for kind, name, q in reversed(seq):
        c.gate(Gate(inv[name] if kind == "1" else name), [q] if kind == "1" else list(q))
    c.u3(0.3, 0.2, 0.1, 0)
    hp = HostPlan(c)
    assert hp.stats["cancelled_gates"] == 2 * len(seq)
    psi = W.gaussian_state(n, 9)
    assert max_abs(emulate(hp, psi), oracle_state(oracle_lib, c, psi)) < 1e-12
