"""GPU parity: the HIP backend against the pinned CPU oracle, through the C ABI.

FP64 amplitudes must match the oracle within 1e-12 max-abs (the Aer
fixtures within 1e-10, harness.rs:18); FP32 within 1e-5 (harness.rs:27).
Sizes are ones the oracle finishes in seconds; large-n checks use
size-independent properties (norm, mirror circuits, determinism).
"""

import math

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd import backend as B
from aleph_amd import workloads as W
from aleph_amd.circuit import Circuit, DiagonalPhase, Gate, GateInstance, PauliString, PhaseTerm
from tests.helpers import (DISTRIBUTION_SHOTS, PARITY_TOL, STATE_TOLERANCE, assert_distribution_close,
                           fixture_amplitudes, fixture_circuit, fixture_names, has_measure, load_fixture, max_abs)
from tests.test_oracle import random_circuit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def oracle_run(orc, circuit, psi0=None):
    o = orc.OracleSv()
    st = o.allocate(circuit.num_qubits)
    if psi0 is not None:
        st.load(psi0)
    o.run(circuit, st)
    return st.amplitudes()


def gpu_program(hip, circuit, psi0=None, **kw):
    st = hip.allocate(circuit.num_qubits)
    if psi0 is not None:
        st.load(psi0)
    stats = hip.run_program(st, circuit, **kw)
    a = st.amplitudes()
    st.free()
    return a, stats


def gpu_gatewise(hip, circuit, psi0=None):
    st = hip.allocate(circuit.num_qubits)
    if psi0 is not None:
        st.load(psi0)
    for inst in circuit.instructions:
        if isinstance(inst, GateInstance):
            hip.apply_gate(st, inst)
        elif isinstance(inst, DiagonalPhase):
            hip.apply_diagonal_phase(st, inst)
    a = st.amplitudes()
    st.free()
    return a


def mirror(circuit, barrier: bool = False):
    """circuit followed by its inverse (U3(θ,φ,λ)^-1 = U3(−θ,−λ,−φ), kinds.rs inverse()).
    With `barrier`, a barrier on every qubit separates the halves, so the
    planner's inverse-pair cancellation cannot collapse the mirror and every
    gate is executed."""
    m = Circuit(circuit.num_qubits)
    gates = circuit.gates()
    for g in gates:
        m.gate(g.gate, g.qubits, g.controls)
    if barrier:
        m.barrier(range(circuit.num_qubits))
    for g in reversed(gates):
        gt = g.gate
        if gt.name == "U3":
            inv = Gate.U3(-gt.params[0], -gt.params[2], -gt.params[1])
        elif gt.name in ("Rx", "Ry", "Rz", "Phase"):
            inv = Gate(gt.name, (-gt.params[0],))
        elif gt.name in ("H", "X", "Y", "Z", "Cnot", "Cz", "Swap", "Toffoli", "Ccz"):
            inv = gt
        else:
            raise ValueError(gt.name)
        m.gate(inv, g.qubits, g.controls)
    return m


# ---------------- allocation and per-gate path ----------------

def test_allocate_zero_state(hip):
    st = hip.allocate(5)
    a = st.amplitudes()
    assert a[0] == 1 and np.all(a[1:] == 0)


def test_too_many_qubits(hip):
    b = B.HipSvBackend(seed=0).with_qubit_cap(10)
    with pytest.raises(B.TooManyQubits) as e:
        b.allocate(11)
    assert (e.value.requested, e.value.limit) == (11, 10)


@pytest.mark.parametrize("seed", range(16))
def test_apply_gate_random_circuits(hip, orc, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 11))
    c = random_circuit(n, 50, rng)
    psi = W.gaussian_state(n, seed)
    print(f"apply_gate seed={seed} n={n}", flush=True)
    assert max_abs(gpu_gatewise(hip, c, psi), oracle_run(orc, c, psi)) < PARITY_TOL


@pytest.mark.parametrize("name", [n for n in fixture_names()])
def test_aer_fixtures_gatewise_and_planned(hip, name):
    c = fixture_circuit(name)
    if has_measure(c):
        pytest.skip("mid-circuit measurement")
    want = fixture_amplitudes(load_fixture(name))
    assert max_abs(gpu_gatewise(hip, c), want) < STATE_TOLERANCE
    got, _ = gpu_program(hip, c)
    assert max_abs(got, want) < STATE_TOLERANCE


def test_gate_errors_match_oracle_variants(hip):
    st = hip.allocate(2)
    with pytest.raises(B.QubitOutOfRange) as e:
        hip.apply_gate(st, GateInstance(Gate.X(), [5]))
    assert (e.value.qubit, e.value.num_qubits) == (5, 2)
    with pytest.raises(B.DuplicateQubit):
        hip.apply_gate(st, GateInstance(Gate.X(), [0], [0]))
    with pytest.raises(B.ArityMismatch) as e:
        hip.apply_gate(st, GateInstance(Gate.Cnot(), [0]))
    assert (e.value.expected, e.value.got) == (2, 1)
    with pytest.raises(B.NonFiniteParam):
        hip.apply_gate(st, GateInstance(Gate.Rx(float("nan")), [0]))
    with pytest.raises(B.NonUnitaryMatrix):
        hip.apply_gate(st, GateInstance(Gate.unitary(2 * np.eye(2)), [0]))
    with pytest.raises(B.QubitOutOfRange):
        hip.run_program(st, Circuit(2).h(0).cx(0, 3))


# ---------------- planned programs (fused LDS passes) ----------------

@pytest.mark.parametrize("n", [4, 5, 7, 9, 12, 13, 16, 20])
def test_brickwall_planned(hip, orc, n):
    c = W.brickwall(n, 20, seed=n)
    print(f"brickwall n={n}", flush=True)
    got, stats = gpu_program(hip, c)
    assert max_abs(got, oracle_run(orc, c)) < PARITY_TOL
    assert stats["n_passes"] <= len(c)


@pytest.mark.parametrize("n", [4, 6, 11, 14, 18])
def test_qft_gaussian_planned(hip, orc, n):
    psi = W.gaussian_state(n, 7)
    c = W.qft(n)
    print(f"qft n={n}", flush=True)
    got, _ = gpu_program(hip, c, psi)
    assert max_abs(got, oracle_run(orc, c, psi)) < PARITY_TOL


@pytest.mark.parametrize("seed", range(8))
def test_trailing_swap_involution_sweep(hip, hip32, orc, seed):
    """A program ending in disjoint SWAPs (at least two) runs them as one
    in-place permutation sweep (launch kind 7), with the one-qubit gates
    just before them on tile qubits applied in the same sweep; the result
    equals the oracle's gate-by-gate run in both precisions."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(13, 21))
    c = random_circuit(n, 30, rng, kq=False, controls=False)
    perm = [int(q) for q in rng.permutation(n)]
    pairs = int(rng.integers(2, n // 2 + 1))
    low = [q for q in range(min(5, n))]
    for q in low[: int(rng.integers(0, 4))]:
        c.h(q)
    if seed % 2:
        c.ry(0.3 + seed, perm[0])
    for j in range(pairs):
        c.swap(perm[2 * j], perm[2 * j + 1])
    psi = W.gaussian_state(n, seed)
    want = oracle_run(orc, c, psi)
    print(f"bitperm seed={seed} n={n} pairs={pairs}", flush=True)
    for b, tol in ((hip, PARITY_TOL), (hip32, 2e-5)):
        for jit in (False, True):
            st = b.allocate(n)
            st.load(psi)
            plan = b.compile(c, jit=jit)
            _, kinds = plan.execute_timed(st)
            b.synchronize()
            plan.free()
            got = st.amplitudes()
            st.free()
            assert kinds[-1] == 7
            assert max_abs(got, want) < tol


@pytest.mark.parametrize("seed", range(20))
def test_random_circuits_planned(hip, orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 13))
    c = random_circuit(n, 80, rng)
    psi = W.gaussian_state(n, 1000 + seed)
    opts = dict(local_qubits=int(rng.integers(4, min(n, 12) + 1)), max_ops=int(rng.choice([1, 3, 24, 500])),
                coalesce_bits=int(rng.integers(0, 4)), fuse_1q=bool(rng.integers(0, 2)))
    print(f"random planned seed={seed} n={n} opts={opts}", flush=True)
    got, _ = gpu_program(hip, c, psi, **opts)
    assert max_abs(got, oracle_run(orc, c, psi)) < PARITY_TOL


@pytest.mark.parametrize("planner", [L.PLANNER_UNFUSED, L.PLANNER_RULE])
def test_hea_and_qaoa_planned(hip, orc, planner):
    c = W.hea(10, 4, seed=3)
    got, _ = gpu_program(hip, c, planner=planner)
    assert max_abs(got, oracle_run(orc, c)) < PARITY_TOL
    edges = W.random_3_regular(10, 5)
    q = W.qaoa_maxcut(10, edges, [0.3, 1.1], [0.7, 2.0])
    got, _ = gpu_program(hip, q, planner=planner)
    assert max_abs(got, oracle_run(orc, q)) < PARITY_TOL


def test_diagonal_phase_gpu(hip, orc):
    n = 9
    rng = np.random.default_rng(2)
    psi = W.gaussian_state(n, 2)
    # parity conditions (general) and single-bit AND conditions
    general = DiagonalPhase(n, [PhaseTerm(tuple(int(rng.integers(1, 1 << n)) for _ in range(int(rng.integers(0, 3)))),
                                          float(rng.uniform(-3, 3))) for _ in range(12)])
    single = DiagonalPhase(n, [PhaseTerm(tuple(1 << int(q) for q in rng.choice(n, int(rng.integers(0, 4)),
                                                                                replace=False)),
                                         float(rng.uniform(-3, 3))) for _ in range(20)])
    for dp in (general, single):
        c = Circuit(n).h(0).add(dp).cx(0, 3)
        want = oracle_run(orc, c, psi)
        assert max_abs(gpu_gatewise(hip, c, psi), want) < PARITY_TOL
        got, _ = gpu_program(hip, c, psi)
        assert max_abs(got, want) < PARITY_TOL


def test_program_determinism(hip):
    c = W.brickwall(16, 10, seed=5)
    a, _ = gpu_program(hip, c)
    b, _ = gpu_program(hip, c)
    assert np.array_equal(a, b)


def test_large_mirror_and_norm(hip):
    n = 26
    c = mirror(W.brickwall(n, 6, seed=9), barrier=True)
    st = hip.allocate(n)
    stats = hip.run_program(st, c)
    assert stats["cancelled_gates"] == 0
    assert abs(hip.norm_sqr(st) - 1.0) < 1e-12
    amp0 = hip.probabilities(st, [])  # validates the state
    assert amp0[0] == 1.0
    p = hip.probabilities(st, list(range(8)))
    assert abs(p[0] - 1.0) < 1e-10
    st.free()


# ---------------- readout ----------------

def test_norm_measure_expectation_probabilities(hip, orc):
    n = 10
    c = W.brickwall(n, 6, seed=3)
    st = hip.allocate(n)
    hip.run_program(st, c)
    o = orc.OracleSv()
    ost = o.run(c)
    assert abs(hip.norm_sqr(st) - 1.0) < 1e-13
    rng = np.random.default_rng(8)
    for _ in range(12):
        k = int(rng.integers(0, 5))
        qs = rng.choice(n, k, replace=False)
        ps = PauliString.new(float(rng.uniform(-2, 2)), [(int(q), "XYZ"[rng.integers(3)]) for q in qs])
        assert abs(hip.expectation_value(st, ps) - o.expectation_value(ost, ps)) < 1e-12
    for qs in ([0], [3, 1], [9, 0, 4], list(range(10))):
        assert max_abs(hip.probabilities(st, qs), o.probabilities(ost, qs)) < 1e-13
    for q, u in ((2, 0.1), (7, 0.9), (0, 0.5)):
        g_out, g_used = hip.measure_with_uniform(st, q, u)
        o_out, o_used = o.measure_with_uniform(ost, q, u)
        assert (g_out, g_used) == (o_out, o_used)
        assert max_abs(st.amplitudes(), ost.amplitudes()) < 1e-12
    st.free()


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 1e-5)])
def test_expectation_single_strings_every_x_mask(orc, precision, tol):
    """Backend::expectation_value (measure.rs:155-235) of single strings
    whose X/Y parts sit on low, high and mixed qubits (the pair sums over
    {i, i ^ x}), identity and Z-only strings (x = 0), with even and odd
    numbers of Y: equal to the oracle."""
    n = 16
    hip = B.HipSvBackend(seed=4, precision=precision)
    c = W.brickwall(n, 6, seed=11)
    st = hip.allocate(n)
    hip.run_program(st, c)
    o = orc.OracleSv()
    ost = o.run(c)
    strings = [[], [(15, "X")], [(15, "Y")], [(0, "Y"), (15, "Y")], [(14, "Z"), (3, "X")],
               [(15, "X"), (7, "Y"), (0, "Z")], [(2, "Z"), (9, "Z")], [(8, "Y"), (12, "Y"), (13, "Y")]]
    for terms in strings:
        ps = PauliString.new(1.3, terms)
        assert abs(hip.expectation_value(st, ps) - o.expectation_value(ost, ps)) < tol, terms
    st.free()


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 1e-5)])
def test_expectation_pauli_sum(orc, precision, tol):
    """Backend::expectation_pauli_sum (crates/aleph-backend/src/lib.rs:422-432):
    one call over many Pauli strings equals the sum of the oracle's
    per-string expectations; a MaxCut Hamiltonian sum_edges (1 - Z_u Z_v)/2
    (maxcut_pauli_sum, crates/aleph-ir/src/ansatz.rs) equals sum_x |psi_x|^2 cut(x)."""
    n = 11
    hip = B.HipSvBackend(seed=4, precision=precision)
    c = W.brickwall(n, 6, seed=5)
    st = hip.allocate(n)
    hip.run_program(st, c)
    o = orc.OracleSv()
    ost = o.run(c)
    rng = np.random.default_rng(21)
    terms = []
    for _ in range(40):
        k = int(rng.integers(0, 6))
        qs = rng.choice(n, k, replace=False)
        terms.append(PauliString.new(float(rng.uniform(-2, 2)), [(int(q), "XYZ"[rng.integers(3)]) for q in qs]))
    want = sum(o.expectation_value(ost, p) for p in terms)
    assert abs(hip.expectation_pauli_sum(st, terms) - want) < tol * len(terms)
    assert hip.expectation_pauli_sum(st, []) == 0.0
    edges = W.random_3_regular(10, 7)
    ham = [PauliString.new(0.5, []) for _ in edges] + [PauliString.new(-0.5, [(u, "Z"), (v, "Z")]) for u, v in edges]
    probs = np.abs(ost.amplitudes()) ** 2
    idx = np.arange(1 << n)
    cut = sum(((idx >> u) ^ (idx >> v)) & 1 for u, v in edges)
    assert abs(hip.expectation_pauli_sum(st, ham) - float(np.sum(probs * cut))) < tol * len(ham)
    with pytest.raises(B.QubitOutOfRange):
        hip.expectation_pauli_sum(st, [PauliString.new(1.0, [(n + 3, "Z")])])
    # The drift budget: FP64 sqrt(2^n) * 1e-10, FP32 a fixed 1e-3
    # (crates/aleph-cuda/src/sv/readout_f32.rs:34-39); a state scaled by 0.99
    # (norm² 0.98) is rejected in both precisions.
    st.load(0.99 * ost.amplitudes())
    with pytest.raises(B.InvalidState):
        hip.expectation_pauli_sum(st, terms)
    st.free()


def test_readout_errors(hip):
    st = hip.allocate(2)
    with pytest.raises(B.QubitOutOfRange):
        hip.probabilities(st, [5])
    with pytest.raises(B.DuplicateQubit):
        hip.probabilities(st, [0, 0])
    with pytest.raises(B.InvalidPauliString):
        hip.expectation_value(st, PauliString(float("nan"), []))
    assert hip.expectation_value(st, PauliString.identity(2.5)) == pytest.approx(2.5, abs=1e-12)
    bad = hip.allocate(1)
    bad.load([1.0, float("nan")])
    with pytest.raises(B.InvalidState):
        hip.measure(bad, 0)
    out, used = hip.measure_with_uniform(hip.allocate(2), 0, 0.3)
    assert out is False and used is False


def test_sampling_kats(hip):
    assert np.all(hip.sample(hip.allocate(3), 100) == 0)
    st = hip.allocate(2)
    hip.run_program(st, Circuit(2).h(0).cx(0, 1))
    shots = hip.sample(st, 1000)
    assert set(np.unique(shots)) <= {0, 3}
    assert (shots == 0).sum() > 100 and (shots == 3).sum() > 100
    assert hip.sample(st, 0).size == 0


@pytest.mark.parametrize("n", [3, 12, 14, 17])
def test_sampler_matches_inverse_cdf(hip, n):
    """With the same uniforms, the hierarchical sampler returns the first index
    whose cumulative probability exceeds u * total (sampling.rs:61-91)."""
    psi = W.gaussian_state(n, 40 + n)
    psi[np.random.default_rng(n).random(1 << n) < 0.2] = 0     # zero-mass bins must be skipped
    psi /= np.linalg.norm(psi)
    st = hip.allocate(n)
    st.load(psi)
    u = np.random.default_rng(n + 1).random(4096)
    u[:3] = [0.0, 0.5, 1.0 - 2 ** -53]
    got = hip.sample(st, len(u), uniforms=u).astype(np.int64)
    p = np.abs(st.amplitudes()) ** 2
    cum = np.cumsum(p)
    want = np.searchsorted(cum, u * cum[-1], side="right")
    want = np.minimum(want, np.flatnonzero(p > 0)[-1])
    close = np.abs(cum[np.minimum(want, len(cum) - 1)] - u * cum[-1]) < 1e-12
    assert np.all((got == want) | close)
    assert np.all(p[got] > 0)


@pytest.mark.parametrize("name", ["bell_phi_plus", "ghz_5", "qft_5", "random_nonclifford_n4_d20"])
def test_sampling_distribution_vs_aer(hip, name):
    fx = load_fixture(name)
    st = hip.allocate(fx["num_qubits"])
    hip.run_program(st, fixture_circuit(name))
    shots = hip.sample(st, DISTRIBUTION_SHOTS)
    counts = np.bincount(shots.astype(np.int64), minlength=1 << fx["num_qubits"])
    assert_distribution_close(counts, np.abs(fixture_amplitudes(fx)) ** 2, DISTRIBUTION_SHOTS, name)


def test_unpermute(hip, orc):
    n = 7
    psi = W.gaussian_state(n, 77)
    perm = [3, 0, 6, 1, 5, 2, 4]
    st = hip.allocate(n)
    st.load(psi)
    hip.unpermute_state(st, perm)
    o = orc.OracleSv()
    ost = o.allocate(n)
    ost.load(psi)
    o.unpermute_state(ost, perm)
    assert np.array_equal(st.amplitudes(), ost.amplitudes())


# ---------------- FP32 amplitudes ----------------

@pytest.mark.parametrize("n", [6, 12, 16])
def test_fp32_brickwall(hip32, orc, n):
    c = W.brickwall(n, 20, seed=n)
    got, _ = gpu_program(hip32, c)
    want = oracle_run(orc, c)
    err = max_abs(got, want)
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real)
    print(f"fp32 brickwall n={n} max_abs={err:.3e} 1-F={1 - fid:.3e}", flush=True)
    assert err < 1e-5
    assert 1 - fid < 1e-6


def test_normalize_and_compare(hip):
    c = W.brickwall(14, 8, seed=3)
    a = hip.allocate(14)
    hip.run_program(a, c, jit=True)
    b = hip.allocate(14, precision=32)            # one context, two precisions
    hip.run_program(b, c, jit=True)
    r = hip.compare(a, a)
    assert abs(r["fidelity"] - 1.0) < 1e-12 and r["max_abs"] == 0.0 and r["l2_sq"] == 0.0
    r = hip.compare(a, b)          # FP64 against FP32, accumulated in FP64
    assert 1.0 - r["fidelity"] < 1e-6 and r["max_abs"] < 1e-5
    want = a.amplitudes()
    assert abs(r["overlap"] - np.vdot(want, b.amplitudes())) < 1e-9
    b.load(1.5 * b.amplitudes())
    assert abs(hip.normalize(b) - 2.25) < 1e-5
    assert abs(hip.norm_sqr(b) - 1.0) < 1e-6
    assert max_abs(b.amplitudes(), want) < 1e-5
    a.free()
    b.free()


def test_renormalized_and_generic_phase_plans(hip, hip32):
    """RENORMALIZE ends every execution at norm 1 in FP64; GENERIC_PHASE
    kernels equal the per-class generated ones."""
    edges = W.random_3_regular(16, 7)
    c = W.qaoa_maxcut(16, edges, [0.4, 1.1, 0.7], [0.9, 0.2, 1.3])
    outs = []
    for generic in (False, True):
        st = hip.allocate(16)
        p = hip.compile(c, jit=True, generic_phase=generic)
        p.execute(st)
        outs.append(st.amplitudes())
        p.free()
        st.free()
    assert max_abs(outs[0], outs[1]) < PARITY_TOL
    st = hip32.allocate(18)
    p = hip32.compile(W.brickwall(18, 40, seed=5), jit=True, renormalize=True)
    for _ in range(3):
        p.execute(st)
    assert abs(hip32.norm_sqr(st) - 1.0) < 2e-7
    p.free()
    st.free()
