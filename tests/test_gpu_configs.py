"""The benchmark workloads at their own sizes on the device, each against the
oracle or a size-independent property:

* random brickwall n=20 depth 20 (FP64): amplitudes vs the oracle, and 4096
  shots accepted by k-qubit marginal TVD against a simulated multinomial
  null and by linear XEB within 3 sigma of the exact value;
* the reference's published linear XEB of sycamore_n{20,24,28,30}_d20
  (docs/perf/phase4.md:51-56; XEB = 2^n sum p^2 - 1, benches/src/lib.rs:29);
* QFT on a seeded Gaussian state: n=24 amplitudes vs the oracle, n=30 norm
  and the closed form QFT|k> at n=30;
* brickwall depth 40: n=28 FP64 (generated pass kernels) vs the oracle run
  through its restated CPU pipeline, FP32 1-F and max-abs vs the same, and
  dense-block plans (kq_width 3..6) vs the oracle at n=20 and vs the FP64
  rule plan at n=28;
* mirror circuits U U^dag |0> = |0>: the bench's own n=30 depth-20 FP64 plan
  (seed 2024, generated kernels) and n=32 FP32 depth 40 (32 GiB).
"""

import math

import numpy as np
import pytest

from aleph_amd import workloads as W
from aleph_amd.backend import HipSvBackend
from oracle import pipeline as P
from tests.helpers import GOLDEN, PARITY_TOL, max_abs
from tests.test_gpu_parity import mirror

pytestmark = pytest.mark.gpu


def oracle_final(orc, circuit, psi0=None, cap=30):
    """Final state through the restated reference CPU pipeline (FP64)."""
    o = orc.OracleSv(cap=cap)
    st = o.allocate(circuit.num_qubits)
    if psi0 is not None:
        st.load(psi0)
    P.run_final_state(o, st, circuit)
    return st.amplitudes()


def gpu_final(b, circuit, psi0=None, **kw):
    st = b.allocate(circuit.num_qubits)
    if psi0 is not None:
        st.load(psi0)
    plan = b.compile(circuit, **kw)
    plan.execute(st)
    plan.free()
    a = st.amplitudes()
    st.free()
    return a


def bit_reverse(x, n):
    x = np.asarray(x, dtype=np.int64)
    r = np.zeros_like(x)
    for b in range(n):
        r |= ((x >> b) & 1) << (n - 1 - b)
    return r


def fidelity(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


# ---------------------------------------------------------------- brickwall n=20


def test_brickwall_n20_amplitudes_and_sampling(oracle_lib):
    n, shots = 20, 4096
    c = W.brickwall(n, 20, seed=2024)
    want = oracle_final(oracle_lib, c)
    b = HipSvBackend(seed=77)
    st = b.allocate(n)
    b.run_program(st, c, jit=True)
    got = st.amplitudes()
    err = max_abs(got, want)
    print(f"brickwall n=20 d20: max_abs {err:.2e}", flush=True)
    assert err < PARITY_TOL
    p = np.abs(want) ** 2
    x = b.sample(st, shots).astype(np.int64)
    st.free()
    assert np.all((x >= 0) & (x < (1 << n)))
    rng = np.random.default_rng(5)
    idx = np.arange(1 << n)
    for qs in ([0, 1, 2, 3, 4, 5, 6, 7], [6, 7, 8, 9, 10, 11, 12, 13], [12, 13, 14, 15, 16, 17, 18, 19]):
        key = np.zeros(1 << n, dtype=np.int64)
        for j, q in enumerate(qs):
            key |= ((idx >> q) & 1) << j
        pm = np.bincount(key, weights=p, minlength=1 << len(qs))
        sk = np.zeros(shots, dtype=np.int64)
        for j, q in enumerate(qs):
            sk |= ((x >> q) & 1) << j
        tvd = 0.5 * np.abs(np.bincount(sk, minlength=1 << len(qs)) / shots - pm).sum()
        null = 0.5 * np.abs(rng.multinomial(shots, pm / pm.sum(), size=2000) / shots - pm).sum(axis=1)
        q99 = float(np.quantile(null, 0.99))
        print(f"marginal {qs[0]}..{qs[-1]}: TVD {tvd:.4f}, null 99th percentile {q99:.4f}", flush=True)
        assert tvd <= q99
    dim = float(1 << n)
    s2, s3 = float(np.sum(p * p)), float(np.sum(p ** 3))
    xeb_exact = dim * s2 - 1.0
    xeb_shots = dim * float(np.mean(p[x])) - 1.0
    sigma = dim * math.sqrt((s3 - s2 * s2) / shots)
    print(f"linear XEB of shots {xeb_shots:.4f}, exact {xeb_exact:.4f}, sigma {sigma:.4f}", flush=True)
    assert abs(xeb_shots - xeb_exact) <= 3 * sigma


# ---------------------------------------------------------------- published XEB

PUBLISHED_XEB = {20: 1.0195, 24: 1.0245, 28: 1.0379, 30: 1.0355}   # docs/perf/phase4.md:53-56


@pytest.mark.parametrize("n", [20, 24, 28, 30])
def test_sycamore_published_linear_xeb(n):
    c = W.load_qasm(GOLDEN / "corpus" / f"sycamore_n{n}_d20.qasm")
    b = HipSvBackend(seed=1)
    st = b.allocate(n)
    stats = b.run_program(st, c, jit=True)
    xeb = b.linear_xeb(st)
    norm = b.norm_sqr(st)
    st.free()
    print(f"sycamore n={n}: {stats['input_gates']} gates, {stats['n_passes']} passes, XEB {xeb:.6f} "
          f"(published {PUBLISHED_XEB[n]}), norm-1 {norm - 1:.1e}", flush=True)
    assert abs(norm - 1.0) < 1e-12
    assert abs(xeb - PUBLISHED_XEB[n]) <= 1e-4


def test_linear_xeb_small_states(oracle_lib):
    """The device collision sum equals 2^n sum p^2 - 1 computed on the host,
    in both precisions; a basis state has XEB 2^n - 1, a uniform state 0."""
    c = W.brickwall(12, 8, seed=3)
    want = oracle_final(oracle_lib, c)
    ref = (1 << 12) * float(np.sum(np.abs(want) ** 4)) - 1.0
    for prec in (64, 32):
        b = HipSvBackend(seed=1, precision=prec)
        st = b.allocate(12)
        assert abs(b.linear_xeb(st) - ((1 << 12) - 1)) < 1e-9
        b.run_program(st, c)
        assert abs(b.linear_xeb(st) - ref) < (1e-10 if prec == 64 else 1e-4)
        st.load(np.full(1 << 12, 1 / 64.0, dtype=np.complex128))
        assert abs(b.linear_xeb(st)) < 1e-6
        st.free()


# ---------------------------------------------------------------- QFT


@pytest.mark.parametrize("jit", [False, True])
def test_qft_n24_gaussian_vs_oracle(oracle_lib, jit):
    n = 24
    c = W.qft(n)
    psi = W.gaussian_state(n, 2024)
    want = oracle_final(oracle_lib, c, psi)
    got = gpu_final(HipSvBackend(seed=1), c, psi, jit=jit)
    err = max_abs(got, want)
    print(f"QFT n=24 Gaussian (jit={jit}): max_abs {err:.2e}", flush=True)
    assert err < PARITY_TOL


def test_qft_n30_norm_and_basis_state_closed_form():
    """QFT on a seeded Gaussian state keeps the norm at n=30; QFT|k> has
    amplitude 2^(-n/2) e^{2 pi i r(j) r(k) / 2^n} at index j, r reversing the
    n index bits (qubit 0 is the index LSB; the convention is checked against
    the oracle by test_qft_is_dft in tests/test_oracle.py: QFT = R F R)."""
    n = 30
    b = HipSvBackend(seed=1)
    c = W.qft(n)
    plan = b.compile(c, jit=True)
    st = b.allocate(n)
    st.load(W.gaussian_state(n, 2024))
    plan.execute(st)
    norm = b.norm_sqr(st)
    print(f"QFT n=30 Gaussian: norm-1 {norm - 1:.2e}", flush=True)
    assert abs(norm - 1.0) < 1e-12
    k = 0x2A5F3C1
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[k] = 1.0
    st.load(psi)
    del psi
    plan.execute(st)
    plan.free()
    got = st.amplitudes()
    st.free()
    rng = np.random.default_rng(3)
    j = np.concatenate([np.arange(4096), rng.integers(0, 1 << n, 1 << 20)])
    ph = (bit_reverse(j, n) * bit_reverse(np.array([k]), n)[0]) % (1 << n)
    want = np.exp(2j * np.pi * ph / float(1 << n)) / math.sqrt(1 << n)
    err = float(np.max(np.abs(got[j] - want)))
    print(f"QFT n=30 |k>: max_abs over {j.size} sampled amplitudes {err:.2e}", flush=True)
    assert err < 1e-12


# ---------------------------------------------------------------- depth 40, fusion widths


@pytest.fixture(scope="module")
def brickwall28(oracle_lib):
    c = W.brickwall(28, 40, seed=2024)
    print("oracle: brickwall n=28 depth 40 through the restated CPU pipeline", flush=True)
    return c, oracle_final(oracle_lib, c)


def test_brickwall_n28_d40_fp64_vs_oracle(brickwall28):
    c, want = brickwall28
    got = gpu_final(HipSvBackend(seed=1), c, jit=True)
    err = max_abs(got, want)
    print(f"brickwall n=28 d40 FP64 generated kernels: max_abs {err:.2e}", flush=True)
    assert err < PARITY_TOL


def test_brickwall_n28_d40_fp32_fidelity(brickwall28):
    c, want = brickwall28
    got = gpu_final(HipSvBackend(seed=1, precision=32), c, jit=True)
    err = max_abs(got, want)
    inf = 1.0 - fidelity(got, want)
    print(f"brickwall n=28 d40 FP32: 1-F {inf:.2e}, max_abs {err:.2e}", flush=True)
    assert inf <= 1e-6
    assert err < 1e-5


@pytest.mark.parametrize("kq", [3, 4, 5, 6])
def test_dense_block_plans_vs_oracle_n20(oracle_lib, kq):
    c = W.brickwall(20, 40, seed=7)
    want = oracle_final(oracle_lib, c)
    b = HipSvBackend(seed=1)
    got = gpu_final(b, c, kq_width=kq)
    err = max_abs(got, want)
    print(f"brickwall n=20 d40 kq_width={kq}: max_abs {err:.2e}", flush=True)
    assert err < PARITY_TOL
    got32 = gpu_final(HipSvBackend(seed=1, precision=32), c, kq_width=kq)
    assert max_abs(got32, want) < 1e-5


def test_dense_block_plans_vs_rule_plan_n28(brickwall28):
    c, want = brickwall28
    b = HipSvBackend(seed=1)
    for kq in (3, 4, 5, 6):
        got = gpu_final(b, c, kq_width=kq)
        err = max_abs(got, want)
        print(f"brickwall n=28 d40 kq_width={kq}: max_abs vs oracle {err:.2e}", flush=True)
        assert err < PARITY_TOL


# ---------------------------------------------------------------- mirrors at the bench sizes


def _mirror_fidelity_lower_bound(b, st, n):
    """P(all qubits 0) >= 1 - sum over two halves of P(half not all 0)."""
    lo = b.probabilities(st, list(range(n // 2)))
    hi = b.probabilities(st, list(range(n // 2, n)))
    return 1.0 - (1.0 - lo[0]) - (1.0 - hi[0])


def test_mirror_headline_plan_n30_fp64():
    n = 30
    c = mirror(W.brickwall(n, 20, seed=2024), barrier=True)
    b = HipSvBackend(seed=1)
    st = b.allocate(n)
    stats = b.run_program(st, c, jit=True)
    assert stats["cancelled_gates"] == 0 and stats["n_passes"] >= 30
    f = _mirror_fidelity_lower_bound(b, st, n)
    norm = b.norm_sqr(st)
    st.free()
    print(f"mirror n=30 d20 FP64: {stats['n_passes']} passes, 1-F <= {1 - f:.2e}, norm-1 {norm - 1:.1e}", flush=True)
    assert 1.0 - f <= 1e-12
    assert abs(norm - 1.0) < 1e-12


def test_mirror_depth40_n32_fp32():
    n = 32
    c = mirror(W.brickwall(n, 40, seed=2024), barrier=True)
    b = HipSvBackend(seed=1, precision=32)
    st = b.allocate(n)
    stats = b.run_program(st, c, jit=True)
    assert stats["cancelled_gates"] == 0 and stats["n_passes"] >= 60
    f = _mirror_fidelity_lower_bound(b, st, n)
    norm = b.norm_sqr(st)
    st.free()
    print(f"mirror n=32 d40 FP32: {stats['n_passes']} passes, 1-F <= {1 - f:.2e}, norm-1 {norm - 1:.1e}", flush=True)
    assert 1.0 - f <= 1e-5
    assert abs(norm - 1.0) < 1e-4
