"""Distributed state vector on one device (the LocalExchange pattern of
crates/aleph-cuda/src/dist/exchange.rs: every rank slice on one GPU): the
exchange kernel against the CPU chunk form (dist_ref.rs:78-104), and routed
plans against the oracle's physical execution and against the single-state
GPU run of the same circuit."""

import ctypes as C

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd import dist as D
from aleph_amd import workloads as W
from aleph_amd.backend import HipSvBackend, InvalidArgument, _ptr
from oracle import dist_oracle as DO
from tests.helpers import PARITY_TOL, max_abs
from tests.test_dist import random_logical_circuit

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n,g,bits", [(12, 1, [11]), (12, 2, [10, 11]), (14, 3, [12, 11, 13]), (20, 2, [19])])
def test_exchange_kernel_equals_chunk_form(precision, n, g, bits):
    b = HipSvBackend(seed=1, precision=precision)
    dx = D.DistSvBackend(b)
    layout = D.DistLayout(n, g)
    m = layout.m()
    rng = np.random.default_rng(n + g)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    if precision == 32:
        v = v.astype(np.complex64).astype(np.complex128)
    ranks = dx.alloc_ranks(layout)
    for r, st in enumerate(ranks):
        st.load(v[r << m:(r + 1) << m])
    print(f"exchange n={n} g={g} bits={bits} p{precision}", flush=True)
    dx.exchange(ranks, layout, bits)
    got = np.concatenate([st.amplitudes() for st in ranks])
    want = np.concatenate(DO.exchange_cpu([v[r << m:(r + 1) << m] for r in range(1 << g)], m, bits))
    assert np.array_equal(got, want)
    for st in ranks:
        st.free()


def test_alloc_rank_and_bad_exchanges():
    b = HipSvBackend(seed=1)
    dx = D.DistSvBackend(b)
    layout = D.DistLayout(10, 2)
    ranks = dx.alloc_ranks(layout)
    amps = [st.amplitudes() for st in ranks]
    assert amps[0][0] == 1 and np.count_nonzero(amps[0]) == 1 and all(not np.any(a) for a in amps[1:])
    for bits in ([7], [8, 8], [10], []):
        with pytest.raises(InvalidArgument):
            dx.exchange(ranks, layout, bits)
    with pytest.raises(InvalidArgument):
        dx.exchange(ranks[:3], layout, [8])
    for st in ranks:
        st.free()


@pytest.mark.parametrize("seed", range(4))
def test_routed_random_circuits_match_oracle(oracle_lib, seed):
    rng = np.random.default_rng(50 + seed)
    n = int(rng.integers(9, 13))
    g = int(rng.integers(1, 4))
    c = random_logical_circuit(n, 60, rng)
    p = D.plan(c, D.DistLayout(n, g))
    print(f"dist random seed={seed} n={n} g={g} exchanges={p.stats.exchanges}", flush=True)
    st = D.DistSvBackend(HipSvBackend(seed=1)).run(p)
    assert max_abs(st.amplitudes(), DO.run_physical(p, n)) < PARITY_TOL
    st.free()


@pytest.mark.parametrize("name,n,g", [("brickwall", 20, 2), ("qft", 18, 3), ("hea", 16, 1), ("brickwall", 22, 3)])
def test_distributed_run_equals_single_state_run(name, n, g):
    c = {"qft": W.qft(n), "brickwall": W.brickwall(n, 12, seed=n), "hea": W.hea(n, 4, seed=n)}[name]
    b = HipSvBackend(seed=1)
    one = b.allocate(n)
    b.run_program(one, c, jit=True)
    want = one.amplitudes()
    one.free()
    p = D.plan(c, D.DistLayout(n, g))
    print(f"dist {name} n={n} g={g} exchanges={p.stats.exchanges} moved/rank={p.stats.amps_moved_per_rank}",
          flush=True)
    st = D.DistSvBackend(b, jit=True).run(p)
    assert max_abs(st.amplitudes(), want) < PARITY_TOL
    st.free()
