"""Host-side checks of batch packing (no device): every work item stays
inside its instance's state and pass, items cover every chunk once, the
dependency targets count exactly the chunks of earlier passes, and the
queue lists every dependency before its dependants (so a persistent
workgroup only ever waits on items already claimed).  Each instance's
packed passes are emulated and compared with the oracle."""

import numpy as np
import pytest

from aleph_amd import workloads as W
from aleph_amd.plan_tools import HostBatch
from tests.helpers import max_abs
from tests.plan_emulator import emulate


def prepared_state(n, plus):
    """The batch reset's state: 2^(-|plus|/2) on every index with no bit outside `plus`."""
    x = np.arange(1 << n, dtype=np.uint64)
    psi = np.where((x & np.uint64(~plus & ((1 << n) - 1))) == 0, 1.0, 0.0).astype(np.complex128)
    return psi / np.sqrt(float(1 << bin(plus).count("1")))


def test_leading_hadamards_become_the_reset():
    """The QAOA H layer is not planned as passes: instance plans lose the
    passes that applied it (n=16, 12 local qubits: 11 -> 9 passes at p=4)
    and record every qubit in the instance's |+> mask."""
    graphs, gammas, betas = W.qaoa_batch(4, seed=3, n_min=16, n_max=16, p=4)
    hb = HostBatch(graphs, gammas, betas, precision=32, local_qubits=12)
    assert hb.plus_masks == [(1 << 16) - 1] * 4
    from aleph_amd.plan_tools import HostPlan
    for i, (n, edges) in enumerate(graphs):
        full = HostPlan(W.qaoa_maxcut(n, edges, gammas[i], betas[i]), precision=32, local_qubits=12)
        assert len(hb.instance_passes[i]) < full.stats["n_passes"]


@pytest.mark.parametrize("local_qubits,cpi,n_max", [(0, 0, 22), (9, 1, 16), (11, 4, 22), (13, 64, 24)])
def test_batch_items_cover_and_order(local_qubits, cpi, n_max):
    graphs, gammas, betas = W.qaoa_batch(40, seed=17, n_min=6, n_max=n_max, p=4)
    hb = HostBatch(graphs, gammas, betas, precision=32, local_qubits=local_qubits, chunks_per_item=cpi)
    sizes = [1 << n for n, _ in graphs]
    assert hb.total_amps == sum(sizes)
    assert hb.offsets == list(np.cumsum([0] + sizes[:-1]))
    max_m = max(p.m for p in hb.passes)
    assert hb.threads >= (1 << (max_m - 4)) and hb.threads in (256, 512)
    assert hb.lds_bytes == 8 << max_m
    seen = {}
    done_before = {}
    position = {}
    for k, it in enumerate(hb.items):
        n = hb.num_qubits[it.inst]
        p = hb.passes[it.pass_]
        assert p.m <= n and p.reg_bits == 4
        chunks = 1 << (n - p.m)
        assert it.state_off == hb.offsets[it.inst]
        assert 0 < it.chunk_count and it.chunk_begin + it.chunk_count <= chunks
        key = (it.inst, it.pass_)
        seen.setdefault(key, set()).update(range(it.chunk_begin, it.chunk_begin + it.chunk_count))
        position.setdefault(key, []).append(k)
    for i, plist in enumerate(hb.instance_passes):
        total = 0
        for r, g in enumerate(plist):
            chunks = 1 << (hb.num_qubits[i] - hb.passes[g].m)
            assert seen[(i, g)] == set(range(chunks))
            for k in position[(i, g)]:
                assert hb.items[k].dep_target == total
                if r:
                    assert max(position[(i, plist[r - 1])]) < k     # dependencies queue first
            total += chunks
    assert hb.round_begin[0] == 0 and hb.round_begin[-1] == len(hb.items)


def test_batch_instances_emulate_to_oracle(oracle_lib):
    graphs, gammas, betas = W.qaoa_batch(6, seed=23, n_min=6, n_max=12, p=3)
    for prec, tol in ((64, 1e-12), (32, 1e-5)):
        hb = HostBatch(graphs, gammas, betas, precision=prec, local_qubits=7)
        for i, (n, edges) in enumerate(graphs):
            c = W.qaoa_maxcut(n, edges, gammas[i], betas[i])
            assert hb.plus_masks[i] == (1 << n) - 1      # the H layer became the reset's |+>^n
            psi0 = prepared_state(n, hb.plus_masks[i])
            o = oracle_lib.OracleSv()
            st = o.allocate(n)
            o.run(c, st)
            got = emulate(hb.instance_plan(i), psi0, precision=prec)
            assert max_abs(got, st.amplitudes()) < tol, (prec, i)
