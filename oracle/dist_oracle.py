"""CPU references for the distributed state vector — TEST INFRASTRUCTURE ONLY
(used by tests/; the HIP product never imports it).

* `exchange_cpu` restates crates/aleph-sv/src/dist_ref.rs:78-104: chunk c of
  rank r moves to rank r with the chosen global bits set to c, at the chunk
  index made of r's old values of those bits.
* `swap_bits` is the same exchange as the mathematics states it: a
  permutation of the physical index that swaps bit global_bits[j] with bit
  m - k + j.  Tests check the two forms against each other.
* `run_physical` executes a DistPlan on one full 2^n physical state with the
  pinned C oracle: a Local step applies its (physical) instructions to the
  whole state, which is what running the rank-specialised instructions on
  every rank slice must equal; an Exchange is `swap_bits`.  The result is
  returned in logical order through the plan's final map (dist_ref.rs:63-72).
"""

from __future__ import annotations

import numpy as np

from . import oracle as orc


def exchange_cpu(ranks: list, m: int, global_bits) -> list:
    k = len(global_bits)
    chunk = 1 << (m - k)
    out = [r.copy() for r in ranks]
    for r, src in enumerate(ranks):
        for c in range(1 << k):
            dst_rank, dst_chunk = r, 0
            for j, gb in enumerate(global_bits):
                rb = gb - m
                dst_chunk |= ((r >> rb) & 1) << j
                dst_rank = (dst_rank & ~(1 << rb)) | (((c >> j) & 1) << rb)
            out[dst_rank][dst_chunk * chunk:(dst_chunk + 1) * chunk] = src[c * chunk:(c + 1) * chunk]
    return out


def swap_bits(state: np.ndarray, n: int, m: int, global_bits) -> np.ndarray:
    k = len(global_bits)
    x = np.arange(1 << n, dtype=np.int64)
    y = x.copy()
    for j, gb in enumerate(global_bits):
        lb = m - k + j
        bg, bl = (x >> gb) & 1, (x >> lb) & 1
        y = (y & ~((1 << gb) | (1 << lb))) | (bg << lb) | (bl << gb)
    out = np.empty_like(state)
    out[y] = state[x]
    return out


def run_physical(plan, n: int) -> np.ndarray:
    o = orc.OracleSv()
    st = o.allocate(n)
    m = plan.layout.m()
    for step in plan.steps:
        if hasattr(step, "global_bits"):
            st.load(swap_bits(st.amplitudes(), n, m, step.global_bits))
            continue
        for inst in step:
            kind = type(inst).__name__
            if kind == "GateInstance":
                o.apply_gate(st, inst)
            elif kind == "DiagonalPhase":
                o.apply_diagonal_phase(st, inst)
    phys = st.amplitudes()
    x = np.arange(1 << n, dtype=np.int64)
    p = np.zeros_like(x)
    for lq, pq in enumerate(plan.final_map):
        p |= ((x >> lq) & 1) << pq
    return phys[p]
