"""CPU emulation of a plan's device semantics, for host-side planner tests.

Each device op is applied to the whole state vector with numpy, reading its
qubits back from the stage descriptors exactly as the kernel does (register
bit g -> physical mask rg[g]; thread-side masks as stored).  Agreement with
the oracle checks gate lowering, fusion, stage assignment, register-bit
mapping, 2-qubit matrix re-orientation and phase-term classes on the CPU;
chunk addressing and LDS movement are left to the GPU parity tests.
"""

from __future__ import annotations

import numpy as np

from aleph_amd.plan_tools import HostPlan


def _bit(mask: int) -> int:
    assert mask and mask & (mask - 1) == 0, mask
    return mask.bit_length() - 1


def _apply_matrix(psi, n, qubits, m, cmask):
    """Dense k-qubit matrix on qubits (MSB first) where the control mask holds."""
    k = len(qubits)
    idx = np.arange(1 << n, dtype=np.int64)
    out = psi.copy()
    tmask = sum(1 << q for q in qubits)
    base_sel = ((idx & tmask) == 0) & ((idx & cmask) == cmask)
    bases = idx[base_sel]
    offs = []
    for s in range(1 << k):
        o = 0
        for j, q in enumerate(qubits):
            if (s >> (k - 1 - j)) & 1:
                o |= 1 << q
        offs.append(o)
    vals = np.stack([psi[bases | o] for o in offs])      # (2^k, groups)
    new = m @ vals
    for s, o in enumerate(offs):
        out[bases | o] = new[s]
    return out


def emulate(plan: HostPlan, psi: np.ndarray, precision: int = 64) -> np.ndarray:
    n = plan.num_qubits
    psi = np.asarray(psi, dtype=np.complex128).copy()
    idx = np.arange(1 << n, dtype=np.int64)
    for st in plan.steps:
        if st.kind == 0:
            p = plan.passes[st.pass_]
            G = p.reg_bits
            for sd in plan.stages[p.stage_begin:p.stage_begin + p.n_stages]:
                rg = [int(sd.rg[g]) for g in range(G)]
                for oi in range(sd.op_begin, sd.op_end):
                    op = plan.ops[oi]
                    raw = np.frombuffer(bytes(op.m), dtype=np.float64 if precision == 64 else np.float32)
                    m = raw.astype(np.float64)[:32].view(np.complex128)
                    if op.type == 1:
                        psi = _apply_matrix(psi, n, [_bit(rg[op.a])], m[:4].reshape(2, 2), int(op.cmask))
                    elif op.type == 2:
                        psi = _apply_matrix(psi, n, [_bit(rg[op.a]), _bit(rg[op.b])], m[:16].reshape(4, 4),
                                            int(op.cmask))
                    elif op.type == 3:
                        hi = (idx & int(op.ma)) != 0
                        lo = (idx & int(op.mb)) != 0
                        d = m[:4][hi.astype(int) * 2 + lo.astype(int)]
                        on = (idx & int(op.cmask)) == int(op.cmask)
                        psi = np.where(on, psi * d, psi)
                    elif op.type == 4:
                        mask = int(op.cmask) | sum(rg[g] for g in range(G) if (op.a >> g) & 1)
                        psi = np.where((idx & mask) == mask, -psi, psi)
                    elif op.type == 5:
                        phi = np.zeros(1 << n)
                        for t in plan.terms[op.term_off: op.term_off + op.n_terms + op.b]:
                            mask = int(t.mask_out) | sum(rg[g] for g in range(G) if (t.rmask >> g) & 1)
                            phi += np.where((idx & mask) == mask, t.angle, 0.0)
                        psi = psi * np.exp(1j * phi)
                    else:
                        raise AssertionError(f"unknown op type {op.type}")
        elif st.kind == 1:      # dense k-qubit block, matrix in the plan payload
            k = st.k
            dt = np.float64 if precision == 64 else np.float32
            raw = np.frombuffer(plan.payload, dtype=dt, count=2 * (1 << (2 * k)), offset=st.dev_off)
            m = raw.astype(np.float64).view(np.complex128).reshape(1 << k, 1 << k)
            psi = _apply_matrix(psi, n, list(st.qubits[:k]), m, int(st.cmask))
        elif st.kind == 2:      # diagonal
            k = st.k
            d = np.frombuffer(plan.payload, dtype=np.float64, count=2 << k, offset=st.dev_off).view(np.complex128)
            sel = np.zeros(1 << n, dtype=np.int64)
            for q in st.qubits[:k]:
                sel = (sel << 1) | ((idx >> q) & 1)
            on = (idx & int(st.cmask)) == int(st.cmask)
            psi = np.where(on, psi * d[sel], psi)
        elif st.kind == 3:      # parity phase polynomial
            nt = st.count
            ang = np.frombuffer(plan.payload, dtype=np.float64, count=nt, offset=st.dev_off)
            off = np.frombuffer(plan.payload, dtype=np.uint32, count=nt + 1, offset=st.dev_off3)
            conds = np.frombuffer(plan.payload, dtype=np.uint64, count=int(off[-1]), offset=st.dev_off2)
            phi = np.zeros(1 << n)
            for t in range(nt):
                ok = np.ones(1 << n, dtype=bool)
                for c in conds[off[t]:off[t + 1]]:
                    par = np.zeros(1 << n, dtype=np.int64)
                    v = idx & int(c)
                    while np.any(v):
                        par ^= v & 1
                        v = v >> 1
                    ok &= par == 1
                phi += np.where(ok, ang[t], 0.0)
            psi = psi * np.exp(1j * phi)
        elif st.kind in (4, 5):
            d = 2 if st.kind == 4 else 4
            m = np.array(st.mat8[:2 * d * d]).view(np.complex128).reshape(d, d)
            qubits = [st.qubits[0]] if d == 2 else [st.qubits[0], st.qubits[1]]
            psi = _apply_matrix(psi, n, qubits, m, int(st.cmask))
        elif st.kind == 7:      # 1q gates on tile qubits, then the involutive permutation (bit q -> sigma[q])
            sigma = np.frombuffer(plan.payload, dtype=np.uint32, count=n, offset=st.dev_off)
            gq = np.frombuffer(plan.payload, dtype=np.uint32, count=st.k, offset=st.dev_off2)
            gm = np.frombuffer(plan.payload, dtype=np.float64, count=8 * st.k, offset=st.dev_off3)
            for g in range(st.k):
                m = gm[8 * g:8 * g + 8].view(np.complex128).reshape(2, 2)
                psi = _apply_matrix(psi, n, [int(gq[g])], m, 0)
            dst = np.zeros(1 << n, dtype=np.int64)
            for q in range(n):
                dst |= ((idx >> q) & 1) << int(sigma[q])
            out = np.empty_like(psi)
            out[dst] = psi
            psi = out
        else:
            raise AssertionError(f"unknown step kind {st.kind}")
    return psi
