"""Host-side plan inspection: plan a circuit without a device and read the
descriptor arrays (layouts of aleph_amd/csrc/plan.hpp and engine.hpp)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

u32, u64, f64 = C.c_uint32, C.c_uint64, C.c_double

STEP_KINDS = ["fused", "kq", "diag", "poly", "1q", "2q", "swap", "bitperm"]
OP_NAMES = {1: "1q", 2: "2q", 3: "diag", 4: "sign", 5: "phase"}


class PassDev(C.Structure):
    _fields_ = [("m", u32), ("n_stages", u32), ("stage_begin", u32), ("features", u32), ("local_pos", u32 * 14),
                ("reg_bits", u32), ("pad2", u32)]


class StageDev(C.Structure):
    _fields_ = [("op_begin", u32), ("op_end", u32), ("n_thread_bits", u32), ("pad", u32), ("tg", u64 * 10),
                ("tl", u32 * 10), ("pad2", u32 * 2), ("rg", u64 * 5), ("rl", u32 * 5), ("pad3", u32),
                ("go", u64 * 32), ("ls", u32 * 32), ("lb", u32 * 32)]


class DevOp(C.Structure):
    _fields_ = [("type", u32), ("a", u32), ("b", u32), ("n_terms", u32), ("cmask", u64), ("ma", u64), ("mb", u64),
                ("term_off", u32), ("pad", u32), ("m", f64 * 32)]


class PhaseTerm(C.Structure):
    _fields_ = [("mask_out", u64), ("angle", f64), ("cls", u32), ("rmask", u32), ("c", f64), ("s", f64),
                ("pad", u64)]


class Step(C.Structure):
    _fields_ = [("kind", u32), ("pass_", u32), ("m", u32), ("n_stages", u32), ("k", u32), ("qubits", u32 * 8),
                ("cmask", u64), ("mat8", f64 * 32), ("dev_off", C.c_size_t), ("dev_off2", C.c_size_t),
                ("dev_off3", C.c_size_t), ("count", u32)]


class PlanView(C.Structure):
    _fields_ = [("steps", C.c_void_p), ("n_steps", u32), ("step_bytes", u32),
                ("passes", C.c_void_p), ("n_passes", u32), ("pass_bytes", u32),
                ("stages", C.c_void_p), ("n_stages", u32), ("stage_bytes", u32),
                ("ops", C.c_void_p), ("n_ops", u32), ("op_bytes", u32),
                ("terms", C.c_void_p), ("n_terms", u32), ("term_bytes", u32),
                ("payload", C.c_void_p), ("payload_bytes", u64)]


def _array(ptr, count, cls, nbytes):
    if C.sizeof(cls) != nbytes:
        raise RuntimeError(f"{cls.__name__} layout mismatch: {C.sizeof(cls)} != {nbytes}")
    if count == 0:
        return []
    arr = (cls * count).from_address(ptr)
    return [arr[i] for i in range(count)]


class HostPlan:
    """A plan built on the host; arrays are copied out on construction."""

    def __init__(self, circuit, precision: int = 64, planner: int = L.PLANNER_RULE, local_qubits: int = 0,
                 max_ops: int = 0, coalesce_bits: int = 0, fuse_1q: bool = True, reg_bits: int = 0, nontemporal: bool = False,
                fuse_2q: bool = True, fuse_diag: bool = True, kq_width: int = 0, flags: int = 0):
        lib = L.load()
        words, reals = circuit.encode()
        nw, nr = int(words.size), int(reals.size)
        words = np.ascontiguousarray(words if nw else np.zeros(1, np.uint32))
        reals = np.ascontiguousarray(reals if nr else np.zeros(1))
        opts = L.PlanOpts(planner, local_qubits, max_ops, coalesce_bits, int(fuse_1q), reg_bits,
                          (L.PLAN_NONTEMPORAL if nontemporal else 0) | (0 if fuse_2q else L.PLAN_NO_FUSE_2Q)
                          | (0 if fuse_diag else L.PLAN_NO_FUSE_DIAG) | flags, kq_width)
        h = C.c_void_p()
        err = L.Error()
        code = lib.aleph_hip_plan_build_host(circuit.num_qubits, precision, words.ctypes.data_as(L.PU32), nw,
                                             reals.ctypes.data_as(L.PF64), nr, C.byref(opts), C.byref(h),
                                             C.byref(err))
        if code != L.OK:
            raise RuntimeError(f"plan failed with status {code}: {err.message.decode()}")
        try:
            v = PlanView()
            lib.aleph_hip_plan_view_get(h, C.byref(v))
            self.steps = [Step.from_buffer_copy(s) for s in _array(v.steps, v.n_steps, Step, v.step_bytes)]
            self.passes = [PassDev.from_buffer_copy(s) for s in _array(v.passes, v.n_passes, PassDev, v.pass_bytes)]
            self.stages = [StageDev.from_buffer_copy(s) for s in _array(v.stages, v.n_stages, StageDev,
                                                                         v.stage_bytes)]
            self.ops = [DevOp.from_buffer_copy(s) for s in _array(v.ops, v.n_ops, DevOp, v.op_bytes)]
            self.terms = [PhaseTerm.from_buffer_copy(s) for s in _array(v.terms, v.n_terms, PhaseTerm,
                                                                         v.term_bytes)]
            self.payload = C.string_at(v.payload, v.payload_bytes) if v.payload_bytes else b""
            nrows = C.c_uint32()
            lib.aleph_hip_plan_features(h, None, None, 0, C.byref(nrows))
            self.features = np.zeros((nrows.value, L.MODEL_INPUTS), dtype=np.float32)
            self.feature_launch = np.zeros(nrows.value, dtype=np.uint32)
            lib.aleph_hip_plan_features(h, self.features.ctypes.data_as(L.PF32),
                                        self.feature_launch.ctypes.data_as(L.PU32), nrows.value, C.byref(nrows))
            self.jit_sources = []
            for i in range(len(self.passes)):
                n = C.c_size_t()
                lib.aleph_hip_plan_jit_source(h, i, None, 0, C.byref(n))
                buf = C.create_string_buffer(n.value + 1)
                lib.aleph_hip_plan_jit_source(h, i, buf, n.value + 1, C.byref(n))
                self.jit_sources.append(buf.value.decode())
            st = L.PlanStats()
            lib.aleph_hip_plan_stats_get(h, C.byref(st))
            self.stats = {f: getattr(st, f) for f, _ in L.PlanStats._fields_ if f != "reserved"}
        finally:
            lib.aleph_hip_plan_free(h)
        self.num_qubits = circuit.num_qubits
        self.precision = precision

    def summary(self) -> list[dict]:
        out = []
        for s in self.steps:
            d = {"kind": STEP_KINDS[s.kind]}
            if s.kind == 0:
                p = self.passes[s.pass_]
                stages = self.stages[p.stage_begin:p.stage_begin + p.n_stages]
                d.update(m=p.m, chunk=list(p.local_pos[:p.m]), stages=p.n_stages,
                         ops=[[OP_NAMES[self.ops[o].type] for o in range(sd.op_begin, sd.op_end)] for sd in stages])
            out.append(d)
        return out


class BatchItem(C.Structure):
    _fields_ = [("state_off", u64), ("pass_", u32), ("inst", u32), ("chunk_begin", u32), ("chunk_count", u32),
                ("dep_target", u32), ("pad", u32)]


class _Steps:
    """A minimal plan-like view of one batch instance (its fused passes in order)."""

    def __init__(self, hb: "HostBatch", i: int):
        self.num_qubits = hb.num_qubits[i]
        self.passes, self.stages, self.ops, self.terms, self.payload = hb.passes, hb.stages, hb.ops, hb.terms, b""
        self.steps = []
        for g in hb.instance_passes[i]:
            s = Step()
            s.kind = 0
            s.pass_ = g
            self.steps.append(s)


class HostBatch:
    """A QAOA batch planned on the host (no device); packed arrays copied out."""

    def __init__(self, graphs, gammas, betas, precision: int = 32, mode: int = 0, local_qubits: int = 0,
                 chunks_per_item: int = 0):
        lib = L.load()
        nq = np.ascontiguousarray(np.array([g[0] for g in graphs], dtype=np.uint32))
        offs = np.zeros(len(graphs) + 1, dtype=np.uint64)
        flat = []
        for i, (_, edges) in enumerate(graphs):
            offs[i + 1] = offs[i] + len(edges)
            for u, v in edges:
                flat += [u, v]
        flat = np.ascontiguousarray(np.array(flat or [0], dtype=np.uint32))
        ga = np.ascontiguousarray(np.asarray(gammas, dtype=np.float64).reshape(len(graphs), -1))
        be = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(len(graphs), -1))
        opts = L.BatchOpts(mode, local_qubits, chunks_per_item, 0, 0, 0, 0, 0)
        h = C.c_void_p()
        err = L.Error()
        code = lib.aleph_hip_batch_build_host_qaoa(len(graphs), nq.ctypes.data_as(L.PU32),
                                                   offs.ctypes.data_as(L.PU64), flat.ctypes.data_as(L.PU32),
                                                   ga.shape[1], ga.ctypes.data_as(L.PF64), be.ctypes.data_as(L.PF64),
                                                   precision, C.byref(opts), C.byref(h), C.byref(err))
        if code != L.OK:
            raise RuntimeError(f"batch plan failed with status {code}: {err.message.decode()}")
        try:
            v = L.BatchView()
            lib.aleph_hip_batch_view_get(h, C.byref(v))
            self.items = [BatchItem.from_buffer_copy(s) for s in _array(v.items, v.n_items, BatchItem, v.item_bytes)]
            self.passes = [PassDev.from_buffer_copy(s) for s in _array(v.passes, v.n_passes, PassDev, v.pass_bytes)]
            self.stages = [StageDev.from_buffer_copy(s) for s in _array(v.stages, v.n_stages, StageDev,
                                                                         v.stage_bytes)]
            self.ops = [DevOp.from_buffer_copy(s) for s in _array(v.ops, v.n_ops, DevOp, v.op_bytes)]
            self.terms = [PhaseTerm.from_buffer_copy(s) for s in _array(v.terms, v.n_terms, PhaseTerm,
                                                                         v.term_bytes)]
            self.offsets = [int(v.offsets[i]) for i in range(v.n_instances)]
            self.round_begin = [int(v.round_begin[i]) for i in range(v.n_rounds + 1)]
            self.threads, self.lds_bytes, self.total_amps = v.threads, v.lds_bytes, v.total_amps
            self.plus_masks = [int(v.plus_masks[i]) for i in range(v.n_instances)]
        finally:
            lib.aleph_hip_batch_free(h)
        self.num_qubits = nq.tolist()
        self.precision = precision
        self.instance_passes = [[] for _ in graphs]
        for it in self.items:
            if not self.instance_passes[it.inst] or self.instance_passes[it.inst][-1] != it.pass_:
                self.instance_passes[it.inst].append(it.pass_)

    def instance_plan(self, i: int):
        return _Steps(self, i)
