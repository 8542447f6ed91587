"""Benchmark: dense state-vector simulation on one MI355X.

Default workload (one step = one whole circuit, state reset to |0...0>
included): the seeded random brickwall — U3(θ,φ,λ) on every qubit each
layer, CZ on alternating pairs — at n=30 qubits, depth 20, complex128
amplitudes, planned by the rule-based pass planner, each fused pass run as a
kernel generated for that pass and compiled with hiprtc at plan creation
(`--no-jit`: the compiled op-loop kernel).  Also reported on the
same line: the FP32 run of the same circuit, the n=30 QFT on a Gaussian
state, the event-timed roofline of the fused pass kernel, and the CPU
oracle (a C++ restatement of the reference CPU `sv` path and its final-state
pipeline) running the same circuit whole on this host's cores, with
single-thread runs of two reference corpus circuits beside their published
times.

`python bench.py --gpus 1 --steps K --warmup W` prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64, vector and matrix alike
FMA_PER_AMP = {1: 8, 2: 16, 3: 4}  # OP_1Q (2x2), OP_2Q (4x4), OP_DIAG: real FMAs per amplitude


def fused_pass_flops(circuit, precision: int) -> float:
    """Algorithmic FLOPs (2 per real FMA) of the plan's fused passes, averaged
    per launch: a 2x2 complex matrix costs 2 complex multiply-adds = 8 real
    FMAs per amplitude, a 4x4 one 16, a diagonal 4; a controlled op counts
    only the fraction 2^-c of amplitudes its c controls select; sign flips
    and phase polynomials are counted as zero."""
    from aleph_amd.plan_tools import HostPlan
    hp = HostPlan(circuit, precision=precision)
    fma = 0.0
    for o in hp.ops:
        fma += FMA_PER_AMP.get(o.type, 0) * 0.5 ** bin(o.cmask).count("1")
    return 2.0 * fma * float(1 << circuit.num_qubits) / max(1, len(hp.passes))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--local-qubits", type=int, default=0)
    ap.add_argument("--max-ops", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the FP32 and QFT legs")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 traffic measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched QAOA leg")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--no-noise", action="store_true", help="skip the noise-trajectory leg")
    ap.add_argument("--no-dist", action="store_true", help="skip the rank-slice exchange leg")
    ap.add_argument("--noise-shots", type=int, default=100_000)
    ap.add_argument("--batch-probe", action="store_true",
                    help="only time each size class of the QAOA batch alone, per executor (one JSON line each)")
    ap.add_argument("--no-jit", action="store_true",
                    help="run fused passes with the compiled op-loop kernel instead of per-pass generated kernels")
    ap.add_argument("--traffic-child", action="store_true",
                    help="only reset, execute one plan and take the norm (the process rocprofv3 profiles)")
    ap.add_argument("--child-precision", type=int, default=64, choices=(32, 64))
    ap.add_argument("--child-workload", default="brickwall",
                    choices=("brickwall", "qft", "noise", "exchange", "qaoa_batch"))
    ap.add_argument("--child-kq", type=int, default=0, help="plan the child's circuit into dense blocks of this width")
    return ap.parse_args()


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def time_circuit(backend, circuit, steps: int, warmup: int, **plan_kw):
    """Returns (ms per step, plan stats, per-launch ms of the last step, kinds,
    norm, state); plan stats gain `compile_s`, the wall time of plan creation
    (including kernel generation when asked for)."""
    t_plan = time.perf_counter()
    plan = backend.compile(circuit, **plan_kw)
    compile_s = time.perf_counter() - t_plan
    st = backend.allocate(circuit.num_qubits)
    import aleph_amd._lib as L
    lib = L.load()
    for _ in range(warmup):
        backend._check(lib.aleph_hip_state_reset(backend._ctx, st.handle))
        plan.execute(st)
    backend.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        backend._check(lib.aleph_hip_state_reset(backend._ctx, st.handle))
        plan.execute(st)
    backend.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / max(steps, 1)
    backend._check(lib.aleph_hip_state_reset(backend._ctx, st.handle))
    launch_ms, kinds = plan.execute_timed(st)
    norm = backend.norm_sqr(st)
    stats = plan.stats()
    stats["compile_s"] = compile_s
    plan.free()
    return ms, stats, launch_ms, kinds, norm, st


# Published single-thread times of the reference's CPU `sv` path
# (run_optimized, AMD EPYC 8124P; docs/perf/phase4.md:27, 44).
PUBLISHED_1T_MS = {"qft_n20": 320.03, "sycamore_n20_d20": 6028.89}


def _cpu_threads() -> int:
    """The host threads the CPU leg may use: OMP_NUM_THREADS when set (the
    GPU hosts cap a job at their CPU share), else every visible core."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = os.cpu_count() or 1
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


def _cpu_run(circuit, label: str) -> dict:
    """The reference's CPU final-state path restated (oracle/pipeline.py:
    RelabelQubits, CancelInversePairs, Fuse1qRuns, Fuse2q, FuseKq(4),
    TileBlock(15), as final_state_cpu_pipeline, crates/aleph-ir/src/passes/
    mod.rs:192-200), then the oracle's OpenMP kernels per instruction, the
    whole list from |0...0> to the final (un-permuted) state: planning,
    execution and allocation timed separately."""
    from oracle import oracle as orc
    from oracle import pipeline as P
    t = time.perf_counter()
    instrs, perm = P.cpu_pipeline(circuit)
    plan_s = time.perf_counter() - t
    o = orc.OracleSv(cap=64)
    t = time.perf_counter()
    st = o.allocate(circuit.num_qubits)
    alloc_s = time.perf_counter() - t
    t0 = time.perf_counter()
    last = t0
    for i, ins in enumerate(instrs):
        P.apply(o, st, [ins], perm=None)
        now = time.perf_counter()
        if now - last > 20.0:
            log(f"bench: cpu {label}: {i + 1}/{len(instrs)} instructions, {now - t0:.0f} s")
            last = now
    if perm is not None:
        o.unpermute_state(st, perm)
    exec_s = time.perf_counter() - t0
    del st
    return {"exec_s": exec_s, "plan_s": plan_s, "alloc_s": alloc_s, "gates": len(circuit.gates()),
            "instructions": len(instrs), "relabelled": perm is not None}


def cpu_baseline(circuit) -> dict:
    """The benchmark circuit run whole on the host's cores (measured, not
    extrapolated), plus single-thread runs of the reference's own corpus
    circuits qft_n20 and sycamore_n20_d20 beside their published times."""
    from oracle import oracle as orc
    from aleph_amd import workloads as W
    threads = _cpu_threads()
    orc.set_threads(threads)
    full = _cpu_run(circuit, f"n={circuit.num_qubits}")
    log(f"bench: cpu full run {full['exec_s']:.1f} s on {threads} threads")
    calib = {}
    orc.set_threads(1)
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "corpus")
    for name, pub in PUBLISHED_1T_MS.items():
        r = _cpu_run(W.load_qasm(os.path.join(golden, f"{name}.qasm")), name)
        calib[name] = {"ms_1thread": r["exec_s"] * 1e3, "published_ms_1thread": pub,
                       "ratio_to_published": r["exec_s"] * 1e3 / pub, "instructions": r["instructions"]}
        log(f"bench: cpu {name} 1 thread {r['exec_s'] * 1e3:.0f} ms (published {pub} ms)")
    orc.set_threads(threads)
    return {
        "value": full["exec_s"],
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"the whole circuit, measured: reference CPU final-state pipeline restated (RelabelQubits, "
                   f"CancelInversePairs, Fuse1qRuns, Fuse2q, FuseKq(4), TileBlock(15)) turns {full['gates']} gates "
                   f"into {full['instructions']} instructions (relabelled: {full['relabelled']}), all applied to one "
                   f"n={circuit.num_qubits} state from |0...0> on {threads} threads; planning "
                   f"{full['plan_s']:.2f} s and allocation {full['alloc_s']:.1f} s not included"),
        "plan_s": full["plan_s"],
        "alloc_s": full["alloc_s"],
        "calibration_1thread": calib,
    }


FUSED_KERNEL = "k_fused_pass<double"   # compiled op-loop kernel
JIT_KERNEL = "aleph_pass_f64"           # per-pass generated kernels


def traffic_child(args) -> None:
    """Run under rocprofv3 (--pmc or --kernel-trace): one |0..0> reset
    (k_init_basis writes the whole state), one execution of the plan, one
    norm (k_reduce_branch reads the whole state).  The two full-state kernels
    calibrate the counters.  --child-precision and --child-workload pick the
    plan (FP64 brickwall by default), --child-kq plans it into dense blocks
    (matrix-core sweeps for widths 4..6); the noise, exchange and qaoa_batch
    workloads run those legs' kernels once instead."""
    from aleph_amd import workloads as W
    from aleph_amd.backend import HipSvBackend
    if args.child_workload == "noise":
        from aleph_amd import noise as N
        c, m = W.noisy_brickwall(NOISE_N, NOISE_DEPTH, seed=args.seed)
        N.outcomes(c, m, args.noise_shots, 7, backend=HipSvBackend(seed=1))
        return
    if args.child_workload == "exchange":
        from aleph_amd import dist as D
        layout = D.DistLayout(args.n, 3)
        dx = D.DistSvBackend(HipSvBackend(seed=1))
        ranks = dx.alloc_ranks(layout)
        for k in (1, 2, 3):
            dx.exchange(ranks, layout, list(range(args.n - 1, args.n - 1 - k, -1)))
        return
    if args.child_workload == "qaoa_batch":
        from aleph_amd.batch import HipBatch
        graphs, ga, be = W.qaoa_batch(args.batch_count, seed=args.seed)
        bt = HipBatch.qaoa(HipSvBackend(seed=1, precision=32), graphs, ga, be, precision=32, mode=1,
                           local_qubits=12, chunks_per_item=16)
        bt.run()
        bt.run()
        bt.expect_maxcut()
        bt.free()
        return
    b = HipSvBackend(seed=1, precision=args.child_precision)
    c = W.qft(args.n) if args.child_workload == "qft" else W.brickwall(args.n, args.depth, seed=args.seed)
    plan = b.compile(c, local_qubits=args.local_qubits, max_ops=args.max_ops, jit=not args.no_jit,
                     kq_width=args.child_kq)
    st = b.allocate(args.n)
    plan.execute(st)
    b.norm_sqr(st)
    b.synchronize()


def _pmc_pass(counter: str, args, out_dir: str) -> dict:
    """Per-kernel mean of one counter (KiB per dispatch) from a rocprofv3 pass."""
    import csv
    import subprocess
    cmd = ["timeout", "-s", "KILL", "150", "rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out_dir,
           "-o", counter, "--", sys.executable, os.path.abspath(__file__), "--traffic-child", "--n", str(args.n),
           "--depth", str(args.depth), "--seed", str(args.seed), "--local-qubits", str(args.local_qubits),
           "--max-ops", str(args.max_ops)] + (["--no-jit"] if args.no_jit else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd="/tmp")
    sums, counts = {}, {}
    for root, _, files in os.walk(out_dir):
        for f in files:
            if f.endswith("counter_collection.csv"):
                for row in csv.DictReader(open(os.path.join(root, f))):
                    if row["Counter_Name"] != counter:
                        continue
                    name = row["Kernel_Name"]
                    key = (name, row["Dispatch_Id"])
                    sums[key] = sums.get(key, 0.0) + float(row["Counter_Value"])
    per_kernel = {}
    for (name, _), v in sums.items():
        per_kernel.setdefault(name, []).append(v)
    return {k: float(np.mean(v)) for k, v in per_kernel.items()}


def measure_traffic(args, state_bytes: float):
    """HBM bytes per fused-pass launch from two rocprofv3 passes (FETCH_SIZE,
    WRITE_SIZE), each scaled so that the full-state kernels read / write
    exactly the state: the write factor from k_init_basis, the fetch factor
    from k_reduce_branch."""
    import shutil
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None
    tmp = tempfile.mkdtemp(prefix="aleph_pmc_")
    try:
        fetch = _pmc_pass("FETCH_SIZE", args, os.path.join(tmp, "fetch"))
        write = _pmc_pass("WRITE_SIZE", args, os.path.join(tmp, "write"))
    except Exception as e:  # profiler unavailable or refused: no traffic figure
        log(f"bench: traffic measurement skipped ({e})")
        return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)

    def pick(d, key):
        vals = [v for k, v in d.items() if key in k]
        return float(np.mean(vals)) if vals else None
    fused_name = FUSED_KERNEL if args.no_jit else JIT_KERNEL
    f_fused, w_fused = pick(fetch, fused_name), pick(write, fused_name)
    f_cal, w_cal = pick(fetch, "k_reduce_branch<double"), pick(write, "k_init_basis<double")
    if None in (f_fused, w_fused, f_cal, w_cal) or f_cal <= 0 or w_cal <= 0:
        return None
    kib = state_bytes / 1024.0
    f_scale, w_scale = kib / f_cal, kib / w_cal
    return {"bytes": (f_fused * f_scale + w_fused * w_scale) * 1024.0,
            "fetch_bytes": f_fused * f_scale * 1024.0, "write_bytes": w_fused * w_scale * 1024.0,
            "fetch_scale": f_scale, "write_scale": w_scale}


def batch_leg(args) -> dict:
    """Batched QAOA MaxCut: seeded instances, n even in [12, 24], p = 4,
    FP32 amplitudes, <C> in FP64; tile shape and executor autotuned on every
    8th instance, then the whole batch timed."""
    from aleph_amd import workloads as W
    from aleph_amd.backend import HipSvBackend
    from aleph_amd.batch import HipBatch
    graphs, ga, be = W.qaoa_batch(args.batch_count, seed=args.seed)
    hip = HipSvBackend(seed=1, precision=32)
    sub = list(range(0, len(graphs), 8))
    tuner = HipBatch.qaoa(hip, [graphs[i] for i in sub], ga[sub], be[sub], precision=32)
    cands = tuner.autotune(reps=2)
    tuner.free()
    best = min(cands, key=lambda r: r["ms"])
    b = HipBatch.qaoa(hip, graphs, ga, be, precision=32, mode=best["mode"], local_qubits=best["local_qubits"],
                      chunks_per_item=best["chunks_per_item"])
    b.run()
    times = [b.run_timed() for _ in range(3)]
    ms = min(times)
    cvals = b.expect_maxcut()
    st = b.stats()
    b.free()
    return {"instances": len(graphs), "ms": ms, "instances_per_s": len(graphs) / (ms * 1e-3),
            "hbm_gbs": st["sweep_bytes"] / (ms * 1e-3) / 1e9, "sweep_bytes": st["sweep_bytes"],
            "passes": st["n_passes"], "items": st["n_items"], "solo_instances": st["solo_instances"],
            "pool_gb": st["pool_bytes"] / 1e9,
            "plan_ms": st["plan_ms"], "mean_cut": float(np.mean(cvals)),
            "tuned": {"mode": ["persistent", "rounds"][best["mode"]], "local_qubits": best["local_qubits"],
                      "chunks_per_item": best["chunks_per_item"]},
            "config": "QAOA MaxCut, seeded random 3-regular graphs, n even in [12,24], p=4, FP32, <C> FP64"}


NOISE_N, NOISE_DEPTH = 12, 10


def noise_leg(args) -> dict:
    """Noise trajectories: the seeded 12-qubit brickwall of depth 10 under
    depolarizing, amplitude-damping and readout noise (workloads.noisy_brickwall),
    one trajectory per shot, each in one workgroup with the state in LDS."""
    from aleph_amd import noise as N
    from aleph_amd import workloads as W
    from aleph_amd.backend import HipSvBackend
    c, m = W.noisy_brickwall(NOISE_N, NOISE_DEPTH, seed=args.seed)
    hip = HipSvBackend(seed=1)
    N.outcomes(c, m, 1024, 1, backend=hip)                      # warm-up
    times = []
    for rep in range(3):
        t0 = time.perf_counter()
        out = N.outcomes(c, m, args.noise_shots, 7 + rep, backend=hip)
        times.append(time.perf_counter() - t0)
    s = min(times)
    _, att = N.resolve(c, m)
    return {"workload": f"noisy brickwall n={NOISE_N} depth={NOISE_DEPTH}: depolarizing 1e-3 (U3) / 1e-2 (CZ), "
                        f"amplitude damping 2e-3 (U3), readout 2-3 %", "shots": args.noise_shots, "s": s,
            "trajectories_per_s": args.noise_shots / s, "gates": len(c.gates()), "channels_per_shot": len(att),
            "distinct_outcomes": int(np.unique(out).size),
            "note": "host-timed call: program upload, all trajectories, outcome copy back"}


def exchange_leg(args) -> dict:
    """Rank-slice exchange of a distributed n=30 FP64 state held as 8 ranks of
    2^27 amplitudes on this device: k = 1, 2, 3 global bits swapped with the
    top local bits in one in-place kernel.  Moved amplitudes are read once
    and written once: 2 * (1 - 2^-k) * 2^30 * 16 bytes per exchange."""
    from aleph_amd import dist as D
    from aleph_amd import workloads as W
    from aleph_amd.backend import HipSvBackend
    n, g = 30, 3
    layout = D.DistLayout(n, g)
    hip = HipSvBackend(seed=1)
    dx = D.DistSvBackend(hip)
    ranks = dx.alloc_ranks(layout)
    rows = []
    for bits in ([n - 1], [n - 1, n - 2], [n - 1, n - 2, n - 3]):
        k = len(bits)
        dx.exchange(ranks, layout, bits)
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            dx.exchange(ranks, layout, bits)
            best = min(best, time.perf_counter() - t0)
        moved = 2.0 * (1.0 - 2.0 ** -k) * float(1 << n) * 16.0
        rows.append({"k": k, "ms": best * 1e3, "gbs": moved / best / 1e9,
                     "frac_of_hbm_peak": moved / best / 1e9 / HBM_PEAK_GBS})
    for st in ranks:
        st.free()
    p = D.plan(W.brickwall(n, args.depth, seed=args.seed), layout)
    return {"layout": "n=30 FP64, 8 ranks x 2^27 amplitudes on one device", "exchanges": rows,
            "brickwall_plan": {"exchanges": p.stats.exchanges, "amps_moved_per_rank": p.stats.amps_moved_per_rank,
                               "local_swaps": p.stats.local_swaps},
            "note": "host-timed synchronous call (kernel + pointer upload); bytes = moved amplitudes read + written"}


def batch_probe(args) -> None:
    """Each size class of the QAOA batch workload alone, per executor and
    with or without solo instances: one JSON line per run."""
    from aleph_amd import _lib as L
    from aleph_amd import workloads as W
    from aleph_amd.backend import HipSvBackend
    from aleph_amd.batch import HipBatch
    hip = HipSvBackend(seed=1, precision=32)
    graphs, ga, be = W.qaoa_batch(args.batch_count, seed=args.seed)
    ns = np.array([g[0] for g in graphs])
    for n in sorted(set(ns.tolist())):
        idx = np.flatnonzero(ns == n)
        sub = [graphs[i] for i in idx]
        for mode in (L.BATCH_PERSISTENT, L.BATCH_ROUNDS):
            for solo in (64, 0):
                b = HipBatch.qaoa(hip, sub, ga[idx], be[idx], precision=32, mode=mode, solo_qubits=solo)
                b.run()
                ms = min(b.run_timed() for _ in range(3))
                st = b.stats()
                b.free()
                print(json.dumps({"n": n, "instances": len(idx), "mode": ["persistent", "rounds"][mode],
                                  "solo_instances": st["solo_instances"], "ms": round(ms, 3),
                                  "tbs": round(st["sweep_bytes"] / ms / 1e9, 3), "passes": st["n_passes"]}),
                      flush=True)


def main():
    args = parse()
    if args.batch_probe:
        batch_probe(args)
        return
    if args.traffic_child:
        traffic_child(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; --gpus must be 1 (sharding is a design note)"}))
        sys.exit(2)
    from aleph_amd import workloads as W
    from aleph_amd.backend import HipSvBackend

    n, depth = args.n, args.depth
    circuit = W.brickwall(n, depth, seed=args.seed)
    n_gates = len(circuit.gates())
    log(f"bench: brickwall n={n} depth={depth} gates={n_gates} fp64")
    b64 = HipSvBackend(seed=1)
    plan_kw = dict(local_qubits=args.local_qubits, max_ops=args.max_ops, jit=not args.no_jit)
    ms, stats, launch_ms, kinds, norm, st = time_circuit(b64, circuit, args.steps, args.warmup, **plan_kw)
    log(f"bench: fp64 {ms:.2f} ms/circuit passes={stats['n_passes']} norm-1={norm - 1:.2e}")
    fp32_vs_fp64 = None
    if not args.no_extra:
        # The FP32 run of the same plan beside the FP64 state, compared on the
        # device (overlap and max |difference| accumulated in FP64).
        s32 = b64.allocate(n, precision=32)
        p32 = b64.compile(circuit, precision=32, **plan_kw)
        p32.execute(s32)
        cmp = b64.compare(st, s32)
        p32.free()
        fp32_vs_fp64 = {"one_minus_fidelity": 1.0 - cmp["fidelity"] / (norm * b64.norm_sqr(s32)),
                        "max_abs": cmp["max_abs"]}
        s32.free()
    st.free()

    fused = launch_ms[kinds == 0]
    sweep_bytes = 2.0 * (1 << n) * 16.0
    # The practical in-place streaming ceiling of a fused pass on this device:
    # one pass whose only op is a diagonal on qubit 0 (the pass skeleton,
    # every amplitude read and written once, no arithmetic to speak of).
    from aleph_amd.circuit import Circuit
    _, _, s_launch, s_kinds, _, s_st = time_circuit(b64, Circuit(n).rz(0.25, 0), 1, 1, **plan_kw)
    s_st.free()
    stream_ms = float(np.min(s_launch[s_kinds == 0])) if np.any(s_kinds == 0) else float("nan")
    avg_ms = float(fused.mean()) if fused.size else float("nan")
    achieved = sweep_bytes / (avg_ms * 1e-3) / 1e9 if fused.size else 0.0
    total_bytes = stats["sweep_bytes"]
    flops = fused_pass_flops(circuit, 64)
    out = {
        "metric": "n=30 FP64 circuit s + HBM GB/s %peak; FP32+learned vs FP64 rule speedup",
        "value": ms / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random brickwall, complex128 amplitudes from |0...0>)",
        "config": {"workload": f"random brickwall U3+CZ n={n} depth={depth} seed={args.seed}, FP64, rule planner",
                   "qubits": n, "depth": depth, "gates": n_gates, "passes": stats["n_passes"],
                   "planner": "rule",
                   "codegen": "compiled op loop" if args.no_jit else "per-pass generated kernels (hiprtc)",
                   "plan_compile_s": stats["compile_s"]},
        "gates_per_s": n_gates / (ms / 1e3),
        "hbm_gbs_circuit": total_bytes / (ms / 1e3) / 1e9,
        "hbm_frac_circuit": total_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused_pass<double>" if args.no_jit else "aleph_pass_f64 (generated per pass)",
                     "stream_ceiling": {"pass_ms": stream_ms, "gbs": sweep_bytes / (stream_ms * 1e-3) / 1e9,
                                        "frac": sweep_bytes / (stream_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                        "what": "one fused pass with a single diagonal op at the same n: "
                                                "the in-place read+write streaming rate of the pass skeleton"},
                     "frac_of_stream_ceiling": (stream_ms / avg_ms) if fused.size else None,
                     "avg_launch_ms": avg_ms, "launches": int(fused.size),
                     "algorithmic_bytes_per_launch": sweep_bytes,
                     "fp64": {"algorithmic_flops_per_launch": flops,
                              "achieved_tflops": flops / (avg_ms * 1e-3) / 1e12 if fused.size else 0.0,
                              "peak_tflops": FP64_PEAK_TFLOPS,
                              "frac": flops / (avg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS if fused.size else 0.0,
                              "hbm_floor_ms": sweep_bytes / (HBM_PEAK_GBS * 1e9) * 1e3,
                              "fp64_floor_ms": flops / (FP64_PEAK_TFLOPS * 1e12) * 1e3}},
        "norm_error": norm - 1.0,
    }
    if not args.no_extra:
        import aleph_amd._lib as L
        from aleph_amd.cost_model import DEFAULT_PATH, CostModel
        b32 = HipSvBackend(seed=1, precision=32)
        ms32, st32, lm32, k32, norm32, s = time_circuit(b32, circuit, args.steps, args.warmup, **plan_kw)
        s.free()
        log(f"bench: fp32 {ms32:.2f} ms/circuit passes={st32['n_passes']}")
        out["fp32"] = {"ms_per_step": ms32, "passes": st32["n_passes"], "speedup_vs_fp64_rule": ms / ms32,
                       "norm_error": norm32 - 1.0, "vs_fp64_state": fp32_vs_fp64}
        msr, _, _, _, normr, sr = time_circuit(b32, circuit, args.steps, args.warmup, renormalize=True, **plan_kw)
        sr.free()
        out["fp32"]["renormalized"] = {"ms_per_step": msr, "norm_error": normr - 1.0,
                                       "what": "the same plan with the FP64-normed renormalisation after it"}
        if DEFAULT_PATH.exists():
            for prec, bk in ((32, b32), (64, b64)):
                model = CostModel.load(bk)
                model.install()
                msl, stl, _, _, norml, s = time_circuit(bk, circuit, args.steps, args.warmup,
                                                        planner=L.PLANNER_LEARNED, jit=not args.no_jit)
                s.free()
                log(f"bench: fp{prec} learned {msl:.2f} ms/circuit passes={stl['n_passes']}")
                out[f"fp{prec}_learned"] = {
                    "ms_per_step": msl, "passes": stl["n_passes"], "speedup_vs_fp64_rule": ms / msl,
                    "chosen": {"local_qubits": stl["chosen_local_qubits"], "max_ops": stl["chosen_max_ops"],
                               "fuse_2q": not stl["chosen_flags"] & L.PLAN_NO_FUSE_2Q,
                               "nontemporal": bool(stl["chosen_flags"] & L.PLAN_NONTEMPORAL),
                               "generated_kernels": bool(stl["chosen_flags"] & L.PLAN_JIT),
                               "kq_width": stl["chosen_kq_width"]},
                    "kept_rule": bool(stl["chosen_rule"]),
                    "rule_predicted_ms": stl["rule_predicted_us"] / 1e3,
                    "predicted_ms": stl["predicted_us"] / 1e3, "plan_ms": stl["plan_ms"], "norm_error": norml - 1.0}
                model.free()
        del b32
        q = W.qft(n)
        msq, stq, lmq, kq, normq, sq = time_circuit(b64, q, args.steps, args.warmup, **plan_kw)
        sq.free()
        log(f"bench: qft fp64 {msq:.2f} ms/circuit passes={stq['n_passes']}")
        out["qft_fp64"] = {"ms_per_step": msq, "passes": stq["n_passes"], "gates": len(q.gates()),
                           "hbm_gbs": stq["sweep_bytes"] / (msq / 1e3) / 1e9}
    if not args.no_batch:
        log("bench: batched qaoa")
        out["qaoa_batch"] = batch_leg(args)
        log(f"bench: qaoa batch {out['qaoa_batch']['ms']:.1f} ms")
    if not args.no_noise:
        log("bench: noise trajectories")
        out["noise"] = noise_leg(args)
        log(f"bench: noise {out['noise']['trajectories_per_s']:.0f} trajectories/s")
    if not args.no_dist:
        log("bench: rank-slice exchange")
        out["dist_exchange"] = exchange_leg(args)
    if not args.no_pmc:
        log("bench: rocprofv3 traffic passes")
        tr = measure_traffic(args, float(1 << n) * 16.0)
        if tr is not None:
            out["roofline"]["traffic"] = tr["bytes"]
            out["roofline"]["traffic_detail"] = tr
    if not args.no_cpu:
        log("bench: cpu baseline")
        cpu = cpu_baseline(circuit)
        out["cpu_baseline"] = cpu
        out["speedup_vs_cpu"] = cpu["value"] / (ms / 1e3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
