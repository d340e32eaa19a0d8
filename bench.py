"""Benchmark: 30-qubit random circuit on one MI355X.

Workload (the headline configuration): the 30-qubit depth-20 layered random
circuit (600 one-qubit gates from {H,X,S,T,RX,RY,RZ} + 290 nearest-neighbour
CX), statevector resident in HBM.  A step is one application of the whole
circuit to the resident state.

  * plain  : complex128 state, per-gate FP64 arithmetic, gates grouped into
             LDS/register tile sweeps, each sweep run by the kernel
             specialised to its program at run time (the reported `value`,
             in HBM GB/s; compiled during the warm-up steps); the same
             circuit on the interpreter kernels is reported beside it
  * fused  : complex64 state, fusion width 4: per sweep the learned cost
             model (trained in this run on sweeps the simulator records
             itself) picks register phases or 3-/4-qubit FP32-MFMA blocks;
             its time, speedup over plain and infidelity / max-abs error
             against the plain complex128 state are reported, beside greedy
             fixed-width fusion for k = 3..6 and the per-gate complex64
             path; `roofline_fused` puts greedy 4-qubit fusion's kernel
             against the FP32 MFMA peak.

Also reported: the 30-qubit QFT on a random normal state, FP64 sampling of
2^16 runs (measure_all, peek semantics), a batch of 4096 circuits with
mid-circuit measurements (circuits/s), and the learned planner against greedy
fusion on a 32-qubit random-pair circuit.  `roofline` is the dominant kernel's
algorithmic bytes per launch over its average HIP-event duration on the
state's stream; `cpu_baseline` times the CPU restatement of the reference
(the oracle, one core) on a bounded sample.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "30-qubit circuit time: fused c64 vs plain c128 at infidelity<1e-6; HBM GB/s"
PEAK_GBPS = 8000.0


def apply_circuit(state, ops):
    from q1tsim_amd.gvs import gate_matrix
    for name, bits, params in ops:
        state.apply_gate(gate_matrix(name, params), bits, name)
    state.flush()


def timed_steps(state, ops, steps, warmup, profile=True):
    """Median-able step times of the whole circuit on the resident state.
    The warm-up steps also compile the specialised sweep kernels when the
    state uses them (the first flush of a program structure compiles it)."""
    for _ in range(warmup):
        apply_circuit(state, ops)
    state.set_profiling(profile)
    times, sweeps, nbytes = [], 0, 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        apply_circuit(state, ops)
        times.append(time.perf_counter() - t0)
        st = state.stats()
        sweeps, nbytes = st["sweeps"], st["bytes"]
    prof = state.kernel_profile()
    state.set_profiling(False)
    return times, sweeps, nbytes, prof


def roofline_of(prof):
    name = max(prof, key=lambda k: prof[k][1])
    launches, ms, nbytes = prof[name]
    avg_ms = ms / launches
    achieved = nbytes / launches / (avg_ms * 1e-3) / 1e9
    return name, {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                  "frac": round(achieved / PEAK_GBPS, 4), "traffic": None,
                  "kernel": name, "avg_ms": round(avg_ms, 4), "bytes_per_launch": nbytes / launches,
                  "launches": launches}


def profile_values(kernel, n, what):
    """The `key: value` lines of the committed rocprofv3 summary
    profiles/<kernel>_layered<n>_<what>.txt when it covers this kernel and
    size; else None."""
    path = os.path.join(ROOT, "profiles", "%s_layered%d_%s.txt" % (kernel, n, what))
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith(" "):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    if vals.get("kernel") != kernel or vals.get("n_qubits") != str(n):
        return None
    return vals


def profiled_traffic(kernel, n):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE / WRITE_SIZE, scaled by a calibration kernel), or None."""
    vals = profile_values(kernel, n, "traffic")
    return None if vals is None else float(vals["traffic_bytes_per_launch"])


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(ops_full, n_full, n=22, layers=2, reps=3):
    """The oracle (C restatement of VectorState, 1 core, the reference's data
    movement: per-block temporaries, table-based permutations) timed as the
    CPU-baseline protocol asks: the gate stream and the sampling separately,
    each the median of `reps` runs after one warm-up.  Gate stream: the first
    `layers` layers of the same layered recipe at n qubits; per-gate times
    scale with 2^n, so the 30-qubit circuit time is that median x (gates of
    the full circuit / gates timed) x 2^(30-n), stated in `sample`.
    Sampling: peek_all of 2^16 runs at n qubits (the reference's
    measure_all collapse at 30 qubits would need ~1 PB)."""
    from oracle import oracle as orc
    from q1tsim_amd import workloads
    ops = [op for op in workloads.random_layered(n, layers, 1)]
    gate_t = []
    for rep in range(reps + 1):
        st = orc.RefState(n, 1)
        t0 = time.perf_counter()
        for name, bits, params in ops:
            st.apply_gate(name, bits, params)
        if rep:
            gate_t.append(time.perf_counter() - t0)
    samp_t = []
    res = np.zeros(1 << 16, np.uint64)
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        st.measure_all_into(list(range(n)), res, orc.Rng(rep + 1), collapse=False)
        if rep:
            samp_t.append(time.perf_counter() - t0)
    g_s = float(np.median(gate_t))
    s_s = float(np.median(samp_t))
    scale_n = float(1 << (n_full - n))
    est_circuit_s = g_s * len(ops_full) / len(ops) * scale_n
    S = float(16 << n_full)
    unf = workloads.unfused_bytes(ops_full, S)
    model, nproc = host_cpu()
    return {"value": round(unf / est_circuit_s / 1e9, 3), "unit": "GB/s (unfused-equivalent)", "cores": 1,
            "kind": "port",
            "sample": "oracle (C restatement of q1tsim VectorState, FP64, 1 thread): gate stream = first %d "
                      "layers (%d gates) of the layered recipe at n=%d, median of %d runs after 1 warm-up = %.3f s; "
                      "extrapolated x%.2f (890/%d gates) x%d (2^30/2^%d) -> %.1f s per 30-qubit circuit. "
                      "Sampling: peek_all of 2^16 runs at n=%d, median %.4f s"
                      % (layers, len(ops), n, reps, g_s, len(ops_full) / len(ops), len(ops), int(scale_n), n,
                         est_circuit_s, n, s_s),
            "gate_stream_s": round(g_s, 4), "gate_stream_runs_s": [round(t, 4) for t in gate_t],
            "sampling_65536_s": round(s_s, 5), "circuit_seconds_est": round(est_circuit_s, 2),
            "host_cpu": model, "host_nproc": nproc}


def batch_extra(count, seed=2024):
    """Config: `count` independent circuits, qubits uniform in [10,22], depth
    in [10,40], 1-3 mid-circuit Z measurements each followed by a classically
    conditioned X or RZ, measure_all at the end, nr_runs = 64 branch columns;
    all executed by one Batch.  Reports circuits/s per precision."""
    from q1tsim_amd import workloads
    from q1tsim_amd.q1tsim import Batch, Circuit
    t0 = time.perf_counter()
    wl = workloads.batch_workload(count, seed)
    gen_s = time.perf_counter() - t0
    out = {"circuits": count, "nr_runs": 64, "qubits": [10, 22], "depth": [10, 40], "measurements": [1, 3],
           "program_ops": int(sum(len(p) for _, _, p, _ in wl)), "generate_s": round(gen_s, 2)}
    from q1tsim_amd.gvs import set_device_cache_limit
    for prec in ("c64", "c128"):
        # each precision starts from an empty device block cache, as it would
        # in a process of its own: the blocks the earlier legs released (the
        # 30- and 32-qubit states) are no use to the batch's small states
        set_device_cache_limit(1)
        set_device_cache_limit(0)
        circs = []
        for n, nc, prog, cs in wl:
            c = Circuit(n, nc)
            workloads.build_program(c, prog)
            c.seed(cs)
            circs.append(c)
        b = Batch(prec, 0, 0)
        for c in circs:
            b.add(c)
        t0 = time.perf_counter()
        tried = b.autotune(64, sample=min(128, count))
        tune_s = time.perf_counter() - t0
        # one untimed pass (device block cache filled, as in a serving loop),
        # then every circuit's random stream restarted from its seed
        b.execute(64)
        for c, (_, _, _, cs) in zip(circs, wl):
            c.seed(cs)
        t0 = time.perf_counter()
        b.execute(64)
        dt = time.perf_counter() - t0
        st = b.stats()
        out[prec] = {"seconds": round(dt, 3), "circuits_per_s": round(count / dt, 1), "warmup_runs": 1,
                     "waves": st["waves"],
                     "rounds": st["rounds"], "sync_ops": st["sync_ops"], "gate_launches": st["launches"],
                     "column_launches": st["column_launches"], "solo_ops": st["solo_ops"],
                     "plan_s": round(st["plan_s"], 3), "pack_s": round(st["pack_s"], 3),
                     "gates": st["gates"], "sweeps": st["sweeps"], "sweep_hbm_bytes": st["bytes"],
                     "split_s": {k: round(st[k], 3) for k in ("start_s", "queue_s", "flush_s", "measure_s", "other_s")},
                     "autotune": {"seconds": round(tune_s, 3), "chosen_M_R_L_fusion": list(b.config()),
                                  "tried": [[t[0], t[1], t[2], t[3], round(t[4], 4)] for t in tried]}}
        del b, circs
    return out


def cost_model(retrain):
    """The learned planner's model: the shipped MI355X model
    (q1tsim_amd/data/costmodel_gfx950.npz, written by
    q1tsim_amd/train_costmodel.py from sweeps the simulator recorded on this
    part: seeded exploration of the candidate programs on 28-qubit complex64
    states, interpreter and specialised kernels, held-out circuits for
    validation), or with `retrain` one trained in this run the same way."""
    from q1tsim_amd import costmodel, train_costmodel
    if retrain:
        t0 = time.perf_counter()
        cm, meta = train_costmodel.train(evaluate=False, log=lambda s: print(s, file=sys.stderr, flush=True))
        meta["source"] = "trained in this run"
        meta["seconds"] = round(time.perf_counter() - t0, 1)
    else:
        cm = costmodel.load()
        meta = dict(cm.meta)
        meta["source"] = "q1tsim_amd/data/costmodel_gfx950.npz"
    keep = ("source", "n_qubits", "training_sweeps", "training_blocks", "train_mse_log_time", "held_out",
            "held_out_circuits_specialised", "device", "seconds")
    return cm, {k: meta[k] for k in keep if k in meta}


def learned_extra(cm, info, n=32, depth=40, seed=5):
    """Config: n-qubit depth-40 circuit with 25% CX on random pairs, complex64:
    the learned plan against greedy fixed-width fusion and per-gate sweeps on
    a held-out circuit, each compared with the complex128 run."""
    from q1tsim_amd import workloads
    from q1tsim_amd.gvs import GpuVectorState, compare
    ops = workloads.random_pairs(n, depth, seed)
    ref = GpuVectorState(n, 1, "c128")
    ref.set_jit(True)
    apply_circuit(ref, ops)   # warm-up (compiles the specialised kernels)
    ref.reset_all()
    t0 = time.perf_counter()
    apply_circuit(ref, ops)
    c128_s = time.perf_counter() - t0
    res = {"n_qubits": n, "depth": depth, "gates": len(ops), "c128_s": round(c128_s, 3)}
    res["cost_model"] = info["source"]
    for fk, mode, name in ((4, 0, "greedy"), (0, 0, "per_gate"), (4, 1, "learned")):
        s = GpuVectorState(n, 1, "c64")
        s.set_plan(fk, True)
        s.set_cost_model(cm, mode, seed)
        s.set_jit(True)
        apply_circuit(s, ops)   # warm-up (compiles the specialised kernels)
        s.reset_all()
        t0 = time.perf_counter()
        apply_circuit(s, ops)
        dt = time.perf_counter() - t0
        st = s.stats()
        infid, maxabs = compare(ref, s)
        res[name] = {"seconds": round(dt, 4), "sweeps": st["sweeps"], "blocks": st["blocks"],
                     "infidelity_vs_c128": infid, "max_abs_vs_c128": maxabs}
        del s
    res["learned_gain"] = round(res["greedy"]["seconds"] / res["learned"]["seconds"], 4)
    res["learned_vs_per_gate"] = round(res["per_gate"]["seconds"] / res["learned"]["seconds"], 4)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--fusion-k", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the QFT, sampling, batch and learned lines")
    ap.add_argument("--batch-circuits", type=int, default=4096)
    ap.add_argument("--train-cost-model", action="store_true",
                    help="train the learned planner's model in this run instead of loading the shipped one")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batch leg (both precisions), one JSON line; for kernel-trace profiles")
    ap.add_argument("--profile-workload", action="store_true",
                    help="for rocprofv3 counter runs: k_scale over the c128 state (a read+write of known size, "
                         "for scaling FETCH_SIZE/WRITE_SIZE), then the plain circuit twice (specialised kernels: "
                         "the first pass compiles them, the second is the profiled one)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU sharding is not part of this build (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from q1tsim_amd import workloads
    from q1tsim_amd.gvs import GpuVectorState, compare

    if args.batch_only:
        print(json.dumps({"batch": batch_extra(args.batch_circuits)}))
        return

    if args.profile_workload:
        n = args.n
        st = GpuVectorState(n, 1, "c128")
        st.set_plan(0, True)
        st.set_jit(True)
        st.set_random_normal(3)   # k_random_normal, k_norm_partial, k_scale (reads and writes 16 * 2^n bytes)
        lay = workloads.random_layered(n, args.depth, 1)
        apply_circuit(st, lay)    # compiles the specialised sweep kernels
        apply_circuit(st, lay)    # the profiled pass: the last `sweeps` q1t_jit_sweep_c128 dispatches
        s = st.stats()
        print(json.dumps({"profile_workload": True, "n": n, "state_bytes": 16 << n, "sweeps": s["sweeps"],
                          "sweep_bytes_per_launch": 2 * (16 << n)}))
        return

    n = args.n
    ops = workloads.random_layered(n, args.depth, 1)
    S128 = float(16 << n)
    from q1tsim_amd.gvs import jit_stats

    # plain complex128 (the reported `value`): per-gate FP64 arithmetic in
    # tile sweeps, each sweep run by the kernel specialised to its program
    # (compiled during the warm-up steps)
    jit0 = jit_stats()
    plain = GpuVectorState(n, 1, "c128")
    plain.set_plan(0, True)
    plain.set_jit(True)
    t_plain, sw_plain, bytes_plain, prof_plain = timed_steps(plain, ops, args.steps, max(args.warmup, 1))
    jit1 = jit_stats()
    ms_plain = 1e3 * float(np.median(t_plain))
    kname, roof = roofline_of(prof_plain)
    roof["traffic"] = profiled_traffic(kname, n)
    if roof["traffic"] is not None:
        roof["traffic_source"] = "profiles/%s_layered%d_traffic.txt (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)" % (
            kname, n)
    # the sweep also issues FP64 VALU work for its gates; its issue time per
    # launch (SQ_INSTS_VALU x 4 cycles / 1024 SIMDs / 2.4 GHz) next to the
    # HBM figure says which of the two bounds the launch
    valu = profile_values(kname, n, "valu")
    if valu is not None:
        issue = float(valu["valu_issue_ms_per_launch"])
        roof["valu_issue"] = {"ms_per_launch": round(issue, 4), "frac_of_launch": round(issue / roof["avg_ms"], 4),
                              "source": "profiles/%s_layered%d_valu.txt (rocprofv3 SQ_INSTS_VALU)" % (kname, n)}
    jit_info = {"kernels_compiled": jit1[0] - jit0[0], "compile_cpu_s": round(jit1[2] - jit0[2], 2),
                "note": "compiled in the warm-up steps; the timed steps reuse them"}
    # the same circuit on the interpreter kernels (no run-time compilation)
    interp = GpuVectorState(n, 1, "c128")
    interp.set_plan(0, True)
    t_int, _, _, prof_int = timed_steps(interp, ops, 2, 1)
    del interp

    # complex64: the per-gate path, greedy fixed-width fusion into k-qubit
    # FP32-MFMA blocks for k = 3..6, and the learned planner's choice per
    # sweep (register phases or 3-/4-qubit blocks), all on specialised
    # kernels; errors against the complex128 state
    cm, cm_info = cost_model(args.train_cost_model)
    c64 = {}
    legs = [("per_gate", 0, 0)] + [("greedy_k%d" % k, k, 0) for k in (3, 4, 5, 6)] + [("learned", args.fusion_k, 1)]
    for name, fk, mode in legs:
        st64 = GpuVectorState(n, 1, "c64")
        st64.set_plan(fk, True)
        st64.set_cost_model(cm, mode, 5)
        st64.set_jit(True)
        t, sw, nb, prof = timed_steps(st64, ops, args.steps, max(args.warmup, 1))
        infid, maxabs = compare(plain, st64)
        stt = st64.stats()
        c64[name] = {"ms": round(1e3 * float(np.median(t)), 3), "sweeps": sw, "hbm_bytes": nb, "fusion_k": fk,
                     "blocks": stt["blocks"], "mfma_flops": stt["mfma_flops"], "infidelity_vs_c128": infid,
                     "max_abs_vs_c128": maxabs,
                     "kernels": {k: {"launches": v[0], "ms": round(v[1], 3)} for k, v in prof.items()}}
        del st64
    ms_fused = c64["learned"]["ms"]
    # the fused kernel against its own roofline: greedy 4-qubit fusion (a
    # block in every sweep), MFMA flops per launch over the launch time, and
    # the same launches' HBM bytes
    g4 = c64["greedy_k4"]
    kf = max(g4["kernels"], key=lambda k: g4["kernels"][k]["ms"])
    lf, msf = g4["kernels"][kf]["launches"], g4["kernels"][kf]["ms"]
    # flops and bytes are one pass's (the state's stats cover its last
    # flush), the kernel profile all timed passes': per launch, one sweep
    # of the single column is one launch
    flops_launch = g4["mfma_flops"] / g4["sweeps"]
    bytes_launch = g4["hbm_bytes"] / g4["sweeps"]
    avg_s = msf / lf * 1e-3
    tflops = flops_launch / avg_s / 1e12
    roof_fused = {"bound": "mfma", "achieved": round(tflops, 2), "peak": 157.3, "unit": "TFLOP/s",
                  "frac": round(tflops / 157.3, 4), "traffic": None, "kernel": kf, "workload": "greedy 4-qubit fusion",
                  "avg_ms": round(msf / lf, 4), "mfma_flops_per_launch": flops_launch,
                  "hbm_GBps": round(bytes_launch / avg_s / 1e9, 1),
                  "hbm_frac": round(bytes_launch / avg_s / 1e9 / PEAK_GBPS, 4)}

    extra = {}
    if not args.no_extra:
        # QFT on a random normal state + FP64 sampling of 2^16 runs
        q = workloads.qft(n)
        plain.set_random_normal(7)
        apply_circuit(plain, q)   # warm-up: compiles the QFT's specialised kernels
        plain.set_random_normal(7)
        t0 = time.perf_counter()
        apply_circuit(plain, q)
        extra["qft30_ms"] = round(1e3 * (time.perf_counter() - t0), 2)
        extra["qft30_sweeps"] = plain.stats()["sweeps"]
        # the same QFT in complex64: per-gate sweeps, greedy 4-qubit fusion and
        # the learned plan, against the complex128 result
        qft64 = {}
        for name, fk, mode in (("per_gate", 0, 0), ("greedy_k4", 4, 0), ("learned", 4, 1)):
            s64 = GpuVectorState(n, 1, "c64")
            s64.set_plan(fk, True)
            s64.set_cost_model(cm, mode, 5)
            s64.set_jit(True)
            s64.set_random_normal(7)
            apply_circuit(s64, q)   # warm-up: compiles the kernels
            s64.set_random_normal(7)
            t0 = time.perf_counter()
            apply_circuit(s64, q)
            dt = time.perf_counter() - t0
            st = s64.stats()
            infid, maxabs = compare(plain, s64)
            qft64[name] = {"ms": round(1e3 * dt, 2), "sweeps": st["sweeps"], "blocks": st["blocks"],
                           "infidelity_vs_c128": infid, "max_abs_vs_c128": maxabs}
            del s64
        extra["qft30_c64"] = qft64
        res = np.zeros(1 << 16, np.uint64)
        from q1tsim_amd.gvs import Xoshiro
        t0 = time.perf_counter()
        plain.measure_all_into(list(range(n)), res, Xoshiro(1), collapse=False)
        extra["sample65536_ms"] = round(1e3 * (time.perf_counter() - t0), 2)
        del plain
        # the part's read+write streaming ceiling: a plain device copy of 8 GiB
        from q1tsim_amd.gvs import copy_bandwidth
        cms, cgbps = copy_bandwidth(8 << 30, 5)
        extra["device_copy"] = {"bytes": 8 << 30, "ms": round(cms, 3), "GBps_read_plus_write": round(cgbps, 1),
                                "frac_of_peak": round(cgbps / PEAK_GBPS, 4)}
        try:
            extra["batch"] = batch_extra(args.batch_circuits)
        except Exception as e:   # reported, not hidden: the headline line still prints
            extra["batch"] = {"error": repr(e)}
        try:
            extra["learned_planner"] = learned_extra(cm, cm_info)
        except Exception as e:
            extra["learned_planner"] = {"error": repr(e)}

    value = bytes_plain / (ms_plain * 1e-3) / 1e9
    unf = workloads.unfused_bytes(ops, S128)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_plain, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "30-qubit depth-20 layered random circuit (600 1q + 290 CX), statevector in HBM; "
                               "value = plain complex128 path",
                   "n_qubits": n, "depth": args.depth, "gates": len(ops), "seed": 1},
        "plain_c128": {"ms": round(ms_plain, 3), "sweeps": sw_plain, "hbm_bytes": bytes_plain,
                       "GBps": round(value, 1), "frac_of_peak": round(value / PEAK_GBPS, 4),
                       "unfused_equivalent_GBps": round(unf / (ms_plain * 1e-3) / 1e9, 1),
                       "kernels": "specialised per sweep program (hipRTC)", "jit": jit_info,
                       "interpreter_ms": round(1e3 * float(np.median(t_int)), 3),
                       "interpreter_kernels": {k: {"launches": v[0], "ms": round(v[1], 3)}
                                               for k, v in prof_int.items()}},
        "fused_c64": {"ms": ms_fused, "planner": "learned (cost model: %s)" % cm_info["source"],
                      "sweeps": c64["learned"]["sweeps"], "blocks": c64["learned"]["blocks"],
                      "GBps": round(c64["learned"]["hbm_bytes"] / (ms_fused * 1e-3) / 1e9, 1),
                      "fusion_k": args.fusion_k, "speedup_vs_plain": round(ms_plain / ms_fused, 3),
                      "speedup_vs_plain_interpreter": round(1e3 * float(np.median(t_int)) / ms_fused, 3),
                      "infidelity_vs_c128": c64["learned"]["infidelity_vs_c128"],
                      "max_abs_vs_c128": c64["learned"]["max_abs_vs_c128"], "dtype": "f32",
                      "per_k": {k: {kk: v[kk] for kk in ("ms", "sweeps", "blocks", "mfma_flops", "infidelity_vs_c128",
                                                         "max_abs_vs_c128")} for k, v in c64.items()},
                      "cost_model": cm_info},
        "roofline": roof,
        "roofline_fused": roof_fused,
        "kernels_plain": {k: {"launches": v[0], "ms": round(v[1], 3)} for k, v in prof_plain.items()},
        "kernels_fused": c64["learned"]["kernels"],
    }
    out.update(extra)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ops, n)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
