"""Train the learned sweep cost model on this GPU and write the model file.

    python -m q1tsim_amd.train_costmodel [--n 28] [--out PATH]

The planner's seeded exploration (plan mode 2) runs each training circuit on
a complex64 state of `n` qubits, picking per flush one of the two tilings
and per sweep one of its candidate programs (register phases, 3-qubit
blocks, 4-qubit blocks, blocks of the configured width), and the state
records every sweep's features and HIP-event time.  Circuits come from three
families (random pairs at several two-qubit fractions and depths, layered,
QFT), run on the interpreter kernels and on kernels specialised to each
program (feature 19 tells them apart); every fifth circuit of each family is
held out.  The fit is the HIP Adam kernels' minibatch training; the file
holds the parameters, the feature normalisation and a metadata record with
the sample counts, the validation errors per kernel kind, and the wall time
of greedy fusion, per-gate sweeps and the learned plan on held-out circuits.
"""
import argparse
import json
import time

import numpy as np

from . import costmodel, workloads
from .costmodel import CostModel


def families(n, jit):
    """Training circuits as (family, ops) pairs; fewer for the specialised
    kernels, whose every new program is compiled first."""
    out = []
    npairs, nlay, nqft = (15, 10, 10) if jit else (100, 50, 50)
    for k in range(npairs):
        frac = (0.1, 0.25, 0.4)[k % 3]
        depth = (20, 40, 60)[(k // 3) % 3]
        out.append(("pairs", workloads.random_pairs(n, depth, 1000 + k, two_qubit_frac=frac)))
    for k in range(nlay):
        out.append(("layered", workloads.random_layered(n, (10, 20)[k % 2], 2000 + k)))
    for k in range(nqft):
        out.append(("qft", workloads.qft(n)))   # a new random choice of programs on every flush
    return out


def split(circs):
    """Every fifth circuit of each family is held out."""
    train, val, seen = [], [], {}
    for fam, ops in circs:
        i = seen.get(fam, 0)
        seen[fam] = i + 1
        (val if i % 5 == 4 else train).append(ops)
    return train, val


def record_all(n, jit, seed):
    train, val = split(families(n, jit))
    t0 = time.perf_counter()
    tr = costmodel.record(n, train, fusion_k=4, seed=seed, jit=jit)
    va = costmodel.record(n, val, fusion_k=4, seed=seed + 1, jit=jit)
    return tr, va, time.perf_counter() - t0


def flush_time(ops, n, cm, plan, jit, reps=3):
    """Median wall time of the circuit's flush (planning included) on a
    complex64 state: plan 'per_gate' (no blocks), 'greedy' (4-qubit blocks)
    or 'learned'; a first flush compiles the specialised kernels.  Returns
    (seconds, sweeps per flush, blocks per flush)."""
    from .gvs import GpuVectorState, gate_matrix
    s = GpuVectorState(n, 1, "c64")
    s.set_plan(0 if plan == "per_gate" else 4, True)
    s.set_cost_model(cm, 1 if plan == "learned" else 0, 1)
    if jit:
        s.set_jit(True)
    ts = []
    for r in range(reps + 1):
        for g, b, p in ops:
            s.apply_gate(gate_matrix(g, p), b, g)
        t0 = time.perf_counter()
        s.flush()
        if r:
            ts.append(time.perf_counter() - t0)
    st = s.stats()
    s.close()
    return float(np.median(ts)), st["sweeps"] // (reps + 1), st["blocks"] // (reps + 1)


def train(n=28, seed=5, epochs=40, evaluate=True, log=print):
    """Record, fit and validate; returns (model, metadata)."""
    rec = {}
    for jit in (False, True):
        tr, va, secs = record_all(n, jit, seed + (10 if jit else 0))
        rec[jit] = (tr, va)
        log("recorded %s: %d training sweeps, %d held-out sweeps in %.1f s"
            % ("specialised" if jit else "interpreter", len(tr[0]), len(va[0]), secs))
    feat = np.concatenate([rec[j][0][0] for j in (False, True)])
    tiles = np.concatenate([rec[j][0][1] for j in (False, True)])
    ms = np.concatenate([rec[j][0][2] for j in (False, True)])
    cm = CostModel(seed)
    t0 = time.perf_counter()
    mse = costmodel.fit(cm, feat, tiles, ms, epochs=epochs, batch=256)
    fit_s = time.perf_counter() - t0
    meta = {"n_qubits": n, "seed": seed, "epochs": epochs, "fit_s": round(fit_s, 2), "train_mse_log_time": mse,
            "features": costmodel.NFEAT, "params": costmodel.NPARAMS,
            "training_sweeps": int(len(feat)),
            "training_blocks": int(feat[:, 1].sum() + feat[:, 2].sum()),
            "held_out": {}}
    for jit in (False, True):
        kind = "specialised" if jit else "interpreter"
        tr, va = rec[jit]
        meta["held_out"][kind] = costmodel.errors(cm, *va)
        meta["held_out"][kind]["blocks"] = int(va[0][:, 1].sum() + va[0][:, 2].sum())
        meta["train_" + kind] = costmodel.errors(cm, *tr)
        log("%s: held-out %s" % (kind, json.dumps(meta["held_out"][kind])))
    if evaluate:
        # wall time of the plans on held-out circuits (seeds outside the
        # training families), kernels specialised to each program
        ev = {}
        for name, ops in (("pairs_d40", workloads.random_pairs(n, 40, 777)),
                          ("layered_d20", workloads.random_layered(n, 20, 778))):
            ev[name] = {}
            for plan in ("greedy", "per_gate", "learned"):
                secs, sweeps, blocks = flush_time(ops, n, cm, plan, True)
                ev[name][plan] = {"seconds": round(secs, 5), "sweeps": sweeps, "blocks": blocks}
            log("%s: %s" % (name, json.dumps(ev[name])))
        meta["held_out_circuits_specialised"] = ev
    return cm, meta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=28)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--out", default=costmodel.DEFAULT_MODEL)
    ap.add_argument("--no-eval", action="store_true")
    args = ap.parse_args()
    import torch
    t0 = time.perf_counter()
    cm, meta = train(args.n, args.seed, args.epochs, evaluate=not args.no_eval,
                     log=lambda s: print(s, flush=True))
    meta["device"] = torch.cuda.get_device_name(0)
    meta["total_s"] = round(time.perf_counter() - t0, 1)
    costmodel.save(cm, args.out, meta)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
