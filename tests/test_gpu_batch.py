"""Batched execution on the GPU: circuits of uneven width with mid-circuit
measurements and classically conditioned gates, run through one Batch, give
the classical state that each gives when executed alone with the same seed
(circuit.rs:562-600 per circuit) and that the oracle's RefCircuit gives; the
end states of the batch agree with the individually run circuits."""
import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd import workloads
from q1tsim_amd.q1tsim import Batch, Circuit

pytestmark = pytest.mark.gpu

SHOTS = 64


def make(n, nc, prog, seed, prec, fusion=0):
    c = Circuit(n, nc)
    workloads.build_program(c, prog)
    c.set_backend(prec, fusion)
    c.seed(seed)
    return c


def make_ref(n, nc, prog):
    r = orc.RefCircuit(n, nc)
    for op in prog:
        if op[0] == "gate":
            r.add_gate(op[1], op[2], op[3])
        elif op[0] == "cgate":
            r.add_conditional_gate(op[1], op[2], op[3], op[4], op[5])
        elif op[0] == "measure":
            r.measure(op[1], op[2], op[3])
        elif op[0] == "measure_all":
            r.measure_all(op[1], op[2])
        elif op[0] == "peek":
            r.peek(op[1], op[2], op[3])
        elif op[0] == "peek_all":
            r.peek_all(op[1], op[2])
        elif op[0] == "reset":
            r.reset(op[1])
        elif op[0] == "reset_all":
            r.reset_all()
    return r


def run_batch(wl, prec, fusion=0, budget=0, keep=False):
    batch = Batch(prec, fusion, budget)
    circs = [make(n, nc, prog, s, prec, fusion) for (n, nc, prog, s) in wl]
    for c in circs:
        batch.add(c)
    batch.execute(SHOTS, keep_states=keep)
    return batch, circs


# widths 3..16 cover both batch paths: whole-state-in-LDS (n <= 13) and
# tiled sweeps over all states' tiles in one launch (n >= 14)
WL = workloads.batch_workload(24, seed=11, qubits=(3, 16), depths=(3, 12), meas=(1, 3))


def test_batch_cstate_c128_matches_solo_and_oracle():
    batch, circs = run_batch(WL, "c128")
    st = batch.stats()
    assert st["waves"] == 1 and st["sync_ops"] > 0 and st["gates"] > 0
    for (n, nc, prog, s), c in zip(WL, circs):
        solo = make(n, nc, prog, s, "c128")
        solo.execute(SHOTS)
        got = np.array(c.cstate(), dtype=np.uint64)
        assert np.array_equal(got, np.array(solo.cstate(), dtype=np.uint64)), (n, s)
        if n <= 12:
            r = make_ref(n, nc, prog)
            r.execute(SHOTS, s)
            assert np.array_equal(got, r.cstate()), (n, s)


def test_batch_end_states_c128():
    wl = [(n, nc, prog[:-1], s) for (n, nc, prog, s) in WL]   # stop before measure_all
    _, circs = run_batch(wl, "c128", keep=True)
    for (n, nc, prog, s), c in zip(wl, circs):
        solo = make(n, nc, prog, s, "c128")
        solo.execute(SHOTS)
        assert c.counts() == solo.counts(), (n, s)
        for col in range(c.nr_columns()):
            a, b = c.amplitudes(col), solo.amplitudes(col)
            assert np.max(np.abs(a - b)) <= 1e-12, (n, s, col)


# fusion 0, 4 and 6: the batched sweep kernels without MFMA code, with the
# 3-4-qubit blocks and with the 5-6-qubit blocks
@pytest.mark.parametrize("fusion", [0, 4, 6])
def test_batch_c64(fusion):
    wl = [(n, nc, prog[:-1], s) for (n, nc, prog, s) in WL]
    _, circs = run_batch(wl, "c64", fusion=fusion, keep=True)
    for (n, nc, prog, s), c in zip(wl, circs):
        ref = make(n, nc, prog, s, "c128")
        ref.execute(SHOTS)
        # branch outcomes come from FP64 sums of FP32 amplitudes: a flip
        # needs a uniform within ~1e-7 of a marginal, absent for these seeds
        assert c.counts() == ref.counts(), (n, s)
        for col in range(c.nr_columns()):
            a, b = c.amplitudes(col), ref.amplitudes(col)
            fid = abs(np.vdot(b, a)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
            assert 1.0 - fid <= 1e-6, (n, s, col, 1.0 - fid)


def test_batch_waves_under_small_budget():
    wl = WL[:10]
    big, circs_big = run_batch(wl, "c128")
    small, circs_small = run_batch(wl, "c128", budget=1 << 20)
    assert small.stats()["waves"] > 1 and big.stats()["waves"] == 1
    for a, b in zip(circs_big, circs_small):
        assert np.array_equal(np.array(a.cstate()), np.array(b.cstate()))


def test_batch_budget_counts_dense_basis_columns():
    """After a collapsing measure_all, X/Y restoring gates, a reset, or a
    later X/Y measurement turn the run's basis columns into up to nr_shots
    dense columns.  The wave planner must count them: 12-qubit circuits with
    64 runs take 64 x 64 KiB = 4 MiB each, so a 10 MiB budget holds at most
    two per wave.  Classical states equal the solo runs and the oracle's."""
    n = 12
    wl = []
    tails = [
        [("measure_all", list(range(n)), "X")],
        [("measure_all", list(range(n)), "Y"), ("peek_all", list(range(n, 2 * n)), "X")],
        [("measure_all", list(range(n)), "Z"), ("reset", 1), ("measure", 1, n, "Z")],
        [("measure_all", list(range(n)), "Z"), ("measure", 2, n, "Y")],
        [("measure_all", list(range(n)), "Z"), ("peek", 3, n, "X")],
        [("measure_all", list(range(n)), "Y")],
    ]
    for k, tail in enumerate(tails):
        ops = workloads.random_layered(n, 5, 70 + k)
        prog = [("gate",) + tuple(o) for o in ops] + tail
        wl.append((n, 2 * n, prog, 500 + k))
    batch, circs = run_batch(wl, "c128", budget=10 << 20)
    assert batch.stats()["waves"] >= 3
    for (n_, nc, prog, s), c in zip(wl, circs):
        solo = make(n_, nc, prog, s, "c128")
        solo.execute(SHOTS)
        got = np.array(c.cstate(), dtype=np.uint64)
        assert np.array_equal(got, np.array(solo.cstate(), dtype=np.uint64)), s
        r = make_ref(n_, nc, prog)
        r.execute(SHOTS, s)
        assert np.array_equal(got, r.cstate()), s


def test_batch_same_circuit_many_seeds():
    # one program, 32 seeds: every copy's histogram is the one it gives alone
    n, nc, prog, _ = WL[0]
    wl = [(n, nc, prog, 1000 + k) for k in range(32)]
    _, circs = run_batch(wl, "c128")
    total = 0
    for (n_, nc_, prog_, s), c in zip(wl, circs):
        solo = make(n_, nc_, prog_, s, "c128")
        solo.execute(SHOTS)
        assert c.histogram() == solo.histogram()
        total += sum(c.histogram_vec())
    assert total == 32 * SHOTS


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_autotune_keeps_results(prec):
    # the autotuner executes a sample under each configuration; afterwards
    # every circuit still gives what it gives alone with its seed
    batch = Batch(prec)
    circs = [make(n, nc, prog, s, prec) for (n, nc, prog, s) in WL]
    for c in circs:
        batch.add(c)
    tried = batch.autotune(SHOTS, sample=8)
    assert len(tried) == (2 if prec == "c128" else 5) and all(t[4] > 0 for t in tried)
    best = min(tried, key=lambda t: t[4])
    assert batch.config() == best[:4]
    batch.execute(SHOTS)
    for (n, nc, prog, s), c in zip(WL, circs):
        solo = make(n, nc, prog, s, prec, best[3])
        solo.execute(SHOTS)
        assert np.array_equal(np.array(c.cstate()), np.array(solo.cstate())), (n, s)


@pytest.mark.parametrize("chunk", [1, 3, 64])
def test_tile_chunk_keeps_results(chunk):
    """Runs of consecutive tiles per workgroup (the batched sweep's schedule)
    change nothing: end states equal the default schedule's bit for bit and
    the solo path's to FP32 rounding, including chunks that straddle states
    and a chunk longer than most states."""
    wl = [(n, nc, prog[:-1], s) for (n, nc, prog, s) in WL]
    runs = []
    for ch in (None, chunk):
        batch = Batch("c64", 0, 0)
        circs = [make(n, nc, prog, s, "c64") for (n, nc, prog, s) in wl]
        for c in circs:
            batch.add(c)
        if ch is not None:
            batch.set_tile_chunk(ch)
        batch.execute(SHOTS, keep_states=True)
        runs.append(circs)
    for (n, nc, prog, s), c0, c1 in zip(wl, runs[0], runs[1]):
        assert c0.counts() == c1.counts(), (n, s)
        for col in range(c1.nr_columns()):
            assert np.array_equal(c0.amplitudes(col), c1.amplitudes(col)), (n, s, col)
        solo = make(n, nc, prog, s, "c64")
        solo.execute(SHOTS)
        assert c1.counts() == solo.counts(), (n, s)
        for col in range(c1.nr_columns()):
            assert np.max(np.abs(c1.amplitudes(col) - solo.amplitudes(col))) <= 1e-5, (n, s, col)


def test_tile_chunk_is_checked():
    batch = Batch("c64", 0, 0)
    for bad in (0, 2000):
        with pytest.raises(Exception):
            batch.set_tile_chunk(bad)


def test_block_cache_after_stream_teardown():
    """Blocks cached from streams that no longer exist (a finished batch's
    shared stream, states' own streams) are reused or freed safely when later
    releases push the cache over its limit (set to 36 GiB here)."""
    import gc
    from q1tsim_amd.gvs import GpuVectorState, device_cache_bytes, set_device_cache_limit
    set_device_cache_limit(36 << 30)
    try:
        batch, circs = run_batch(WL[:6], "c64")
        del batch, circs
        gc.collect()
        a, b = GpuVectorState(30, 1, "c128"), GpuVectorState(30, 1, "c128")   # 16 GiB each, both live
        a.close()
        b.close()
        c = GpuVectorState(29, 1, "c128")   # 8 GiB, a new size: its release goes over the limit
        c.close()
        assert device_cache_bytes() <= 36 << 30
        batch, circs = run_batch(WL, "c128")
        for (n, nc, prog, s), c in zip(WL, circs):
            solo = make(n, nc, prog, s, "c128")
            solo.execute(SHOTS)
            assert np.array_equal(np.array(c.cstate()), np.array(solo.cstate())), (n, s)
        set_device_cache_limit(1 << 30)   # lowering the limit frees down to it at once
        assert device_cache_bytes() <= 1 << 30
    finally:
        set_device_cache_limit(0)


def test_batch_column_work_is_batched():
    """Every measurement, measure_all and conditional gate of the branching
    workload runs as column work of the whole batch (one launch per kind per
    pass): no op circuit by circuit, and a number of launches per round that
    does not grow with the number of circuits; results as alone."""
    wl = workloads.batch_workload(96, seed=5, qubits=(8, 16), depths=(4, 10), meas=(1, 3))
    batch, circs = run_batch(wl, "c64")
    st = batch.stats()
    assert st["solo_ops"] == 0
    assert 0 < st["column_launches"] <= 40 * st["rounds"]
    for (n, nc, prog, s), c in list(zip(wl, circs))[:24]:
        solo = make(n, nc, prog, s, "c64")
        solo.execute(SHOTS)
        assert np.array_equal(np.array(c.cstate(), dtype=np.uint64), np.array(solo.cstate(), dtype=np.uint64)), (n, s)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_batch_conditional_gates_on_split_columns(prec):
    """A conditional gate controlled by a peeked bit: the runs of one column
    disagree, so the column set changes and the gate runs circuit by circuit
    (solo_ops); a conditional gate after a collapsing measurement runs as
    column work.  Mixed in one batch with X/Y measurements, every classical
    state equals the solo run's and the oracle's."""
    n = 9
    wl = []
    for k in range(6):
        ops = [("gate",) + tuple(o) for o in workloads.random_layered(n, 4, 300 + k)]
        prog = ops[:20] + [("peek", k % n, n, "Z"), ("cgate", [n], 1, "rx", [(k + 2) % n], [0.4 + k]),
                           ("measure", (k + 3) % n, n + 1, "X" if k % 2 else "Z"),
                           ("cgate", [n + 1], 1, "x", [(k + 5) % n], [])]
        prog += ops[20:] + [("measure_all", list(range(n)), "Z")]
        wl.append((n, n + 2, prog, 900 + k))
    batch, circs = run_batch(wl, prec)
    st = batch.stats()
    assert st["solo_ops"] >= 1 and st["column_launches"] > 0
    for (n_, nc, prog, s), c in zip(wl, circs):
        solo = make(n_, nc, prog, s, prec)
        solo.execute(SHOTS)
        got = np.array(c.cstate(), dtype=np.uint64)
        assert np.array_equal(got, np.array(solo.cstate(), dtype=np.uint64)), s
        r = make_ref(n_, nc, prog)
        r.execute(SHOTS, s)
        assert np.array_equal(got, r.cstate()), s
