"""GPU parity of the tiled multi-gate sweeps (the plain c128 path's HBM
schedule) and of the per-gate kernels they fall back to."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from oracle import oracle as orc  # noqa: E402
from q1tsim_amd import workloads  # noqa: E402
from q1tsim_amd.gvs import GpuVectorState  # noqa: E402

pytestmark = pytest.mark.gpu


def run(n, ops, amps, prec, tiling, ncols_ctrl=None):
    g = GpuVectorState.from_amplitudes(amps, 4 if ncols_ctrl is not None else 1, prec)
    g.set_plan(0, tiling)
    if ncols_ctrl is not None:
        g.apply_conditional_gate(ncols_ctrl, orc.gate_matrix("rx", [0.9]), [n - 1])
    for name, bits, params in ops:
        g.apply_gate(orc.gate_matrix(name, params), bits)
    g.flush()
    return g


def oracle(n, ops, amps, ncols_ctrl=None):
    r = orc.RefState(n, 4 if ncols_ctrl is not None else 1, amplitudes=amps)
    if ncols_ctrl is not None:
        r.apply_conditional_gate(ncols_ctrl, "rx", [n - 1], [0.9])
    for name, bits, params in ops:
        r.apply_gate(name, bits, params)
    return r.states


def rand_state(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


def mixed_ops(n, count, seed):
    """Every sweepable kind: dense one-qubit, controlled (controls in and out
    of the tile), phases, diagonals on any bit, plus non-sweepable gates
    (swap, dense two-qubit) that cut sweeps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    names = ["h", "x", "y", "rx", "ry", "rz", "s", "t", "u3", "v", "z", "cx", "cz", "ch", "crx", "crz",
             "cu1", "ccx", "ccz", "cy", "cu3", "ccrz", "swap"]
    ops = []
    for _ in range(count):
        name = names[int(rng.integers(len(names)))]
        kind, nctrl, npar = orc.GATES[name]
        k = (2 if kind == orc.SWAP else 1) + nctrl
        bits = [int(b) for b in rng.permutation(n)[:k]]
        ops.append((name, bits, [float(v) for v in rng.uniform(0, 2 * np.pi, npar)]))
    return ops


@pytest.mark.parametrize("n", [14, 16, 19])
def test_sweep_mixed_c128(n):
    a = rand_state(n, n)
    ops = mixed_ops(n, 300, n)
    g = run(n, ops, a, "c128", True)
    st = g.stats()
    assert st["sweeps"] >= 1
    ref = oracle(n, ops, a)
    err = float(np.abs(g.states - ref).max())
    assert err <= 1e-12, err


def test_sweep_matches_per_gate_exactly_layered():
    n = 18
    ops = workloads.random_layered(n, 20, 1)
    g1 = run(n, ops, rand_state(n, 1), "c128", True)
    g0 = run(n, ops, rand_state(n, 1), "c128", False)
    d = float(np.abs(g1.states - g0.states).max())
    assert d <= 1e-13, d
    st = g1.stats()
    print("layered n=18: %d gates in %d sweeps, %d launches" % (st["gates"], st["sweeps"], st["launches"]))
    assert st["sweeps"] < 40


def test_sweep_multicolumn():
    n = 15
    a = rand_state(n, 3)
    ops = mixed_ops(n, 120, 3)
    ctrl = [1, 0, 0, 1]
    g = run(n, ops, a, "c128", True, ncols_ctrl=ctrl)
    ref = oracle(n, ops, a, ncols_ctrl=ctrl)
    assert g.counts == [1, 2, 1]
    err = float(np.abs(g.states - ref).max())
    assert err <= 1e-12, err


@pytest.mark.parametrize("n", [15, 20])
def test_sweep_c64(n):
    a = rand_state(n, 7)
    ops = mixed_ops(n, 200, 7)
    g = run(n, ops, a, "c64", True)
    ref = oracle(n, ops, a)
    col = g.states[:, 0]
    err = float(np.abs(col - ref[:, 0]).max())
    fid = abs(np.vdot(ref[:, 0], col)) ** 2
    assert err <= 1e-5 and 1 - fid < 1e-6, (err, 1 - fid)


@pytest.mark.parametrize("k", [3, 4, 5, 6])
def test_fused_c64_blocks(k):
    """Dense k-qubit blocks composed in FP64, applied with FP32 MFMA."""
    n = 18
    a = rand_state(n, 20 + k)
    ops = mixed_ops(n, 250, 20 + k) + workloads.random_layered(n, 6, k)
    g = GpuVectorState.from_amplitudes(a, 1, "c64")
    g.set_plan(k, True)
    for name, bits, params in ops:
        g.apply_gate(orc.gate_matrix(name, params), bits)
    g.flush()
    st = g.stats()
    col = g.column(0)
    ref = oracle(n, ops, a)[:, 0]
    err = float(np.abs(col - ref).max())
    nrm = float(np.vdot(col, col).real)
    infid = 1 - abs(np.vdot(ref, col)) ** 2 / nrm
    print("fused k=%d: sweeps %d blocks %d max-abs %.2e infidelity %.2e" % (k, st["sweeps"], st["blocks"], err,
                                                                            infid))
    assert st["blocks"] > 0
    assert err <= 1e-5 and infid < 1e-6, (err, infid)


def test_fused_layered_config1():
    n = 18
    ops = workloads.random_layered(n, 20, 1)
    g = GpuVectorState(n, 1, "c64")
    g.set_plan(4, True)
    for name, bits, params in ops:
        g.apply_gate(orc.gate_matrix(name, params), bits)
    col = g.column(0)
    r = orc.RefState(n, 1)
    for name, bits, params in ops:
        r.apply_gate(name, bits, params)
    ref = r.states[:, 0]
    nrm = float(np.vdot(col, col).real)
    infid = 1 - abs(np.vdot(ref, col)) ** 2 / nrm
    err = float(np.abs(col - ref).max())
    assert err <= 1e-5 and infid < 1e-6, (err, infid)


def test_qft_sweeps():
    n = 20
    g = GpuVectorState(n, 1)
    g.set_plan(0, True)
    g.set_random_normal(3)
    a = g.column(0)
    for name, bits, params in workloads.qft(n):
        g.apply_gate(orc.gate_matrix(name, params), bits)
    out = g.column(0)
    err = float(np.abs(out - np.sqrt(1 << n) * np.fft.ifft(a)).max())
    assert err < 1e-11, err


@pytest.mark.parametrize("prec,geom", [("c128", (12, 3, 4)), ("c128", (12, 4, 4)), ("c128", (12, 4, 5)),
                                       ("c64", (12, 4, 4)), ("c64", (12, 3, 4)), ("c64", (13, 4, 5)),
                                       ("c64", (13, 4, 4))])
def test_every_tile_shape(prec, geom):
    # each compiled tile shape, with controls in and out of the tile,
    # against the oracle (n = 16: several tiles per column)
    n = 16
    ops = mixed_ops(n, 120, 71)
    amps = rand_state(n, 72)
    g = GpuVectorState.from_amplitudes(amps, 1, prec)
    g.set_plan(0, True)
    g.set_geometry(*geom)
    for name, bits, params in ops:
        g.apply_gate(orc.gate_matrix(name, params), bits)
    g.flush()
    got = g.column(0)
    ref = oracle(n, ops, amps)[:, 0]
    if prec == "c128":
        assert np.max(np.abs(got - ref)) <= 1e-12
    else:
        fid = abs(np.vdot(ref, got)) ** 2 / np.vdot(got, got).real
        assert 1 - fid <= 1e-6 and np.max(np.abs(got - ref)) <= 1e-5
    assert g.stats()["sweeps"] > 0


def test_unsupported_tile_shape_is_rejected():
    g = GpuVectorState(14, 1, "c128")
    with pytest.raises(Exception):
        g.set_geometry(13, 4, 5)


def swap_network(n, seed):
    """Swaps that reverse the qubit order, then chains that share bits
    (swap(a, b) followed by swap(b, c)), between one-qubit gates."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = [("swap", [j, n - 1 - j], []) for j in range(n // 2)]
    for _ in range(40):
        a, b, c = (int(x) for x in rng.permutation(n)[:3])
        ops += [("swap", [a, b], []), ("h", [b], []), ("swap", [b, c], []), ("t", [a], [])]
    return ops


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("n", [14, 17])
def test_sweep_search_mixed(n, prec):
    """Tiles and register phases refined by local search (on by default only
    for states of 512 MiB and more, forced on here) and SWAP gates applied
    as write-back permutations: amplitudes as the oracle has them."""
    a = rand_state(n, 40 + n)
    ops = mixed_ops(n, 250, 40 + n) + swap_network(n, n) + mixed_ops(n, 100, 41 + n)
    g = GpuVectorState.from_amplitudes(a, 1, prec)
    g.set_plan(0, True)
    g.set_plan_search(1)
    for name, bits, params in ops:
        g.apply_gate(orc.gate_matrix(name, params), bits)
    g.flush()
    st = g.stats()
    ref = oracle(n, ops, a)[:, 0]
    col = g.states[:, 0]
    err = float(np.abs(col - ref).max())
    print("search n=%d %s: %d gates, %d sweeps, %d phases, %d launches, max-abs %.2e" %
          (n, prec, st["gates"], st["sweeps"], st["phases"], st["launches"], err))
    if prec == "c128":
        assert err <= 1e-12, err
    else:
        fid = abs(np.vdot(ref, col)) ** 2 / float(np.vdot(col, col).real)
        assert err <= 1e-5 and 1 - fid < 1e-6, (err, 1 - fid)


def test_search_and_swaps_keep_per_gate_result():
    """The same circuit with searched tiles, greedy tiles and per-gate
    kernels (no tiles): complex128 results agree to rounding."""
    n = 19
    a = rand_state(n, 77)
    ops = workloads.random_layered(n, 12, 77) + swap_network(n, 5) + workloads.qft(n)
    res = {}
    for mode, tiling in ((1, True), (0, True), (0, False)):
        g = GpuVectorState.from_amplitudes(a, 1, "c128")
        g.set_plan(0, tiling)
        g.set_plan_search(mode)
        for name, bits, params in ops:
            g.apply_gate(orc.gate_matrix(name, params), bits)
        g.flush()
        res[(mode, tiling)] = (g.states[:, 0].copy(), g.stats())
    base = res[(0, False)][0]
    for key in ((1, True), (0, True)):
        d = float(np.abs(res[key][0] - base).max())
        assert d <= 1e-12, (key, d)
    assert res[(1, True)][1]["sweeps"] <= res[(0, True)][1]["sweeps"]


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_back_to_back_flushes(prec):
    """Flushes one after another with no amplitudes read between them: each
    flush's sweep program replaces the previous one's in the same device
    table; the result is the whole gate stream's."""
    n = 20
    a = rand_state(n, 31)
    segments = [mixed_ops(n, 12, 100 + s) for s in range(10)]
    g = GpuVectorState.from_amplitudes(a, 1, prec)
    g.set_plan(4 if prec == "c64" else 0, True)
    sweeps = 0
    for ops in segments:
        for name, bits, params in ops:
            g.apply_gate(orc.gate_matrix(name, params), bits)
        g.flush()
        sweeps += g.stats()["sweeps"]   # stats are per flush
    ref = oracle(n, [op for ops in segments for op in ops], a)[:, 0]
    col = g.states[:, 0]
    err = float(np.abs(col - ref).max())
    assert sweeps >= 10
    if prec == "c128":
        assert err <= 1e-12, err
    else:
        fid = abs(np.vdot(ref, col)) ** 2 / float(np.vdot(col, col).real)
        assert err <= 1e-5 and 1 - fid < 1e-6, (err, 1 - fid)


@pytest.mark.parametrize("prec,fk", [("c128", 0), ("c64", 0), ("c64", 4)])
def test_flush_graph_replays(prec, fk):
    """Flush graphs: one circuit shape flushed five times, twice with its own
    angles, twice with new ones, then its own again.  The first flush
    captures, later ones replay with their sweep programs re-uploaded; a
    reduction and a read-back run eagerly between flushes.  Every flush
    equals the eager path's bit for bit and the oracle."""
    n = 18
    a = rand_state(n, 41)
    base = workloads.random_layered(n, 6, 3)
    rng = np.random.Generator(np.random.PCG64(5))

    def new_angles():
        return [(g, b, [float(x) for x in rng.uniform(0.1, 6.0, len(p))]) for g, b, p in base]

    rounds = [base, base, new_angles(), new_angles(), base]
    eager = GpuVectorState.from_amplitudes(a, 1, prec)
    graph = GpuVectorState.from_amplitudes(a, 1, prec)
    for g in (eager, graph):
        g.set_plan(fk, True)
    graph.set_graph(True)
    ref = a
    sweeps_only = True
    for ops in rounds:
        for g in (eager, graph):
            for name, bits, params in ops:
                g.apply_gate(orc.gate_matrix(name, params), bits)
            g.flush()
        st = eager.stats()
        sweeps_only = sweeps_only and st["launches"] == st["sweeps"]
        assert graph.stats()["sweeps"] == st["sweeps"]
        e, x = eager.states[:, 0], graph.states[:, 0]
        assert np.array_equal(e, x)
        assert np.array_equal(eager.marginal_w0(n - 1), graph.marginal_w0(n - 1))
        ref = oracle(n, ops, ref)[:, 0]
        err = float(np.abs(x - ref).max())
        if prec == "c128":
            assert err <= 1e-12, err
        else:
            fid = abs(np.vdot(ref, x)) ** 2 / float(np.vdot(x, x).real)
            assert err <= 1e-5 and 1 - fid < 1e-6, (err, 1 - fid)
    captures, replays = graph.graph_stats()
    assert captures + replays == len(rounds) and captures >= 1
    assert eager.graph_stats() == (0, 0)
    if sweeps_only:
        # no single-gate kernel holds a matrix by value: every later flush replays
        assert replays == len(rounds) - 1, (captures, replays)


def test_flush_graph_mode_is_checked():
    g = GpuVectorState(12, 1, "c128")
    with pytest.raises(Exception):
        g.set_graph(2)


def test_fused_search_c64():
    n = 17
    a = rand_state(n, 9)
    ops = workloads.random_pairs(n, 20, 9) + swap_network(n, 9)
    g = GpuVectorState.from_amplitudes(a, 1, "c64")
    g.set_plan(4, True)
    g.set_plan_search(1)
    for name, bits, params in ops:
        g.apply_gate(orc.gate_matrix(name, params), bits)
    g.flush()
    ref = oracle(n, ops, a)[:, 0]
    col = g.states[:, 0]
    fid = abs(np.vdot(ref, col)) ** 2 / float(np.vdot(col, col).real)
    err = float(np.abs(col - ref).max())
    assert g.stats()["blocks"] > 0
    assert err <= 1e-5 and 1 - fid < 1e-6, (err, 1 - fid)


def test_plan_search_mode_is_checked():
    g = GpuVectorState(12, 1, "c128")
    with pytest.raises(Exception):
        g.set_plan_search(2)


def test_copy_bandwidth_is_plausible():
    """The measured device copy (the streaming ceiling a sweep is compared
    with) lands between 10 % and 100 % of the 8 TB/s HBM peak; bad sizes are
    rejected."""
    from q1tsim_amd.gvs import copy_bandwidth
    ms, gbps = copy_bandwidth(1 << 30, 3)
    assert ms > 0 and 800.0 < gbps < 8000.0, (ms, gbps)
    with pytest.raises(Exception):
        copy_bandwidth(0, 1)
