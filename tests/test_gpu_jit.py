"""Sweep kernels specialised at run time (hipRTC) on the GPU: the same
amplitudes as the interpreter kernels bit for bit (the same op bodies), the
oracle's to 1e-12 in complex128, the same classical results at the circuit
level, and flush graphs replaying them."""
import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd import workloads
from q1tsim_amd.gvs import GpuVectorState, gate_matrix, jit_stats
from q1tsim_amd.q1tsim import Circuit

pytestmark = pytest.mark.gpu


def mixed(n, seed):
    ops = workloads.random_pairs(n, 8, seed) + workloads.qft(min(n, 9)) + workloads.random_layered(n, 3, seed + 1)
    ops += [("y", [1], []), ("rx", [2], [0.3]), ("ccx", [0, 3, n - 1], []), ("cz", [n - 1, 2], []),
            ("swap", [0, n - 2], []), ("u3", [4], [0.1, 0.2, 0.3]), ("crz", [n - 3, 1], [0.7]),
            ("ch", [2, n - 1], []), ("tdg", [n - 1], []), ("cy", [n - 2, 5], []), ("sdg", [n - 4], []),
            ("cu1", [n - 1, n - 5], [0.9]), ("v", [6], [])]
    return ops


def run(n, ops, precision, fusion_k, jit, cols=1, prefetch=False, geometry=None, factors=True):
    s = GpuVectorState(n, max(cols, 1), precision)
    s.set_plan(fusion_k, True)
    s.set_plan_search(1)
    if geometry:
        s.set_geometry(*geometry)
    s.set_jit(jit, prefetch, factors)
    if cols > 1:   # several branch columns: measure qubits in |+>
        for q in range(int(np.log2(cols))):
            s.apply_gate(gate_matrix("h", []), [q])
        res = np.zeros(cols, np.uint64)
        from q1tsim_amd.gvs import Xoshiro
        rng = Xoshiro(3)
        for q in range(int(np.log2(cols))):
            s.measure_into(q, q, res, rng)
    for g, b, p in ops:
        s.apply_gate(gate_matrix(g, p), b, g)
    s.flush()
    out = [s.column(c) for c in range(s.ncols)]
    st = s.stats()
    s.close()
    return out, st


@pytest.mark.parametrize("precision,fusion_k", [("c128", 0), ("c64", 0), ("c64", 4)])
@pytest.mark.parametrize("n", [14, 17])
def test_jit_equals_interpreter(precision, fusion_k, n):
    """Without tile factors the specialised kernels run the interpreter's
    program with the same op bodies: equal amplitudes bit for bit.  With
    them (the default) diagonal ops outside the tile become one product per
    tile, a different rounding order: equal to the interpreter within a few
    ulps, and the prefetching kernels equal the plain ones bit for bit."""
    ops = mixed(n, 40 + n)
    c0 = jit_stats()[0]
    # the same tile shape for all (the specialised kernels' default differs)
    geom = (12, 4, 4)
    a, st_a = run(n, ops, precision, fusion_k, False, geometry=geom)
    b, st_b = run(n, ops, precision, fusion_k, True, geometry=geom, factors=False)
    assert st_a["sweeps"] == st_b["sweeps"] and st_b["sweeps"] > 0
    assert jit_stats()[0] >= c0
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    c, st_c = run(n, ops, precision, fusion_k, True, geometry=geom)
    d, _ = run(n, ops, precision, fusion_k, True, geometry=geom, prefetch=True)
    assert st_c["sweeps"] == st_a["sweeps"]
    tol = 1e-13 if precision == "c128" else 2e-6
    for x, y, z in zip(a, c, d):
        assert np.abs(x - y).max() <= tol
        assert np.array_equal(y, z)
    if precision == "c128":
        r = orc.RefState(n, 1)
        for g, bb, p in ops:
            r.apply_gate(g, bb, p)
        assert np.abs(b[0] - r.states[:, 0]).max() <= 1e-12
        assert np.abs(c[0] - r.states[:, 0]).max() <= 1e-12


def test_jit_several_columns_and_wide_blocks():
    n = 16
    ops = mixed(n, 7)
    a, _ = run(n, ops, "c64", 6, False, cols=4)
    b, _ = run(n, ops, "c64", 6, True, cols=4)
    assert len(a) == len(b) >= 2
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_jit_kernels_are_reused_for_new_angles():
    """One kernel per program structure: new angles reuse it."""
    n = 16
    base = workloads.random_layered(n, 6, 3)
    run(n, base, "c128", 0, True)
    compiled = jit_stats()[0]
    rng = np.random.Generator(np.random.PCG64(5))
    moved = [(g, b, [float(rng.uniform(0, 6))] if p else p) for g, b, p in base]
    x, _ = run(n, moved, "c128", 0, True)
    assert jit_stats()[0] == compiled
    r = orc.RefState(n, 1)
    for g, b, p in moved:
        r.apply_gate(g, b, p)
    assert np.abs(x[0] - r.states[:, 0]).max() <= 1e-12


def test_jit_circuit_and_graph_replays():
    """Circuit-level switch: classical states equal the interpreter's; with
    flush graphs on too, reexecutes replay graphs of specialised kernels."""
    n = 15
    prog = workloads.random_layered(n, 4, 9)
    results = []
    for jit, graph in ((False, False), (True, False), (True, True)):
        c = Circuit(n, 2)
        for g, b, p in prog:
            c.add_gate(g, b, p if p else None)
        c.peek(1, 0)
        for g, b, p in prog[:20]:
            c.add_gate(g, b, p if p else None)
        c.peek(3, 1)
        c.set_jit(jit)
        c.set_graph(graph)
        c.seed(12)
        c.execute(64)
        seq = [(list(c.cstate()), c.amplitudes(0))]
        for _ in range(2):
            c.reexecute()
            seq.append((list(c.cstate()), c.amplitudes(0)))
        results.append((seq, c.graph_stats()))
        c.close()
    # the specialised kernels default to another tile shape, so the plan
    # (and rounding) differs from the interpreter's: same classical results,
    # amplitudes to 1e-12; with and without graphs they agree bit for bit
    ref = results[0][0]
    for (cs0, a0), (cs1, a1), (cs2, a2) in zip(ref, results[1][0], results[2][0]):
        assert cs0 == cs1 == cs2
        assert np.abs(a0 - a1).max() <= 1e-12
        assert np.array_equal(a1, a2)
    assert results[2][1] == (2, 4), results[2][1]
