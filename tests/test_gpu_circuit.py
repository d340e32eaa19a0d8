"""Circuit-level parity on the GPU: the C++ op interpreter over the GPU
statevector vs the oracle's restated interpreter (RefCircuit), same seed ->
identical classical state; plus the reference's circuit tests
(circuit.rs:1456-1985) and the python/ smoke scripts."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd import workloads
from q1tsim_amd.q1tsim import Circuit, Q1tError, RefParam
from tests import stats_helper

pytestmark = pytest.mark.gpu


def build_pair(nq, nc, program):
    c = Circuit(nq, nc)
    r = orc.RefCircuit(nq, nc)
    for op in program:
        kind = op[0]
        if kind == "gate":
            _, name, bits, params = op
            c.add_gate(name, bits, params if params else None)
            r.add_gate(name, bits, params)
        elif kind == "cgate":
            _, control, target, name, bits, params = op
            c.add_conditional_gate(control, target, name, bits, params if params else None)
            r.add_conditional_gate(control, target, name, bits, params)
        elif kind == "measure":
            c.measure_basis(op[1], op[2], op[3])
            r.measure(op[1], op[2], op[3])
        elif kind == "peek":
            c.peek_basis(op[1], op[2], op[3])
            r.peek(op[1], op[2], op[3])
        elif kind == "measure_all":
            c.measure_all_basis(op[1], op[2])
            r.measure_all(op[1], op[2])
        elif kind == "peek_all":
            c.peek_all_basis(op[1], op[2])
            r.peek_all(op[1], op[2])
        elif kind == "reset":
            c.reset(op[1])
            r.reset(op[1])
        elif kind == "reset_all":
            c.reset_all()
            r.reset_all()
    return c, r


def run_pair(c, r, shots, seed):
    c.seed(seed)
    c.execute(shots)
    r.execute(shots, seed)
    return np.array(c.cstate(), dtype=np.uint64), r.cstate()


def test_mid_circuit_measure_conditional_reset():
    prog = [
        ("gate", "h", [0], []), ("gate", "ry", [1], [1.1]), ("gate", "cx", [0, 2], []),
        ("measure", 0, 0, "Z"), ("cgate", [0], 1, "x", [3], []),
        ("gate", "u3", [2], [0.3, 0.2, -0.9]), ("measure", 2, 1, "X"),
        ("cgate", [0, 1], 2, "rz", [1], [0.77]), ("peek", 1, 2, "Y"),
        ("reset", 3), ("gate", "ch", [1, 3], []), ("measure_all", [4, 3, 2, 1], "Z"),
    ]
    c, r = build_pair(4, 5, prog)
    got, exp = run_pair(c, r, 500, 99)
    assert np.array_equal(got, exp)
    assert c.counts() == r.q_state.counts
    assert c.histogram() == r.histogram_string()


def test_measure_all_xy_bases_and_peek_all():
    prog = [("gate", "h", [0], []), ("gate", "s", [1], []), ("gate", "rx", [2], [0.4]),
            ("peek_all", [0, 1, 2], "X"), ("peek_all", [3, 4, 5], "Y"), ("measure_all", [0, 1, 2], "Z")]
    c, r = build_pair(3, 6, prog)
    got, exp = run_pair(c, r, 1000, 5)
    assert np.array_equal(got, exp)


def test_config4_like_branching():
    """Random circuit with 1-3 mid-circuit measurements and conditional X/RZ, 64 runs."""
    rng = np.random.Generator(np.random.PCG64(44))
    for trial in range(3):
        n = int(rng.integers(6, 11))
        ops = workloads.random_layered(n, int(rng.integers(4, 9)), trial)
        prog = [("gate",) + tuple(o) for o in ops]
        for m in range(int(rng.integers(1, 4))):
            pos = int(rng.integers(1, len(prog)))
            q = int(rng.integers(n))
            prog.insert(pos, ("measure", q, m, "Z"))
            if rng.uniform() < 0.5:
                prog.insert(pos + 1, ("cgate", [m], 1, "x", [int(rng.integers(n))], []))
            else:
                prog.insert(pos + 1, ("cgate", [m], 1, "rz", [int(rng.integers(n))], [float(rng.uniform(0, 6))]))
        prog.append(("measure_all", list(range(3, 3 + n)), "Z"))
        c, r = build_pair(n, 3 + n, prog)
        got, exp = run_pair(c, r, 64, 1000 + trial)
        assert np.array_equal(got, exp)


def test_conditional_injected_like_reference():
    """circuit.rs:1619-1651 pattern: control word bits, control[0] is the LSB."""
    prog = [("gate", "x", [0], []), ("measure", 0, 0, "Z"), ("gate", "h", [1], []), ("measure", 1, 1, "Z"),
            ("cgate", [1, 0], 1, "x", [2], []), ("measure", 2, 2, "Z")]
    c, r = build_pair(3, 3, prog)
    got, exp = run_pair(c, r, 400, 3)
    assert np.array_equal(got, exp)
    # word = c1 | c0 << 1 == 1  <=>  c1 == 1 and c0 == 0: never (c0 is always 1)
    assert all(((int(v) >> 2) & 1) == 0 for v in got)


def test_reexecute_and_reference_parameters():
    """python/test.py: RefParam values are read at execute time."""
    ang = RefParam(1.23)
    with Circuit(1, 1) as c:
        c.rx(ang, 0)
        c.peek(0, 0)
        c.seed(1)
        c.execute(4000)
        p1 = c.histogram().get("1", 0) / 4000
        ang.assign(math.pi)
        c.reexecute()
        p2 = c.histogram().get("1", 0) / 4000
    assert abs(p1 - math.sin(1.23 / 2) ** 2) < 0.05
    # reexecute starts from the previous end state (circuit.rs:602-641), so
    # the second pass applies RX(pi) to RX(1.23)|0>.
    assert abs(p2 - math.sin((math.pi + 1.23) / 2) ** 2) < 0.05


def test_reexecute_replays_flush_graphs():
    """Flush graphs on a circuit (hipGraph): reexecute keeps the state, so its
    gate flush replays the previous execute's graph, with the angles of the
    reference parameters re-read each time; every pass equals the same
    circuit run without graphs, bit for bit."""
    n = 16
    base = workloads.random_layered(n, 4, 2)
    runs = []
    for graph in (False, True):
        refs = []
        c = Circuit(n, 1)
        for name, bits, params in base:
            if params:
                r = RefParam(params[0])
                refs.append(r)
                c.add_gate(name, bits, [r])
            else:
                c.add_gate(name, bits)
        c.peek(0, 0)
        c.seed(4)
        c.set_graph(graph)
        c.execute(8)
        amps = [c.amplitudes(0)]
        rng = np.random.Generator(np.random.PCG64(9))
        for _ in range(3):
            for r in refs:
                r.assign(float(rng.uniform(0.1, 6.0)))
            c.reexecute()
            amps.append(c.amplitudes(0))
        runs.append((amps, c.graph_stats(), c))
    (a0, g0, _), (a1, g1, _) = runs
    for x, y in zip(a0, a1):
        assert np.array_equal(x, y)
    assert g0 == (0, 0)
    assert g1 == (1, 3), g1


def test_reexecute_replays_each_flush_graph():
    """A circuit whose gates are split by peeks flushes three times per
    execute; each flush has its own graph, and every reexecute replays all
    three (the state keeps a graph per launch signature).  Classical states
    and amplitudes equal the same circuit run without graphs."""
    n = 15
    parts = [workloads.random_layered(n, 2 + k, 20 + k) for k in range(3)]
    runs = []
    for graph in (False, True):
        c = Circuit(n, 3)
        for k, part in enumerate(parts):
            for name, bits, params in part:
                c.add_gate(name, bits, params if params else None)
            c.peek(k, k)
        c.seed(8)
        c.set_graph(graph)
        c.execute(16)
        res = [(list(c.cstate()), c.amplitudes(0))]
        for _ in range(3):
            c.reexecute()
            res.append((list(c.cstate()), c.amplitudes(0)))
        runs.append((res, c.graph_stats()))
        c.close()
    (r0, g0), (r1, g1) = runs
    for (cs0, a0), (cs1, a1) in zip(r0, r1):
        assert cs0 == cs1
        assert np.array_equal(a0, a1)
    assert g0 == (0, 0)
    assert g1 == (3, 9), g1


def test_cstate_script():
    """python/test_cstate.py, with the binomial band of stats.rs."""
    with Circuit(2, 2) as c:
        c.h(0)
        c.h(1)
        c.measure(0, 0)
        c.measure(1, 1)
        c.execute(1000)
        cs = c.cstate()
    assert len(cs) == 1000
    assert stats_helper.measurement_ok(sum(v & 1 for v in cs), 1000, 0.5, 1e-5)
    assert stats_helper.measurement_ok(sum((v >> 1) & 1 for v in cs), 1000, 0.5, 1e-5)


def test_circuit_measure_all_histogram():  # circuit.rs:1654-1686
    with Circuit(2, 2) as c:
        c.x(0)
        c.measure_all([1, 0])
        c.execute(10)
        assert c.histogram() == {"10": 10} or c.histogram() == {"01": 10}
        hv = c.histogram_vec()
    # qubit 0 -> cbit 1: key 0b10
    assert hv[0b10] == 10


def test_execute_errors_surface():
    with Circuit(2, 1) as c:
        c.measure_all([0])     # 1 cbit for 2 qubits: InvalidNrMeasurementBits at execute
        with pytest.raises(Q1tError, match="Expected 2 measurement bits, but got 1"):
            c.execute(5)


def test_amplitude_readout_and_backend_switch():
    n = 12
    ops = workloads.random_layered(n, 10, 4)
    ref = orc.RefState(n, 1)
    for name, bits, params in ops:
        ref.apply_gate(name, bits, params)
    for prec, tol in (("c128", 1e-12), ("c64", 1e-5)):
        with Circuit(n, 0) as c:
            c.set_backend(prec, 0, True)
            for name, bits, params in ops:
                c.add_gate(name, bits, params if params else None)
            c.execute(1)
            a = c.amplitudes(0)
            assert np.abs(a - ref.states[:, 0]).max() < tol
            st = c.stats()
            assert st["gates"] == len(ops)
