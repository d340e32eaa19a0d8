"""X- and Y-basis measurement folded into the measurement kernels
(circuit.rs:669-748): one-qubit measure / peek read the rotated marginal and
collapse with the rotation back in one pass each; a collapsing measure_all
keeps its basis columns with the rotation back in closed form.  Checked
against the oracle, whose op loop applies the basis-change gates as the
reference does (and is pinned by the reference's circuit tests)."""
import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd import workloads
from q1tsim_amd.gvs import GpuVectorState, Xoshiro, gate_matrix
from q1tsim_amd.q1tsim import Circuit
from tests.test_circuit_kats import make_ref
from tests.test_gpu_circuit_kats import make_gpu

pytestmark = pytest.mark.gpu


def ref_measure(r, q, c, basis, res, rng, collapse):
    """The op loop's basis change around the oracle's measure (circuit.rs:669-718)."""
    if basis == "X":
        r.apply_gate("h", [q])
    elif basis == "Y":
        r.apply_gate("sdg", [q])
        r.apply_gate("h", [q])
    r.measure_into(q, c, res, rng, collapse=collapse)
    if basis == "X":
        r.apply_gate("h", [q])
    elif basis == "Y":
        r.apply_gate("h", [q])
        r.apply_gate("s", [q])


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_state_basis_measure_and_peek_vs_oracle(precision):
    n, shots = 10, 64
    ops = workloads.random_layered(n, 6, 8)
    g = GpuVectorState(n, shots, precision)
    r = orc.RefState(n, shots)
    for name, bits, params in ops:
        g.apply_gate(gate_matrix(name, params), bits)
        r.apply_gate(name, bits, params)
    res_g = np.zeros(shots, np.uint64)
    res_r = np.zeros(shots, np.uint64)
    rng_g, rng_r = Xoshiro(21), orc.Rng(21)
    seq = [(0, 0, "X", True), (3, 1, "Y", False), (5, 2, "Y", True), (1, 3, "X", False), (7, 4, "Z", True),
           (2, 5, "X", True), (9, 6, "Y", True), (4, 7, "X", True)]
    for k, (q, c, basis, collapse) in enumerate(seq):
        g.measure_into(q, c, res_g, rng_g, collapse=collapse, basis=basis)
        ref_measure(r, q, c, basis, res_r, rng_r, collapse)
        assert np.array_equal(res_g, res_r), (k, q, basis)
        assert g.counts == list(r.counts), k
        # a gate between measurements: the next marginal sees the new columns
        g.apply_gate(gate_matrix("ry", [0.3 + k]), [(q + 1) % n])
        r.apply_gate("ry", [(q + 1) % n], [0.3 + k])
    tol = 1e-12 if precision == "c128" else 2e-5
    for col in range(g.ncols):
        assert np.abs(g.column(col) - r.states[:, col]).max() <= tol, col
    g.close()


def test_basis_measurements_apply_no_gates():
    """A circuit of X/Y measurements and peeks only: no basis-change gate is
    ever applied (the state's flushes apply no gate)."""
    with Circuit(12, 4) as c:
        c.set_initial_random(5)
        c.measure_x(2, 0)
        c.peek_y(5, 1)
        c.measure_y(7, 2)
        c.peek_x(11, 3)
        c.seed(3)
        c.execute(32)
        assert c.stats()["gates"] == 0
        cs = list(c.cstate())
    r = orc.RefCircuit(12, 4)
    r.initial_amplitudes = orc.random_normal_state(12, 5)
    r.measure(2, 0, "X")
    r.peek(5, 1, "Y")
    r.measure(7, 2, "Y")
    r.peek(11, 3, "X")
    r.execute(32, 3)
    assert cs == [int(v) for v in r.cstate()]


@pytest.mark.parametrize("n", [7, 8])
@pytest.mark.parametrize("basis", ["X", "Y"])
def test_measure_all_basis_columns_vs_oracle(n, basis):
    """A collapsing X/Y measure_all keeps compact columns with the rotation
    back in closed form: the columns written out equal the oracle's dense
    columns (H^n|k> or S^n H^n|k>, odd and even n), and later gates and
    measurements continue from them."""
    prog = [("gate", name, bits, params) for name, bits, params in workloads.random_layered(n, 4, 2)]
    prog.append(("measure_all", list(range(n)), basis))
    shots = 48
    c = make_gpu(n, 2 * n + 2, prog)
    c.seed(9)
    c.execute(shots)
    r = make_ref(orc.RefCircuit, n, 2 * n + 2, prog)
    r.execute(shots, 9)
    assert [int(v) for v in c.cstate()] == [int(v) for v in r.cstate()]
    assert c.counts() == list(r.q_state.counts)
    for col in range(c.nr_columns()):
        assert np.abs(c.amplitudes(col) - r.q_state.states[:, col]).max() <= 1e-12, col
    c.close()
    more = prog + [("gate", "h", [1], []), ("measure", 1, 2 * n, "Z"), ("measure", 0, 2 * n + 1, basis),
                   ("measure_all", list(range(n, 2 * n)), "Z")]
    c = make_gpu(n, 2 * n + 2, more)
    c.seed(10)
    c.execute(shots)
    r = make_ref(orc.RefCircuit, n, 2 * n + 2, more)
    r.execute(shots, 10)
    assert [int(v) for v in c.cstate()] == [int(v) for v in r.cstate()]
    c.close()


def test_repeated_measure_all_in_same_basis_is_deterministic():
    """measure_all X twice: the second reads the first's outcomes exactly
    (compact rotated columns measured in their own basis)."""
    n = 9
    prog = [("gate", name, bits, params) for name, bits, params in workloads.random_layered(n, 3, 4)]
    prog += [("measure_all", list(range(n)), "X"), ("measure_all", list(range(n, 2 * n)), "X")]
    c = make_gpu(n, 2 * n, prog)
    c.seed(2)
    c.execute(40)
    mask = (1 << n) - 1
    for v in c.cstate():
        assert int(v) & mask == (int(v) >> n) & mask
    r = make_ref(orc.RefCircuit, n, 2 * n, prog)
    r.execute(40, 2)
    assert [int(v) for v in c.cstate()] == [int(v) for v in r.cstate()]
    c.close()
