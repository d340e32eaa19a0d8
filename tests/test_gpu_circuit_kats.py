"""The reference's circuit tests (src/circuit.rs:1456-1985) through the C-ABI
Circuit on the GPU statevector: the reference's own checks on every
histogram, and the classical state bit-identical to the oracle's op loop
(RefCircuit) for the same seed.  Also the product gate table (the names
Circuit.add_gate resolves) against the dense-simulator fixture."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd.q1tsim import Circuit
from tests import circuit_kats
from tests.test_circuit_kats import make_ref

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def make_gpu(nq, nc, prog, precision="c128"):
    c = Circuit(nq, nc)
    if precision != "c128":
        c.set_backend(precision, 0, True)
    for op in prog:
        kind = op[0]
        if kind == "gate":
            c.add_gate(op[1], op[2], op[3] if op[3] else None)
        elif kind == "cgate":
            c.add_conditional_gate(op[1], op[2], op[3], op[4], op[5] if op[5] else None)
        elif kind == "measure":
            c.measure_basis(op[1], op[2], op[3])
        elif kind == "peek":
            c.peek_basis(op[1], op[2], op[3])
        elif kind == "measure_all":
            c.measure_all_basis(op[1], op[2])
        elif kind == "peek_all":
            c.peek_all_basis(op[1], op[2])
        elif kind == "reset":
            c.reset(op[1])
        elif kind == "reset_all":
            c.reset_all()
        else:
            raise ValueError(kind)
    return c


@pytest.mark.parametrize("entry", circuit_kats.KATS, ids=[k[0] for k in circuit_kats.KATS])
def test_reference_circuit_kats_on_gpu(entry):
    name, nq, nc, shots, prog, check = entry
    for seed in (1, 2):
        c = make_gpu(nq, nc, prog)
        c.seed(seed)
        c.execute(shots)
        cs = [int(v) for v in c.cstate()]
        hv, h = circuit_kats.histograms(cs, nc)
        assert check(hv, h, cs, shots), (name, seed)
        assert c.histogram_vec() == hv
        r = make_ref(orc.RefCircuit, nq, nc, prog)
        r.execute(shots, seed)
        assert cs == [int(v) for v in r.cstate()], (name, seed)
        c.close()


def test_reference_circuit_kats_complex64():
    """The deterministic KATs also hold on the complex64 state."""
    for name, nq, nc, shots, prog, check in circuit_kats.KATS:
        c = make_gpu(nq, nc, prog, "c64")
        c.seed(3)
        c.execute(shots)
        cs = [int(v) for v in c.cstate()]
        hv, h = circuit_kats.histograms(cs, nc)
        assert check(hv, h, cs, shots), name
        c.close()


def test_histogram_string_keys_on_gpu():
    """circuit.rs:1897-1921, and key 0 of a circuit without classical bits
    prints as "0" ({:0width$b} with width 0, circuit.rs:831)."""
    with make_gpu(2, 2, circuit_kats.HIST_PROG) as c:
        c.seed(5)
        c.execute(4096)
        h = c.histogram()
        assert sorted(h) == ["00", "01", "10", "11"] and sum(h.values()) == 4096
    with Circuit(1, 0) as c:
        c.h(0)
        c.execute(8)
        assert c.histogram() == {"0": 8}


@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64", 1e-5)])
def test_product_gate_table_on_fixture(precision, tol):
    """all_gates_n6.npz (dense simulator, 117 gates over all 39 names) run
    through Circuit.add_gate, i.e. the library's own gate table, not the
    oracle's matrices."""
    d = np.load(os.path.join(HERE, "golden", "all_gates_n6.npz"))
    n = int(d["n"])
    with Circuit(n, 0) as c:
        if precision != "c128":
            c.set_backend(precision, 0, True)
        c.set_initial_amplitudes(d["input"])
        for name, bits, params in zip(d["names"], d["bits"], d["params"]):
            name = str(name)
            b = [int(x) for x in bits if x >= 0]
            npar = {"rx": 1, "ry": 1, "rz": 1, "u1": 1, "u2": 2, "u3": 3}.get(name.lstrip("c"), 0)
            c.add_gate(name, b, [float(x) for x in params[:npar]] if npar else None)
        c.execute(1)
        a = c.amplitudes(0)
    assert np.abs(a - d["output"]).max() < tol
