"""The headline workloads at their full sizes on one MI355X, checked through
size-independent properties (a circuit followed by its inverse returns the
input state), plus the same recipes at sizes the oracle runs in seconds,
compared with the oracle directly:

* 30-qubit depth-20 layered random circuit, complex128 (16 GiB): round trip
  within 1e-11;
* 30-qubit QFT on a random normal state: QFT then its inverse returns the
  input;
* 32-qubit depth-40 random-pair circuit, complex64 (32 GiB), greedy and
  learned fusion plans: round trip to |0...0> at infidelity <= 1e-6; the
  same recipe at 20 qubits against the oracle for both plans;
* branching circuits of 20 qubits, depth 30, mid-circuit measurements with
  conditional gates, 64 runs, in a batch: classical state bit-identical to
  the oracle's op loop."""
import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd import costmodel, workloads
from q1tsim_amd.costmodel import CostModel
from q1tsim_amd.gvs import GpuVectorState, compare, gate_matrix
from q1tsim_amd.q1tsim import Batch, Circuit

pytestmark = pytest.mark.gpu


def apply_ops(s, ops):
    for g, b, p in ops:
        s.apply_gate(gate_matrix(g, p), b, g)
    s.flush()


def test_config2_layered30_round_trip_c128():
    n = 30
    ops = workloads.random_layered(n, 20, 1)
    s = GpuVectorState(n, 1, "c128")
    apply_ops(s, ops)
    mid = s.stats()
    apply_ops(s, workloads.inverse(ops))
    zero = GpuVectorState(n, 1, "c128")
    infid, maxabs = compare(s, zero)
    print("30q layered c128 round trip: infidelity %.3e max-abs %.3e (%d sweeps forward)" % (infid, maxabs,
                                                                                            mid["sweeps"]))
    assert maxabs <= 1e-11 and infid <= 1e-12, (infid, maxabs)
    s.close()
    zero.close()


def test_config2_qft30_round_trip_c128():
    n = 30
    s = GpuVectorState(n, 1, "c128")
    s.set_random_normal(3)
    ref = GpuVectorState(n, 1, "c128")
    ref.set_random_normal(3)
    ops = workloads.qft(n)
    apply_ops(s, ops)
    infid0, _ = compare(s, ref)
    assert infid0 > 0.5          # the transform moved the state
    apply_ops(s, workloads.inverse(ops))
    infid, maxabs = compare(s, ref)
    print("30q QFT round trip: infidelity %.3e max-abs %.3e" % (infid, maxabs))
    assert infid <= 1e-12 and maxabs <= 1e-13, (infid, maxabs)
    s.close()
    ref.close()


@pytest.fixture(scope="module")
def small_model():
    """A cost model recorded and trained on 20-qubit states (seconds)."""
    circs = [workloads.random_pairs(20, 12, 900 + k) for k in range(6)]
    circs += [workloads.random_layered(20, 8, 950 + k) for k in range(3)]
    feat, tiles, ms = costmodel.record(20, circs, fusion_k=4, seed=3, reps=2)
    cm = CostModel(5)
    costmodel.fit(cm, feat, tiles, ms, epochs=60)
    return cm


@pytest.mark.parametrize("mode", [0, 1])
def test_config5_random_pairs32_round_trip_c64(small_model, mode):
    n = 32
    ops = workloads.random_pairs(n, 40, 5)
    s = GpuVectorState(n, 1, "c64")
    s.set_plan(4, True)
    s.set_cost_model(small_model, mode, 1)
    apply_ops(s, ops)
    fwd = s.stats()
    apply_ops(s, workloads.inverse(ops))
    zero = GpuVectorState(n, 1, "c64")
    infid, maxabs = compare(s, zero)
    print("32q random pairs c64 (%s): forward %.3f s, %d sweeps, %d blocks; round trip infidelity %.3e "
          "max-abs %.3e" % ("learned" if mode else "greedy", fwd["gate_seconds"], fwd["sweeps"], fwd["blocks"],
                            infid, maxabs))
    assert infid <= 1e-6 and maxabs <= 1e-3, (infid, maxabs)
    s.close()
    zero.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_config5_recipe_vs_oracle_n20(small_model, mode):
    n = 20
    ops = workloads.random_pairs(n, 40, 77)
    r = orc.RefState(n, 1)
    for g, b, p in ops:
        r.apply_gate(g, b, p)
    exp = r.states[:, 0]
    s = GpuVectorState(n, 1, "c64")
    s.set_plan(4, True)
    s.set_cost_model(small_model, mode, 1)
    apply_ops(s, ops)
    a = s.column(0)
    fid = abs(np.vdot(exp, a)) ** 2 / (np.vdot(a, a).real * np.vdot(exp, exp).real)
    assert 1.0 - fid <= 1e-6 and np.abs(a - exp).max() <= 1e-5, (1.0 - fid, np.abs(a - exp).max())
    assert s.stats()["blocks"] > 0
    p = GpuVectorState(n, 1, "c128")
    apply_ops(p, ops)
    assert np.abs(p.column(0) - exp).max() <= 1e-12
    s.close()
    p.close()


BRANCH_SHOTS = 64


@pytest.fixture(scope="module")
def branch_case():
    """One 20-qubit depth-30 branching circuit (2 mid-circuit measurements
    with conditional gates, measure_all), run once on the oracle's op loop."""
    n, seed = 20, 31
    nc, prog = workloads.branching(n, 30, 2, 4242)
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
    r.execute(BRANCH_SHOTS, seed)
    return n, nc, prog, seed, np.array(r.cstate(), dtype=np.uint64), len(r.q_state.counts)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_config4_branching_n20_vs_oracle(branch_case, prec):
    n, nc, prog, seed, expect, ref_cols = branch_case
    batch = Batch(prec)
    c = Circuit(n, nc)
    workloads.build_program(c, prog)
    c.set_backend(prec, 0)
    c.seed(seed)
    batch.add(c)
    # small companions so the batch has several widths and items
    others = []
    for k, (m, d) in enumerate(((14, 12), (17, 20))):
        nc2, prog2 = workloads.branching(m, d, 2, 77 + k)
        o = Circuit(m, nc2)
        workloads.build_program(o, prog2)
        o.set_backend(prec, 0)
        o.seed(5 + k)
        batch.add(o)
        others.append(o)
    batch.execute(BRANCH_SHOTS)
    got = np.array(c.cstate(), dtype=np.uint64)
    # complex64: the branch and sampling draws come from FP64 sums of FP32
    # amplitudes; a flip needs a uniform within ~1e-7 (relative) of a
    # marginal or cumulative weight, absent for this seed
    assert np.array_equal(got, expect), (prec, int(np.sum(got != expect)))
    assert ref_cols >= 2
