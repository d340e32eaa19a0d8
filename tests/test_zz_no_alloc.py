"""The allocation-free marginal (q1t_gvs_marginal_w0_noalloc) through the C
ABI: bit-identical to q1t_gvs_marginal_w0 on two calls in a row, the second
on a smaller state, for both precisions and several branch columns; sizes
beyond its scratch are rejected with an error."""
import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd import workloads
from q1tsim_amd.gvs import GpuVectorState, Xoshiro
from q1tsim_amd._lib import Q1tError

pytestmark = pytest.mark.gpu


def prepared(n, depth, seed, precision, nr_meas, shots=64):
    s = GpuVectorState(n, shots, precision)
    for name, bits, params in workloads.random_layered(n, depth, seed):
        s.apply_gate(orc.gate_matrix(name, params), bits)
    res = np.zeros(shots, np.uint64)
    rng = Xoshiro(seed)
    for m in range(nr_meas):   # each collapse may split the columns
        s.measure_into(m, m, res, rng)
    s.flush()
    return s


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_noalloc_marginal_matches_marginal(precision):
    big = prepared(18, 6, 3, precision, 3)
    small = prepared(12, 5, 4, precision, 1)
    assert big.ncols > 1
    for st in (big, small):       # second call: a smaller state
        for q in (0, 5, st.nr_bits - 1):
            exp = st.marginal_w0(q)
            got = st.marginal_w0_noalloc(q)
            assert got.shape == exp.shape
            assert np.array_equal(got, exp), (st.nr_bits, q, got, exp)
            # a probability, up to the state's norm drift (FP32 amplitudes
            # drift by ~1e-7 per gate)
            assert np.all((got >= 0) & (got <= 1 + (1e-5 if precision == "c64" else 1e-9)))
    big.close()
    small.close()


def test_noalloc_marginal_limits():
    # 9 qubits in |+>, each measured: up to 512 columns > 256
    n, shots = 9, 2048
    s = GpuVectorState(n, shots, "c128")
    for q in range(n):
        s.apply_gate(orc.gate_matrix("h", []), [q])
    res = np.zeros(shots, np.uint64)
    rng = Xoshiro(1)
    for q in range(n):
        s.measure_into(q, q, res, rng)
    assert s.ncols > 256
    with pytest.raises(Q1tError, match="at most 256 columns"):
        s.marginal_w0_noalloc(0)
    s.close()
    t = GpuVectorState(10, 4, "c128")
    t.apply_gate(orc.gate_matrix("h", []), [2])
    with pytest.raises(Q1tError, match="Gates are pending"):
        t.marginal_w0_noalloc(2)
    t.flush()
    assert np.array_equal(t.marginal_w0_noalloc(2), t.marginal_w0(2))
    t.close()
