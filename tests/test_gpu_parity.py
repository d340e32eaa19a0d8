"""GPU parity: the HIP GpuVectorState (through the C ABI) against the CPU
oracle, the golden fixtures and closed forms.

Bars: amplitudes <= 1e-12 max-abs for complex128 (plain path); outcome
indices bit-exact for the same uniforms (canonical F=16 tree order); cstate
bit-exact for the same seed (same Bernoulli/uniform stream); complex64 within
1e-5 max-abs; histograms chi-square consistent with |a|^2.
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import dense_sim  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from q1tsim_amd import workloads  # noqa: E402
from q1tsim_amd.gvs import GpuVectorState, Xoshiro  # noqa: E402
from tests import stats_helper  # noqa: E402

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 1e-5
z, o, i_ = 0j, 1 + 0j, 1j
x = math.sqrt(0.5) + 0j


def close(a, b, tol):
    a, b = np.asarray(a, dtype=np.complex128), np.asarray(b, dtype=np.complex128)
    assert a.shape == b.shape, (a.shape, b.shape)
    err = float(np.abs(a - b).max()) if a.size else 0.0
    assert err <= tol, err
    return err


def M(name, params=()):
    return orc.gate_matrix(name, params)


def run_both(n, ops, amps=None, precision="c128", plan=None):
    """Apply ops to the GPU state and to the oracle; return both columns."""
    if amps is None:
        g = GpuVectorState(n, 1, precision)
        r = orc.RefState(n, 1)
    else:
        g = GpuVectorState.from_amplitudes(amps, 1, precision)
        r = orc.RefState(n, 1, amplitudes=amps)
    if plan is not None:
        g.set_plan(*plan)
    for name, bits, params in ops:
        g.apply_gate(M(name, params), bits)
        r.apply_gate(name, bits, params)
    return g.column(0), r.states[:, 0], g


# ------------------------------------------------------------ reference KATs

def test_new_and_product_state():  # vectorstate.rs:426-470
    s = GpuVectorState(3, 1500)
    assert s.counts == [1500]
    close(s.states, [[o], [z], [z], [z], [z], [z], [z], [z]], 0)
    s = GpuVectorState.from_qubit_coefs([o, o, -i_, z], 9)
    close(s.states, [[-x * i_], [z], [-x * i_], [z]], 1e-15)
    s = GpuVectorState.from_qubit_coefs([z, o, o, z], 13)
    close(s.states, [[z], [z], [o], [z]], 0)


def test_apply_gates_kat():  # vectorstate.rs:641-707
    h = 0.5 * o
    s = GpuVectorState(3, 1)
    s.apply_gate(M("h"), [1])
    close(s.states, [[x], [z], [x], [z], [z], [z], [z], [z]], 1e-15)
    s = GpuVectorState(3, 1)
    s.apply_gate(M("y"), [2])
    close(s.states, [[z], [i_], [z], [z], [z], [z], [z], [z]], 0)
    s = GpuVectorState.from_qubit_coefs([z, o, o, z, o, z], 1)
    s.apply_gate(M("cx"), [0, 2])
    close(s.states, [[z], [z], [z], [z], [z], [o], [z], [z]], 0)
    s = GpuVectorState.from_qubit_coefs([z, o, o, z, o, z], 1)
    s.apply_gate(np.kron(M("h"), M("h")), [1, 2])
    close(s.states, [[z], [z], [z], [z], [h], [h], [h], [h]], 1e-15)
    s = GpuVectorState.from_qubit_coefs([x, -x, x, -x, x, -x], 1)
    s.apply_gate(M("ccx"), [0, 2, 1])
    hx = 0.5 * x
    close(s.states, [[hx], [-hx], [-hx], [hx], [-hx], [-hx], [hx], [hx]], 1e-15)


def test_apply_conditional_gate_kat():  # vectorstate.rs:473-509
    s = GpuVectorState(2, 5)
    s.apply_conditional_gate([0, 0, 1, 1, 0], M("x"), [1])
    assert s.counts == [2, 2, 1]
    close(s.states, [[o, z, o], [z, o, z], [z, z, z], [z, z, z]], 0)
    s = GpuVectorState.from_qubit_coefs([o, z, x, x], 5)
    s.apply_conditional_gate([1, 0, 1, 1, 0], M("cx"), [1, 0])
    assert s.counts == [1, 1, 2, 1]
    close(s.states, [[x, x, x, x], [z, x, z, x], [z, z, z, z], [x, z, x, z]], 1e-15)
    s = GpuVectorState(2, 5)
    s.apply_conditional_gate([1, 1, 1, 0, 0], M("h"), [0])
    assert s.counts == [3, 2]
    s.apply_conditional_gate([0, 0, 1, 1, 1], M("h"), [0])
    assert s.counts == [2, 1, 2]
    close(s.states, [[x, o, x], [z, z, z], [x, z, x], [z, z, z]], 1e-15)


def test_measure_all_bits_kat():  # vectorstate.rs:710-735
    rng = Xoshiro(5)
    s = GpuVectorState.from_qubit_coefs([z, o, z, o, o, z], 5)
    r = np.zeros(5, np.uint64)
    s.measure_all_into([0, 1, 2], r, rng)
    assert all(int(b) == 0b011 for b in r)
    s = GpuVectorState(3, 5)
    s.apply_gate(M("h"), [2])
    r = np.zeros(5, np.uint64)
    s.measure_all_into([0, 1, 2], r, rng)
    assert all((int(b) & 0b011) == 0 for b in r)


def test_errors():
    s = GpuVectorState(3, 4)
    from q1tsim_amd._lib import Q1tError
    with pytest.raises(Q1tError, match="Expected 2 bits"):
        s.apply_gate(M("cx"), [0])
    with pytest.raises(Q1tError, match="Invalid index 3 for a quantum bit"):
        s.apply_gate(M("h"), [3])
    with pytest.raises(Q1tError, match="Not enough space to store 4 measurement results in array of length 2"):
        s.measure_into(0, 0, np.zeros(2, np.uint64), Xoshiro(1))
    with pytest.raises(Q1tError, match="Expected 3 measurement bits, but got 2"):
        s.measure_all_into([0, 1], np.zeros(4, np.uint64), Xoshiro(1))
    with pytest.raises(Q1tError, match="The number of runs is 4, but received 3 control bits"):
        s.apply_conditional_gate([1, 0, 1], M("x"), [0])
    zero = GpuVectorState.from_amplitudes(np.zeros(8, complex), 4)
    with pytest.raises(Q1tError, match="zero or invalid norm"):
        zero.measure_all_uniforms([0, 1, 2], np.full(4, 0.5), np.zeros(4, np.uint64))


# -------------------------------------------------------- oracle parity

@pytest.mark.parametrize("fname", ["layered_n10_d20_s7.npz", "layered_n8_d12_s3.npz"])
def test_layered_fixture(fname):
    d = np.load(os.path.join(HERE, "golden", fname))
    n = int(d["n"])
    ops = workloads.random_layered(n, int(d["depth"]), int(d["seed"]))
    g, r, _ = run_both(n, ops)
    close(g, d["amplitudes"], TOL128)
    close(g, r, TOL128)


def test_all_gates_fixture():
    d = np.load(os.path.join(HERE, "golden", "all_gates_n6.npz"))
    ops = []
    for name, b, p in zip(d["names"], d["bits"], d["params"]):
        name = str(name)
        ops.append((name, [int(v) for v in b if v >= 0], [float(v) for v in p[:orc.GATES[name][2]]]))
    g, r, _ = run_both(6, ops, amps=d["input"])
    close(g, d["output"], TOL128)
    close(g, r, TOL128)


def test_qft_fixture_and_closed_form():
    d = np.load(os.path.join(HERE, "golden", "qft_n8.npz"))
    g, r, _ = run_both(8, workloads.qft(8), amps=d["input"])
    close(g, d["output"], TOL128)
    n = 16
    s = GpuVectorState(n, 1)
    s.set_random_normal(42)
    a = s.column(0)
    close(a, orc.random_normal_state(n, 42), 1e-14)
    for name, bits, params in workloads.qft(n):
        s.apply_gate(M(name, params), bits)
    close(s.column(0), np.sqrt(1 << n) * np.fft.ifft(a), 1e-11)


def test_config1_circuit_amplitudes():
    """18-qubit depth-20 layered circuit (530 gates), complex128 vs oracle."""
    n = 18
    ops = workloads.random_layered(n, 20, 1)
    assert len(ops) == 530
    g, r, s = run_both(n, ops)
    err = close(g, r, TOL128)
    print("config-1 max-abs error vs oracle: %.3e" % err)


def test_config1_c64_amplitudes():
    n = 18
    ops = workloads.random_layered(n, 20, 1)
    g, r, _ = run_both(n, ops, precision="c64")
    err = close(g, r, TOL64)
    # per-gate FP32 arithmetic: FP32-rounded unitaries let the norm drift by
    # ~1e-7 per gate; the state infidelity compares directions
    norm = float(np.vdot(g, g).real)
    fid = abs(np.vdot(r, g)) ** 2 / norm
    print("c64 max-abs %.3e infidelity %.3e norm-1 %.3e" % (err, 1 - fid, norm - 1))
    assert 1 - fid < 1e-6
    assert abs(norm - 1) < 1e-4


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
def test_random_gates_small_and_ragged(n):
    """Every gate kind on random qubits, small n (tree levels smaller than F)."""
    rng = np.random.Generator(np.random.PCG64(100 + n))
    ops = []
    names = [k for k, v in orc.GATES.items() if (2 if v[0] == orc.SWAP else 1) + v[1] <= n]
    for _ in range(60):
        name = names[int(rng.integers(len(names)))]
        kind, nctrl, npar = orc.GATES[name]
        k = (2 if kind == orc.SWAP else 1) + nctrl
        bits = [int(b) for b in rng.permutation(n)[:k]]
        ops.append((name, bits, [float(v) for v in rng.uniform(0, 2 * np.pi, npar)]))
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    g, r, _ = run_both(n, ops, amps=a)
    close(g, r, TOL128)


def test_generic_dense_gates():
    """User gates known only by matrix: dense 3..7-qubit unitaries and controls found from structure."""
    rng = np.random.Generator(np.random.PCG64(3))
    n = 9
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    g = GpuVectorState.from_amplitudes(a, 1)
    ref = a.copy()
    for k in [3, 4, 5, 6, 7, 2, 1]:
        q, _ = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
        bits = [int(b) for b in rng.permutation(n)[:k]]
        g.apply_gate(q, bits)
        ref = dense_sim.apply(ref, q, bits, n)
        # a controlled version: identity block on the first qubit = 0
        cq = np.eye(2 << k, dtype=complex)
        cq[1 << k:, 1 << k:] = q
        cbits = [int(b) for b in rng.permutation(n)[:k + 1]]
        if k + 1 <= 8:
            g.apply_gate(cq, cbits)
            ref = dense_sim.apply(ref, cq, cbits, n)
    close(g.column(0), ref, 1e-12)


def test_multi_column_gates():
    """Gates after a conditional split act on every branch column."""
    n = 6
    rng = np.random.Generator(np.random.PCG64(9))
    a = rng.normal(size=64) + 1j * rng.normal(size=64)
    a /= np.linalg.norm(a)
    g = GpuVectorState.from_amplitudes(a, 6)
    r = orc.RefState(n, 6, amplitudes=a)
    ctrl = [1, 0, 0, 1, 1, 0]
    g.apply_conditional_gate(ctrl, M("ry", [0.4]), [3])
    r.apply_conditional_gate(ctrl, "ry", [3], [0.4])
    for name, bits, params in workloads.random_layered(n, 6, 2) + [("swap", [0, 5], []), ("cu1", [1, 4], [0.3])]:
        g.apply_gate(M(name, params), bits)
        r.apply_gate(name, bits, params)
    assert g.counts == r.counts
    close(g.states, r.states, TOL128)


# ------------------------------------------------------------ measurement

@pytest.mark.parametrize("n,nruns", [(1, 50), (2, 64), (4, 100), (5, 333), (12, 4096), (18, 8192)])
def test_measure_all_uniforms_bit_exact(n, nruns):
    rng = np.random.Generator(np.random.PCG64(n))
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a[rng.integers(0, 1 << n, size=max(1, (1 << n) // 7))] = 0
    a /= np.linalg.norm(a)
    u = rng.uniform(size=nruns)
    cb = list(rng.permutation(n + 2)[:n])
    res_g = np.zeros(nruns, np.uint64)
    res_r = np.zeros(nruns, np.uint64)
    g = GpuVectorState.from_amplitudes(a, nruns)
    r = orc.RefState(n, nruns, amplitudes=a)
    g.measure_all_uniforms(cb, u, res_g, collapse=True)
    r.measure_all_uniforms(cb, u, res_r, collapse=True, mode=0)
    assert np.array_equal(res_g, res_r)
    assert g.counts == r.counts
    # the sequential (rand 0.7 WeightedIndex) order agrees on these uniforms
    r2 = orc.RefState(n, nruns, amplitudes=a)
    res_s = np.zeros(nruns, np.uint64)
    r2.measure_all_uniforms(cb, u, res_s, collapse=False, mode=1)
    assert np.array_equal(res_s, res_r)


def test_sampling_bit_exact_multicolumn_and_c64():
    n = 10
    rng = np.random.Generator(np.random.PCG64(77))
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    for prec in ("c128", "c64"):
        g = GpuVectorState.from_amplitudes(a, 600, prec)
        ctrl = (rng.uniform(size=600) < 0.5).astype(np.uint8)
        g.apply_conditional_gate(ctrl, M("h"), [4])
        # the oracle on exactly the GPU's columns (c64 values widen exactly)
        cols = g.states
        u = rng.uniform(size=600)
        res_g = np.zeros(600, np.uint64)
        g.measure_all_uniforms(list(range(n)), u, res_g, collapse=False)
        off = 0
        exp = np.zeros(600, np.uint64)
        for c, cnt in enumerate(g.counts):
            w = orc.canonical_weights(cols[:, c])
            idx, _ = orc.tree_sample_many(w, u[off:off + cnt])
            idx = np.sort(idx)
            for k in range(cnt):
                exp[off + k] = orc.reverse_bits(int(idx[k]), n)
            off += cnt
        assert np.array_equal(res_g, exp), prec


def test_measure_collapse_matches_oracle():
    n = 8
    rng = np.random.Generator(np.random.PCG64(5))
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    g = GpuVectorState.from_amplitudes(a, 1000)
    r = orc.RefState(n, 1000, amplitudes=a)
    rg, rr = Xoshiro(31), orc.Rng(31)
    for q in (3, 0, 7, 3):
        res_g, res_r = np.zeros(1000, np.uint64), np.zeros(1000, np.uint64)
        g.measure_into(q, q, res_g, rg)
        r.measure_into(q, q, res_r, rr)
        assert np.array_equal(res_g, res_r)
        assert g.counts == r.counts
        close(g.states, r.states, TOL128)
    close(g.marginal_w0(5), r.marginal_w0(5), 1e-13)
    # peek leaves the state alone
    res_g, res_r = np.zeros(1000, np.uint64), np.zeros(1000, np.uint64)
    g.measure_into(2, 9, res_g, rg, collapse=False)
    r.measure_into(2, 9, res_r, rr, collapse=False)
    assert np.array_equal(res_g, res_r)
    close(g.states, r.states, TOL128)


def test_reset_and_reset_all():
    n = 5
    rng = np.random.Generator(np.random.PCG64(6))
    a = rng.normal(size=32) + 1j * rng.normal(size=32)
    a /= np.linalg.norm(a)
    g = GpuVectorState.from_amplitudes(a, 40)
    r = orc.RefState(n, 40, amplitudes=a)
    rg, rr = Xoshiro(8), orc.Rng(8)
    g.reset(2, rg)
    r.reset(2, rr)
    assert g.counts == r.counts
    close(g.states, r.states, TOL128)
    g.reset_all()
    r.reset_all()
    close(g.states, r.states, 0)


def test_measure_all_then_more_ops():
    """measure_all leaves compact basis columns; later ops still agree with the oracle."""
    n = 6
    rng = np.random.Generator(np.random.PCG64(12))
    a = rng.normal(size=64) + 1j * rng.normal(size=64)
    a /= np.linalg.norm(a)
    g = GpuVectorState.from_amplitudes(a, 300)
    r = orc.RefState(n, 300, amplitudes=a)
    rg, rr = Xoshiro(4), orc.Rng(4)
    res_g, res_r = np.zeros(300, np.uint64), np.zeros(300, np.uint64)
    g.measure_all_into(list(range(n)), res_g, rg)
    r.measure_all_into(list(range(n)), res_r, rr)
    assert np.array_equal(res_g, res_r)
    assert g.counts == r.counts
    # measure again on basis columns, then gates force dense columns
    g.measure_into(1, 7, res_g, rg)
    r.measure_into(1, 7, res_r, rr)
    assert np.array_equal(res_g, res_r)
    g.apply_gate(M("h"), [2])
    r.apply_gate("h", [2])
    close(g.states, r.states, TOL128)


def test_histogram_chi2_config1():
    n = 18
    s = GpuVectorState(n, 8192)
    ops = workloads.random_layered(n, 20, 1)
    for name, bits, params in ops:
        s.apply_gate(M(name, params), bits)
    a = s.column(0)
    res = np.zeros(8192, np.uint64)
    s.measure_all_into(list(range(n)), res, Xoshiro(2024))
    # group outcomes into 64 bins by their top 6 qubits (cbits 0..5)
    p = np.abs(a) ** 2
    keys = (np.arange(1 << n) >> (n - 6))          # index bits of qubits 0..5
    pb = np.bincount(keys, weights=p, minlength=64)
    obs_keys = np.array([orc.reverse_bits(int(v), n) for v in res]) >> (n - 6)
    cnt = np.bincount(obs_keys, minlength=64)
    ok, stat, dof = stats_helper.chi2_ok(cnt, pb, 8192)
    assert ok, (stat, dof)


# -------------------------------------------------- full-size properties

def test_qft22_random_vs_fft():
    """22-qubit QFT on the seeded random normal state vs sqrt(N) ifft (numpy)."""
    n = 22
    s = GpuVectorState(n, 1)
    s.set_plan(0, True)
    s.set_random_normal(7)
    a = s.column(0)
    for name, bits, params in workloads.qft(n):
        s.apply_gate(M(name, params), bits)
    err = float(np.abs(s.column(0) - np.sqrt(1 << n) * np.fft.ifft(a)).max())
    print("qft22 max-abs vs fft: %.3e" % err)
    assert err < 1e-12


def test_qft30_basis_state_closed_form():
    """30-qubit QFT (16 GiB complex128) of |x>: amplitude y = exp(2 pi i x y / N) / sqrt(N)."""
    n = 30
    x0 = 0x2AAAAAAB
    N = 1 << n
    s = GpuVectorState(n, 1)
    s.set_plan(0, True)
    ops = [("x", [q], []) for q in range(n) if (x0 >> (n - 1 - q)) & 1] + workloads.qft(n)
    for name, bits, params in ops:
        s.apply_gate(M(name, params), bits)
    out = s.column(0)
    err = 0.0
    for lo in range(0, N, 1 << 26):
        y = np.arange(lo, lo + (1 << 26), dtype=np.uint64)
        ph = ((np.uint64(x0) * y) & np.uint64(N - 1)).astype(np.float64) / N
        ref = np.exp(2j * np.pi * ph) / np.sqrt(N)
        err = max(err, float(np.abs(out[lo:lo + (1 << 26)] - ref).max()))
    print("qft30 basis-state max-abs: %.3e" % err)
    assert err < 1e-12


def test_layered28_roundtrip():
    """Circuit then its inverse returns |0...0> at full scale (n=28)."""
    n = 28
    ops = workloads.random_layered(n, 20, 3)
    inv = {"s": "sdg", "t": "tdg", "h": "h", "x": "x", "cx": "cx"}
    s = GpuVectorState(n, 1)
    for name, bits, params in ops:
        s.apply_gate(M(name, params), bits)
    for name, bits, params in reversed(ops):
        if name in ("rx", "ry", "rz"):
            s.apply_gate(M(name, [-params[0]]), bits)
        else:
            s.apply_gate(M(inv[name]), bits)
    col = s.column(0)
    e0 = np.zeros(1 << n, complex)
    e0[0] = 1
    err = float(np.abs(col - e0).max())
    print("roundtrip n=28 max-abs: %.3e" % err)
    assert err < 1e-11


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_wide_dense_gates(precision):
    """Gates on 11 and 12 qubits known only by their matrix (the widest the
    backend takes; a group of 2^12 complex128 amplitudes fills 64 KiB of LDS),
    plain and with a control found from the matrix, against the dense
    simulator."""
    rng = np.random.Generator(np.random.PCG64(11))
    n = 14
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    g = GpuVectorState.from_amplitudes(a, 1, precision)
    ref = a.copy()
    for k in (11, 12):
        q, _ = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
        bits = [int(b) for b in rng.permutation(n)[:k]]
        g.apply_gate(q, bits)
        ref = dense_sim.apply(ref, q, bits, n)
    q, _ = np.linalg.qr(rng.normal(size=(1 << 11, 1 << 11)) + 1j * rng.normal(size=(1 << 11, 1 << 11)))
    cq = np.eye(1 << 12, dtype=complex)
    cq[1 << 11:, 1 << 11:] = q
    cbits = [int(b) for b in rng.permutation(n)[:12]]
    g.apply_gate(cq, cbits)
    ref = dense_sim.apply(ref, cq, cbits, n)
    close(g.column(0), ref, 1e-12 if precision == "c128" else 2e-5)
    # rejected before the matrix is read (a stand-in of the right row count)
    with pytest.raises(Exception, match="more than 12 qubits"):
        g.apply_gate(np.zeros((1 << 13, 1), dtype=complex), list(range(13)))
