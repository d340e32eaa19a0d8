"""Pins the CPU oracle (oracle/vectorstate_ref.c) before it is trusted.

Checks, by source:
  * the reference's own known-answer tests, transcribed (vectorstate.rs:426-829,
    support.rs:100-121, stats.rs:40-59, gate test_apply blocks);
  * golden fixtures written by the independent dense simulator
    (tests/golden/dense_sim.py) and brute force against it on random states;
  * closed forms: QFT = sqrt(N) * ifft, norm preservation, S.Sdg = T.Tdg = I;
  * the sampling definition: tree walk vs rand 0.7 WeightedIndex restated, and
    a brute-force inverse CDF.
A planted mistake (a transposed gate matrix) must fail the dense comparison.
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
from tests import stats_helper  # noqa: E402

z, o, i_ = 0j, 1 + 0j, 1j
x = math.sqrt(0.5) + 0j
TOL = 1e-15


def close(a, b, tol=TOL):
    a, b = np.asarray(a, dtype=np.complex128), np.asarray(b, dtype=np.complex128)
    assert a.shape == b.shape, (a.shape, b.shape)
    err = np.abs(a - b).max() if a.size else 0.0
    assert err <= tol, err


# ---------------------------------------------------------------- support.rs

def test_reverse_bits():  # support.rs:100-109
    assert orc.reverse_bits(1, 1) == 1
    assert orc.reverse_bits(1, 4) == 8
    assert orc.reverse_bits(10, 4) == 5
    assert orc.reverse_bits(26, 4) == 5
    assert orc.reverse_bits(0xfffffffffffffffa, 4) == 0x5
    assert orc.reverse_bits(0xffffffffffffaaaa, 32) == 0x5555ffff
    assert orc.reverse_bits(0x1ffffffffffffffa, 64) == 0x5ffffffffffffff8


def test_shuffle_bits():  # support.rs:111-118
    assert orc.shuffle_bits(0xc, [0, 3, 1, 2]) == 0x6
    assert orc.shuffle_bits(0xfffffffffffffffa, [8, 9, 10, 11]) == 0xa00
    assert orc.shuffle_bits(0xf555555555555555, [63, 62, 61, 60]) == 0xa000000000000000
    assert orc.shuffle_bits(0x3, [3, 2, 1, 0]) == 0xc


def test_stats_bounds():  # stats.rs:40-59
    assert stats_helper.get_bounds(1024, 0.5, 1.0e-5) == (443, 581)
    assert stats_helper.get_bounds(1234, 0.25, 1.0e-5) == (243, 374)
    assert stats_helper.get_bounds(1234, 0.75, 1.0e-5) == (860, 991)
    assert stats_helper.get_bounds(1234, 0.75, 1.0e-10) == (828, 1023)
    assert stats_helper.get_bounds(1_000_000, 0.43, 1.0e-4) == (428158, 431842)
    assert stats_helper.measurement_ok(512, 1024, 0.5, 1.0e-5)
    assert stats_helper.measurement_ok(444, 1024, 0.5, 1.0e-5)
    assert not stats_helper.measurement_ok(443, 1024, 0.5, 1.0e-5)
    assert stats_helper.measurement_ok(580, 1024, 0.5, 1.0e-5)
    assert not stats_helper.measurement_ok(581, 1024, 0.5, 1.0e-5)


# ------------------------------------------------------------ vectorstate.rs

def test_new():  # vectorstate.rs:426-442
    s = orc.RefState(1, 1024)
    assert s.counts == [1024]
    close(s.states, [[o], [z]])
    s = orc.RefState(3, 1500)
    assert s.counts == [1500]
    close(s.states, [[o], [z], [z], [z], [z], [z], [z], [z]])


def test_from_qubit_coefs():  # vectorstate.rs:445-470
    s = orc.RefState.from_qubit_coefs([o, z, z, o], 1)
    close(s.states, [[z], [o], [z], [z]])
    s = orc.RefState.from_qubit_coefs([z, o, o, z], 13)
    assert s.counts == [13]
    close(s.states, [[z], [z], [o], [z]])
    s = orc.RefState.from_qubit_coefs([o, o, -i_, z], 9)
    xi = x * i_
    close(s.states, [[-xi], [z], [-xi], [z]])


def test_apply_conditional_gate():  # vectorstate.rs:473-509
    s = orc.RefState(2, 5)
    s.apply_conditional_gate([0, 0, 1, 1, 0], "x", [1])
    assert s.counts == [2, 2, 1]
    close(s.states, [[o, z, o], [z, o, z], [z, z, z], [z, z, z]])

    s = orc.RefState(2, 5)
    s.apply_conditional_gate([0, 0, 1, 1, 1], "x", [0])
    assert s.counts == [2, 3]
    close(s.states, [[o, z], [z, z], [z, o], [z, z]])

    s = orc.RefState(2, 5)
    s.apply_conditional_gate([1, 0, 1, 1, 0], "h", [1])
    assert s.counts == [1, 1, 2, 1]
    close(s.states, [[x, o, x, o], [x, z, x, z], [z, z, z, z], [z, z, z, z]])

    s = orc.RefState.from_qubit_coefs([o, z, x, x], 5)
    s.apply_conditional_gate([1, 0, 1, 1, 0], "cx", [1, 0])
    assert s.counts == [1, 1, 2, 1]
    close(s.states, [[x, x, x, x], [z, x, z, x], [z, z, z, z], [x, z, x, z]])

    s = orc.RefState(2, 5)
    s.apply_conditional_gate([1, 1, 1, 0, 0], "h", [0])
    assert s.counts == [3, 2]
    close(s.states, [[x, o], [z, z], [x, z], [z, z]])
    s.apply_conditional_gate([0, 0, 1, 1, 1], "h", [0])
    assert s.counts == [2, 1, 2]
    close(s.states, [[x, o, x], [z, z, z], [x, z, x], [z, z, z]])


def test_apply_unary_gate():  # vectorstate.rs:641-658
    s = orc.RefState(3, 1)
    s.apply_gate("h", [0])
    close(s.states, [[x], [z], [z], [z], [x], [z], [z], [z]])
    s = orc.RefState(3, 1)
    s.apply_gate("h", [1])
    close(s.states, [[x], [z], [x], [z], [z], [z], [z], [z]])
    s = orc.RefState(3, 1)
    s.apply_gate("y", [2])
    close(s.states, [[z], [i_], [z], [z], [z], [z], [z], [z]])


def test_apply_binary_gate():  # vectorstate.rs:661-683
    h = 0.5 * o
    s = orc.RefState(3, 1)
    s.apply_gate("cx", [0, 1])
    close(s.states, [[o], [z], [z], [z], [z], [z], [z], [z]])
    s = orc.RefState.from_qubit_coefs([z, o, o, z, o, z], 1)
    s.apply_gate("cx", [0, 1])
    close(s.states, [[z], [z], [z], [z], [z], [z], [o], [z]])
    s = orc.RefState.from_qubit_coefs([z, o, o, z, o, z], 1)
    s.apply_gate("cx", [0, 2])
    close(s.states, [[z], [z], [z], [z], [z], [o], [z], [z]])
    s = orc.RefState.from_qubit_coefs([z, o, o, z, o, z], 1)
    hh = np.kron(dense_sim.gate_matrix("h"), dense_sim.gate_matrix("h"))
    s.apply_gate(None, [1, 2], matrix=hh)
    close(s.states, [[z], [z], [z], [z], [h], [h], [h], [h]])


def test_apply_n_ary_gate():  # vectorstate.rs:686-707
    hx = 0.5 * x
    s = orc.RefState(3, 1)
    s.apply_gate("ccx", [0, 1, 2])
    close(s.states, [[o], [z], [z], [z], [z], [z], [z], [z]])
    s = orc.RefState.from_qubit_coefs([z, o, z, o, o, z], 1)
    s.apply_gate("ccx", [0, 2, 1])
    close(s.states, [[z], [z], [z], [z], [z], [z], [o], [z]])
    s.apply_gate("ccx", [0, 1, 2])
    close(s.states, [[z], [z], [z], [z], [z], [z], [z], [o]])
    s = orc.RefState.from_qubit_coefs([x, -x, x, -x, x, -x], 1)
    s.apply_gate("ccx", [0, 2, 1])
    close(s.states, [[hx], [-hx], [-hx], [hx], [-hx], [-hx], [hx], [hx]])


def test_measure_structure():  # vectorstate.rs:512-590
    rng = orc.Rng(1234)
    s = orc.RefState(1, 3)
    m = np.zeros(3, np.uint64)
    s.measure_into(0, 0, m, rng)
    assert list(m) == [0, 0, 0]
    close(s.states, [[o], [z]])

    s = orc.RefState.from_qubit_coefs([o, o, o, o], 1024)
    m0 = np.zeros(1024, np.uint64)
    s.measure_into(0, 0, m0, rng)
    sc = 0
    prev = m0[0]
    st = s.states
    for b in m0:
        if b != prev:
            sc += 1
            prev = b
        close(st[:, sc], [x, x, z, z] if b == 0 else [z, z, x, x])
    m0b = np.zeros(1024, np.uint64)
    s.measure_into(0, 0, m0b, rng)
    assert np.array_equal(m0b, m0)
    m1 = np.zeros(1024, np.uint64)
    s.measure_into(1, 0, m1, rng)
    st = s.states
    sc = 0
    pb = (m0[0], m1[0])
    expect = {(0, 0): [o, z, z, z], (0, 1): [z, o, z, z], (1, 0): [z, z, o, z], (1, 1): [z, z, z, o]}
    for j in range(1024):
        b = (int(m0[j]), int(m1[j]))
        if b != pb:
            sc += 1
            pb = b
        close(st[:, sc], expect[b])


def test_peek_into_statistics():  # vectorstate.rs:593-638
    rng = orc.Rng(99)
    n = 1024
    s = orc.RefState.from_qubit_coefs([o, o], n)
    m = np.zeros(n, np.uint64)
    s.measure_into(0, 0, m, rng, collapse=False)
    assert stats_helper.measurement_ok(int(m.sum()), n, 0.5, 1e-5)
    close(s.states, [[x], [x]])
    s = orc.RefState.from_qubit_coefs([x, x, z, o], n)
    m = np.zeros(n, np.uint64)
    s.measure_into(1, 0, m, rng, collapse=False)
    assert int(m.sum()) == n
    close(s.states, [[z], [x], [z], [x]])


def test_measure_all_bits():  # vectorstate.rs:710-735
    rng = orc.Rng(5)
    s = orc.RefState.from_qubit_coefs([z, o, z, o, z, o], 5)
    r = np.zeros(5, np.uint64)
    s.measure_all_into([0, 1, 2], r, rng)
    assert all(int(b) == 0b111 for b in r)
    s = orc.RefState.from_qubit_coefs([z, o, z, o, o, z], 5)
    r = np.zeros(5, np.uint64)
    s.measure_all_into([0, 1, 2], r, rng)
    assert all(int(b) == 0b011 for b in r)
    s = orc.RefState(3, 5)
    s.apply_gate("h", [2])
    r = np.zeros(5, np.uint64)
    s.measure_all_into([0, 1, 2], r, rng)
    assert all((int(b) & 0b011) == 0 for b in r)


def test_peek_all():  # vectorstate.rs:738-771
    h = 0.5 * o
    rng = orc.Rng(17)
    n = 1024
    s = orc.RefState.from_qubit_coefs([z, o, z, o, z, o], n)
    s.apply_gate("h", [0])
    s.apply_gate("h", [2])
    r = np.zeros(n, np.uint64)
    s.measure_all_into([0, 1, 2], r, rng, collapse=False)
    assert s.counts == [n]
    close(s.states, [[z], [z], [h], [-h], [z], [z], [-h], [h]])
    cnt = [sum((int(b) >> k) & 1 for b in r) for k in range(3)]
    assert stats_helper.measurement_ok(cnt[0], n, 0.5, 1e-5)
    assert cnt[1] == n
    assert stats_helper.measurement_ok(cnt[2], n, 0.5, 1e-5)


def test_reset():  # vectorstate.rs:774-811
    rng = orc.Rng(3)
    s = orc.RefState.from_qubit_coefs([o, z], 10)
    s.reset(0, rng)
    close(s.states, [[o], [z]])
    s = orc.RefState.from_qubit_coefs([z, o], 10)
    s.reset(0, rng)
    close(s.states, [[o], [z]])
    s = orc.RefState.from_qubit_coefs([z, o, z, o], 10)
    s.reset(0, rng)
    close(s.states, [[z], [o], [z], [z]])
    s = orc.RefState.from_qubit_coefs([z, o, z, o], 10)
    s.reset(1, rng)
    close(s.states, [[z], [z], [o], [z]])
    s = orc.RefState.from_qubit_coefs([x, -x, o, z], 10)
    s.reset(0, rng)
    if len(s.counts) == 1:
        close(s.states, [[o], [z], [z], [z]])
    else:
        close(s.states, [[o, -o], [z, z], [z, z], [z, z]])
    s = orc.RefState.from_qubit_coefs([x, -x, o, z], 10)
    s.reset(1, rng)
    close(s.states, [[x], [z], [-x], [z]])


def test_reset_all():  # vectorstate.rs:814-829
    s = orc.RefState(5, 100)
    s.apply_gate("h", [2])
    s.apply_gate("x", [0])
    s.apply_gate("h", [4])
    s.reset_all()
    assert s.counts == [100]
    e = np.zeros((32, 1), complex)
    e[0, 0] = 1
    close(s.states, e)


def test_conditional_ranges():  # qustate.rs:100-127
    assert orc.collect_conditional_ranges([5], [0, 0, 1, 1, 0]) == [(0, 2, False), (0, 2, True), (0, 1, False)]
    assert orc.collect_conditional_ranges([2, 3], [1, 1, 1, 0, 0]) == [(0, 2, True), (1, 1, True), (1, 2, False)]


# ------------------------------------------------------ per-gate KATs (gate_test)

def _cols(name, params, cols):
    """gate_test (gates.rs:373-381): apply_slice on each column."""
    cols = np.asarray(cols, dtype=np.complex128)
    k = int(math.log2(cols.shape[0]))
    out = []
    for c in range(cols.shape[1]):
        s = orc.RefState(k, 1, amplitudes=cols[:, c])
        g = int(math.log2(orc.gate_matrix(name, params).shape[0]))
        s.apply_conditional_gate([1], name, list(range(g)), params)
        out.append(s.states[:, 0])
    return np.array(out).T


def test_gate_kats():
    close(_cols("h", [], [[o, z, x, x], [z, o, x, -x]]), [[x, x, o, z], [x, -x, z, o]])  # hadamard.rs
    close(_cols("rx", [math.pi / 2], [[o, z, x, x], [z, o, x, -x]]),
          [[x, -x * i_, 0.5 * (o - i_), 0.5 * (o + i_)], [-x * i_, x, 0.5 * (o - i_), -0.5 * (o + i_)]])
    close(_cols("rz", [math.pi / 2], [[o, z, x, x], [z, o, x, -x]]),
          [[x * (o - i_), z, 0.5 * (o - i_), 0.5 * (o - i_)], [z, x * (o + i_), 0.5 * (o + i_), -0.5 * (o + i_)]])
    res = 0.5 * np.array([[-o - i_, -o], [z, o], [2 * x, (o + i_) * x], [z, -(o + i_) * x]])
    close(_cols("u3", [3 * math.pi / 2, math.pi / 4, math.pi / 2], [[x, x], [z, -x], [x, z], [z, z]]), res)
    h = 0.5
    sw_in = [[o, z, x, x, x, h, z], [z, o, x, -x, z, -h, z], [z, z, z, z, x, h, o], [z, z, z, z, z, -h, z]]
    sw_out = [[o, z, x, x, x, h, z], [z, z, z, z, x, h, o], [z, o, x, -x, z, -h, z], [z, z, z, z, z, -h, z]]
    close(_cols("swap", [], sw_in), sw_out)


# ------------------------------------------- brute force vs dense simulator

def _apply_list(n, ops, amps):
    s = orc.RefState(n, 1, amplitudes=amps)
    for name, bits, params in ops:
        s.apply_gate(name, bits, params)
    return s.states[:, 0]


def _decode(d):
    ops = []
    for name, b, p in zip(d["names"], d["bits"], d["params"]):
        name = str(name)
        bits = [int(v) for v in b if v >= 0]
        npar = orc.GATES[name][2]
        ops.append((name, bits, [float(v) for v in p[:npar]]))
    return ops


def test_oracle_matrices_match_closed_forms():
    rng = np.random.Generator(np.random.PCG64(1))
    for name, (_, _, npar) in orc.GATES.items():
        params = list(rng.uniform(0, 2 * np.pi, size=npar))
        close(orc.gate_matrix(name, params), dense_sim.gate_matrix(name, params), 1e-15)


def test_all_gates_fixture():
    d = np.load(os.path.join(HERE, "golden", "all_gates_n6.npz"))
    out = _apply_list(int(d["n"]), _decode(d), d["input"])
    close(out, d["output"], 1e-13)


@pytest.mark.parametrize("fname", ["layered_n10_d20_s7.npz", "layered_n8_d12_s3.npz"])
def test_layered_fixture(fname):
    d = np.load(os.path.join(HERE, "golden", fname))
    n = int(d["n"])
    ops = workloads.random_layered(n, int(d["depth"]), int(d["seed"]))
    s = orc.RefState(n, 1)
    for name, bits, params in ops:
        s.apply_gate(name, bits, params)
    close(s.states[:, 0], d["amplitudes"], 1e-12)


def test_qft_closed_form():
    d = np.load(os.path.join(HERE, "golden", "qft_n8.npz"))
    out = _apply_list(8, workloads.qft(8), d["input"])
    close(out, d["output"], 1e-13)
    n = 10
    a = orc.random_normal_state(n, 42)
    out = _apply_list(n, workloads.qft(n), a)
    close(out, np.sqrt(1 << n) * np.fft.ifft(a), 1e-12)


def test_random_normal_state_moments():
    a = orc.random_normal_state(14, 7)
    assert abs(np.linalg.norm(a) - 1.0) < 1e-12
    v = a * np.sqrt(2.0 * len(a))       # re, im ~ N(0, 1) before normalisation
    re = v.real
    assert abs(re.mean()) < 0.03 and abs(re.std() - 1.0) < 0.03
    assert abs(np.mean(re ** 4) - 3.0) < 0.15


def test_brute_force_random_multi_column():
    """Gates on a multi-column state (after a conditional split) vs dense sim per column."""
    rng = np.random.Generator(np.random.PCG64(21))
    n = 5
    a = rng.normal(size=32) + 1j * rng.normal(size=32)
    a /= np.linalg.norm(a)
    s = orc.RefState(n, 4, amplitudes=a)
    s.apply_conditional_gate([1, 0, 0, 1], "ry", [2], [0.7])
    cols = s.states.copy()
    ops = [("h", [4], []), ("cx", [3, 0], []), ("ccx", [4, 1, 2], []), ("u3", [1], [0.3, 1.1, -0.4]),
           ("swap", [0, 4], []), ("cu1", [2, 3], [0.9]), ("crx", [0, 3], [2.2])]
    for name, bits, params in ops:
        s.apply_gate(name, bits, params)
    for c in range(cols.shape[1]):
        close(s.states[:, c], dense_sim.run(n, ops, state=cols[:, c].copy()), 1e-13)


def test_planted_transpose_fails():
    """A transposed (non-symmetric) gate matrix fed to the oracle must be caught."""
    n = 3
    rng = np.random.Generator(np.random.PCG64(2))
    a = rng.normal(size=8) + 1j * rng.normal(size=8)
    a /= np.linalg.norm(a)
    m = dense_sim.gate_matrix("u3", [0.4, 1.3, 2.1])
    good = orc.RefState(n, 1, amplitudes=a)
    good.apply_gate(None, [1], matrix=m)
    bad = orc.RefState(n, 1, amplitudes=a)
    bad.apply_gate(None, [1], matrix=m.T)
    ref = dense_sim.apply(a, m, [1], n)
    assert np.abs(good.states[:, 0] - ref).max() < 1e-14
    assert np.abs(bad.states[:, 0] - ref).max() > 1e-3


def test_unitarity_identities():
    rng = np.random.Generator(np.random.PCG64(8))
    n = 4
    a = rng.normal(size=16) + 1j * rng.normal(size=16)
    a /= np.linalg.norm(a)
    ops = [("s", [1], []), ("sdg", [1], []), ("t", [2], []), ("tdg", [2], []),
           ("rx", [3], [0.77]), ("rx", [3], [-0.77]), ("v", [0], []), ("vdg", [0], [])]
    close(_apply_list(n, ops, a), a, 1e-15)
    out = _apply_list(n, workloads.random_layered(n, 5, 4), a)
    assert abs(np.linalg.norm(out) - 1) < 1e-14


# ------------------------------------------------------------------ sampling

def test_tree_sample_matches_weighted_index():
    rng = np.random.Generator(np.random.PCG64(31))
    for n in [1, 3, 4, 5, 9, 13]:
        w = rng.uniform(size=1 << n) ** 3
        w[rng.integers(0, 1 << n, size=max(1, (1 << n) // 5))] = 0.0
        if w.sum() == 0:
            w[0] = 1.0
        cum = np.cumsum(w)
        for u in rng.uniform(size=200):
            idx, tot = orc.tree_sample(w, u)
            j = orc.weighted_index_seq(w, u)
            t = u * cum[-1]
            brute = int(np.searchsorted(cum[:-1], t, side="right"))
            assert w[idx] > 0
            assert idx == j == brute, (n, u, idx, j, brute)


def test_tree_sample_edges():
    w = np.zeros(64)
    w[63] = 2.0
    assert orc.tree_sample(w, 0.999999)[0] == 63
    assert orc.tree_sample(w, 0.0)[0] == 63
    w = np.zeros(64)
    w[0] = 1.0
    assert orc.tree_sample(w, 0.9999999999)[0] == 0
    w = np.ones(17)
    assert orc.tree_sample(w, 0.0)[0] == 0
    assert orc.tree_sample(w, 1.0 - 2.0 ** -52)[0] == 16


def test_measure_all_uniform_modes_agree():
    rng = np.random.Generator(np.random.PCG64(41))
    n = 8
    a = rng.normal(size=256) + 1j * rng.normal(size=256)
    a /= np.linalg.norm(a)
    u = rng.uniform(size=3000)
    res = []
    for mode in (0, 1):
        s = orc.RefState(n, 3000, amplitudes=a)
        r = np.zeros(3000, np.uint64)
        s.measure_all_uniforms(list(range(n)), u, r, collapse=False, mode=mode)
        res.append(r)
    assert np.array_equal(res[0], res[1])
    p = np.abs(a) ** 2
    # bit mapping: qubit q -> cbit q, so key bit q = index bit n-1-q
    keys = [orc.reverse_bits(i, n) for i in range(256)]
    hist = np.bincount(res[0].astype(np.int64), minlength=256)
    ok, stat, dof = stats_helper.chi2_ok([hist[k] for k in keys], p, 3000)
    assert ok, (stat, dof)
