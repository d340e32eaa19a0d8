"""The reference's circuit-level tests (src/circuit.rs:1456-1985) on the
oracle's op loop (RefCircuit), with planted mistakes in its X/Y-basis
wrappers that these tests must catch.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import circuit_kats

SEEDS = (1, 2, 3)


def make_ref(cls, nq, nc, prog):
    r = cls(nq, nc)
    for op in prog:
        kind = op[0]
        if kind == "gate":
            r.add_gate(op[1], op[2], op[3])
        elif kind == "cgate":
            r.add_conditional_gate(op[1], op[2], op[3], op[4], op[5])
        elif kind == "measure":
            r.measure(op[1], op[2], op[3])
        elif kind == "peek":
            r.peek(op[1], op[2], op[3])
        elif kind == "measure_all":
            r.measure_all(op[1], op[2])
        elif kind == "peek_all":
            r.peek_all(op[1], op[2])
        elif kind == "reset":
            r.reset(op[1])
        elif kind == "reset_all":
            r.reset_all()
        else:
            raise ValueError(kind)
    return r


def run_kat(cls, entry, seed):
    name, nq, nc, shots, prog, check = entry
    r = make_ref(cls, nq, nc, prog)
    r.execute(shots, seed)
    cs = r.cstate()
    hv, h = circuit_kats.histograms(cs, nc)
    return check(hv, h, list(int(v) for v in cs), shots)


@pytest.mark.parametrize("entry", circuit_kats.KATS, ids=[k[0] for k in circuit_kats.KATS])
def test_reference_circuit_kats_on_oracle(entry):
    for seed in SEEDS:
        assert run_kat(orc.RefCircuit, entry, seed), (entry[0], seed)


def test_conditional_injected_cstate():
    """circuit.rs:1619-1651: reexecute from an injected c_state on |00>, 5 runs."""
    cases = [([0, 1], 1, [1], [0b10, 0b00, 0b00, 0b00, 0b00]),
             ([0, 1], 2, [1], [0b00, 0b10, 0b10, 0b00, 0b00]),
             ([1], 1, [0], [0b00, 0b01, 0b01, 0b01, 0b00])]
    for control, target, bits, expect in cases:
        r = orc.RefCircuit(2, 2)
        r.add_conditional_gate(control, target, "x", bits)
        r.measure_all([0, 1])
        r.q_state = orc.RefState(2, 5)
        r.c_state = np.array([0b01, 0b10, 0b10, 0b11, 0b00], dtype=np.uint64)
        r.rng = orc.Rng(7)
        r.reexecute()
        assert [int(v) for v in r.c_state] == expect


def test_histogram_string_keys():
    """circuit.rs:1897-1921 and the key format of :831 ({:0width$b})."""
    r = make_ref(orc.RefCircuit, 2, 2, circuit_kats.HIST_PROG)
    r.execute(4096, 5)
    hs = r.histogram_string()
    assert sorted(hs) == ["00", "01", "10", "11"]
    assert sum(hs.values()) == 4096
    r0 = orc.RefCircuit(1, 0)
    r0.add_gate("h", [0])
    r0.execute(8, 1)
    assert r0.histogram_string() == {"0": 8}


# ----------------------------------------------------------------- plants
# Each plant is a plausible slip in the X/Y-basis wrappers of the op loop
# (circuit.rs:669-748).  Every one must make at least one of the transcribed
# tests fail, or the tests would not pin those wrappers.

def _pre(st, basis, b, y_gates):
    if basis == "X":
        st.apply_gate("h", [b])
    elif basis == "Y":
        for gname in y_gates:
            st.apply_gate(gname, [b])


def plant(pre_y=("sdg", "h"), post_y=("h", "s"), post_x=("h",), all_pre_y=("sdg", "h"), all_post_y=("h", "s"),
          all_post_x=("h",)):
    class Planted(orc.RefCircuit):
        def _basis_pre(self, basis, bits):
            for b in bits:
                _pre(self.q_state, basis, b, pre_y)

        def _basis_post(self, basis, bits):
            for b in bits:
                if basis == "X":
                    for gname in post_x:
                        self.q_state.apply_gate(gname, [b])
                elif basis == "Y":
                    for gname in post_y:
                        self.q_state.apply_gate(gname, [b])

        def _all_pre(self, basis):
            if basis == "X":
                self.q_state.apply_unary_gate_all("h")
            elif basis == "Y":
                for gname in all_pre_y:
                    self.q_state.apply_unary_gate_all(gname)

        def _all_post(self, basis):
            if basis == "X":
                for gname in all_post_x:
                    self.q_state.apply_unary_gate_all(gname)
            elif basis == "Y":
                for gname in all_post_y:
                    self.q_state.apply_unary_gate_all(gname)
    return Planted


PLANTS = {
    "pre_y_s_for_sdg": dict(pre_y=("s", "h")),
    "pre_y_order": dict(pre_y=("h", "sdg")),
    "post_y_sdg_for_s": dict(post_y=("h", "sdg")),
    "post_y_order": dict(post_y=("s", "h")),
    "post_x_missing": dict(post_x=()),
    "all_pre_y_s_for_sdg": dict(all_pre_y=("s", "h")),
    "all_pre_y_order": dict(all_pre_y=("h", "sdg")),
    "all_post_y_sdg_for_s": dict(all_post_y=("h", "sdg")),
    "all_post_y_order": dict(all_post_y=("s", "h")),
    "all_post_x_missing": dict(all_post_x=()),
}


def test_unplanted_wrappers_pass():
    cls = plant()
    for entry in circuit_kats.KATS:
        assert run_kat(cls, entry, 1), entry[0]


@pytest.mark.parametrize("name", sorted(PLANTS))
def test_planted_basis_mistake_fails(name):
    cls = plant(**PLANTS[name])
    failed = [e[0] for e in circuit_kats.KATS if not run_kat(cls, e, 1)]
    assert failed, "planted %s passes every circuit KAT" % name
