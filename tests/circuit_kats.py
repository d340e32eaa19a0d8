"""Circuit-level known-answer tests of the reference, transcribed as data.

Each entry restates one circuit of the reference's own circuit tests
(src/circuit.rs:1456-1985) as an op list plus the check the reference makes
on its histogram or classical state.  The same entries run on the oracle's
RefCircuit (CPU, tests/test_circuit_kats.py) and on the C-ABI Circuit over
the GPU statevector (tests/test_gpu_circuit_kats.py).

Op format: ("gate", name, bits, params), ("measure"|"peek", q, c, basis),
("measure_all"|"peek_all", cbits, basis), ("reset", q), ("reset_all",).

Checks get (hist_vec, hist, cstate, nr_shots): hist_vec is histogram_vec()
(list indexed by the u64 key), hist the u64-keyed histogram() dict.
"""
from tests import stats_helper

TOL = 1.0e-5


def ok(count, n, p):
    return stats_helper.measurement_ok(count, n, p, TOL)


def g(name, *bits, params=()):
    return ("gate", name, list(bits), list(params))


def _all(vals, n, p):
    return all(ok(c, n, p) for c in vals)


def _nibbles(hist, width, shift):
    out = [0] * (1 << width)
    for key, count in hist.items():
        out[(key >> shift) & ((1 << width) - 1)] += count
    return out


KATS = []


def kat(name, nq, nc, shots, prog):
    def deco(check):
        KATS.append((name, nq, nc, shots, prog, check))
        return check
    return deco


# --- test_execute, circuit.rs:1456-1468 -----------------------------------
@kat("execute", 2, 2, 5, [g("x", 0), g("x", 1), g("cx", 0, 1), ("measure", 0, 0, "Z"), ("measure", 1, 1, "Z")])
def _(hv, h, cs, n):
    return list(cs) == [0b01] * 5


# --- test_measure, circuit.rs:1471-1517 -----------------------------------
@kat("measure_z", 2, 2, 1024, [g("x", 0), ("measure", 0, 0, "Z"), ("measure", 1, 1, "Z")])
def _(hv, h, cs, n):
    return hv == [0, n, 0, 0]


@kat("measure_x_uniform", 2, 2, 1024, [g("x", 0), ("measure", 0, 0, "X"), ("measure", 1, 1, "X")])
def _(hv, h, cs, n):
    return _all(hv, n, 0.25)


@kat("measure_x_determined", 2, 2, 1024,
     [g("x", 0), g("h", 0), g("h", 1), ("measure", 0, 0, "X"), ("measure", 1, 1, "X")])
def _(hv, h, cs, n):
    return hv == [0, n, 0, 0]


@kat("measure_y_uniform", 2, 2, 1024, [g("x", 0), ("measure", 0, 0, "Y"), ("measure", 1, 1, "Y")])
def _(hv, h, cs, n):
    return _all(hv, n, 0.25)


# --- test_peek, circuit.rs:1520-1574 --------------------------------------
@kat("peek_1q", 1, 3, 1024, [g("h", 0), ("peek", 0, 0, "Z"), g("h", 0), ("peek", 0, 1, "Z"), g("h", 0),
                             ("peek", 0, 2, "Z")])
def _(hv, h, cs, n):
    return (ok(hv[0] + hv[2] + hv[4] + hv[6], n, 0.5) and hv[0] + hv[1] + hv[4] + hv[5] == n
            and ok(hv[0] + hv[1] + hv[2] + hv[3], n, 0.5))


@kat("peek_2q", 2, 6, 1024, [g("h", 0), g("h", 1), ("peek", 0, 0, "Z"), g("h", 0), ("peek", 0, 1, "Z"),
                             g("h", 0), ("peek", 0, 2, "Z")])
def _(hv, h, cs, n):
    n0, n1, n2 = _nibbles(h, 1, 0), _nibbles(h, 1, 1), _nibbles(h, 1, 2)
    return _all(n0, n, 0.5) and n1 == [n, 0] and _all(n2, n, 0.5)


# --- test_peek_basis, circuit.rs:1576-1617 --------------------------------
@kat("peek_x_1q", 1, 3, 1024, [("peek", 0, 0, "X"), g("h", 0), ("peek", 0, 1, "X"), g("h", 0),
                               ("peek", 0, 2, "X")])
def _(hv, h, cs, n):
    return (ok(hv[0] + hv[2] + hv[4] + hv[6], n, 0.5) and hv[0] + hv[1] + hv[4] + hv[5] == n
            and ok(hv[0] + hv[1] + hv[2] + hv[3], n, 0.5))


@kat("peek_y_2q", 2, 6, 1024, [("peek", 0, 0, "Y"), g("h", 0), ("peek", 0, 1, "Y"), g("sdg", 0),
                               ("peek", 0, 2, "Y")])
def _(hv, h, cs, n):
    return (ok(hv[0] + hv[2] + hv[4] + hv[6], n, 0.5) and ok(hv[0] + hv[1] + hv[4] + hv[5], n, 0.5)
            and hv[0] + hv[1] + hv[2] + hv[3] == 0)


# --- test_conditional, circuit.rs:1619-1630 (first circuit; the injected
# c_state cases run on the oracle only, see test_circuit_kats.py) ----------
@kat("conditional_none", 2, 2, 5, [("cgate", [0, 1], 1, "x", [1], []), ("measure_all", [0, 1], "Z")])
def _(hv, h, cs, n):
    return list(cs) == [0] * 5


# --- test_measure_all, circuit.rs:1654-1686 -------------------------------
@kat("measure_all_01", 2, 2, 1024, [g("x", 0), ("measure_all", [0, 1], "Z")])
def _(hv, h, cs, n):
    return hv == [0, n, 0, 0]


@kat("measure_all_10", 2, 2, 1024, [g("x", 0), ("measure_all", [1, 0], "Z")])
def _(hv, h, cs, n):
    return hv == [0, 0, n, 0]


@kat("measure_all_uniform", 2, 2, 1024, [g("h", 0), g("h", 1), ("measure_all", [0, 1], "Z")])
def _(hv, h, cs, n):
    return _all(hv, n, 0.25)


# --- test_measure_all_basis, circuit.rs:1688-1733 -------------------------
@kat("measure_all_x_00", 2, 2, 1024, [g("h", 0), g("h", 1), ("measure_all", [0, 1], "X")])
def _(hv, h, cs, n):
    return hv == [n, 0, 0, 0]


@kat("measure_all_x_01", 2, 2, 1024, [g("x", 0), g("h", 0), g("h", 1), ("measure_all", [0, 1], "X")])
def _(hv, h, cs, n):
    return hv == [0, n, 0, 0]


@kat("measure_all_y_01", 2, 2, 1024, [g("x", 0), g("h", 0), g("h", 1), g("s", 0), g("s", 1),
                                      ("measure_all", [0, 1], "Y")])
def _(hv, h, cs, n):
    return hv == [0, n, 0, 0]


@kat("measure_all_y_uniform", 2, 2, 1024, [("measure_all", [0, 1], "Y")])
def _(hv, h, cs, n):
    return _all(hv, n, 0.25)


# --- test_peek_all, circuit.rs:1736-1792 ----------------------------------
@kat("peek_all_1q", 1, 3, 1024, [g("h", 0), ("peek_all", [0], "Z"), g("h", 0), ("peek_all", [1], "Z"),
                                 g("h", 0), ("peek_all", [2], "Z")])
def _(hv, h, cs, n):
    return (ok(hv[0] + hv[2] + hv[4] + hv[6], n, 0.5) and hv[0] + hv[1] + hv[4] + hv[5] == n
            and ok(hv[0] + hv[1] + hv[2] + hv[3], n, 0.5))


@kat("peek_all_2q", 2, 6, 1024, [g("h", 0), g("h", 1), ("peek_all", [0, 1], "Z"), g("h", 0),
                                 ("peek_all", [2, 3], "Z"), g("h", 0), ("peek_all", [4, 5], "Z")])
def _(hv, h, cs, n):
    n0, n1, n2 = _nibbles(h, 2, 0), _nibbles(h, 2, 2), _nibbles(h, 2, 4)
    return (_all(n0, n, 0.25) and n1[1] == 0 and n1[3] == 0 and ok(n1[0], n, 0.5) and ok(n1[2], n, 0.5)
            and _all(n2, n, 0.25))


# --- test_peek_all_basis, circuit.rs:1794-1850 ----------------------------
@kat("peek_all_x_1q", 1, 3, 1024, [("peek_all", [0], "X"), g("h", 0), ("peek_all", [1], "X"), g("h", 0),
                                   ("peek_all", [2], "X")])
def _(hv, h, cs, n):
    return (ok(hv[0] + hv[2] + hv[4] + hv[6], n, 0.5) and hv[0] + hv[1] + hv[4] + hv[5] == n
            and ok(hv[0] + hv[1] + hv[2] + hv[3], n, 0.5))


@kat("peek_all_y_2q", 2, 6, 1024, [g("h", 0), g("h", 1), ("peek_all", [0, 1], "Y"), g("s", 1),
                                   ("peek_all", [2, 3], "Y"), g("s", 0), ("peek_all", [4, 5], "Y")])
def _(hv, h, cs, n):
    n0, n1, n2 = _nibbles(h, 2, 0), _nibbles(h, 2, 2), _nibbles(h, 2, 4)
    return (_all(n0, n, 0.25) and n1[2] == 0 and n1[3] == 0 and ok(n1[0], n, 0.5) and ok(n1[1], n, 0.5)
            and n2 == [n, 0, 0, 0])


# --- test_histogram / _vec / _string, circuit.rs:1851-1921 ----------------
HIST_PROG = [g("h", 0), g("h", 1), ("measure", 0, 0, "Z"), ("measure", 1, 1, "Z")]


@kat("histogram", 2, 2, 4096, HIST_PROG)
def _(hv, h, cs, n):
    return sorted(h) == [0, 1, 2, 3] and sum(h.values()) == n and _all(h.values(), n, 0.25)


@kat("histogram_vec", 2, 2, 4096, HIST_PROG)
def _(hv, h, cs, n):
    return sum(hv) == n and _all(hv, n, 0.25)


# --- test_reset, circuit.rs:1923-1969 -------------------------------------
@kat("reset_00", 2, 2, 1024, [g("h", 0), g("z", 0), ("reset", 0), ("measure", 0, 0, "Z"), ("measure", 1, 1, "Z")])
def _(hv, h, cs, n):
    return hv == [n, 0, 0, 0]


@kat("reset_10", 2, 2, 1024, [g("h", 0), g("z", 0), g("x", 1), ("reset", 0), ("measure", 0, 0, "Z"),
                              ("measure", 1, 1, "Z")])
def _(hv, h, cs, n):
    return hv == [0, 0, n, 0]


@kat("reset_half", 2, 2, 1024, [g("h", 0), g("z", 0), g("h", 1), ("reset", 0), ("measure", 0, 0, "Z"),
                                ("measure", 1, 1, "Z")])
def _(hv, h, cs, n):
    return ok(hv[0], n, 0.5) and hv[1] == 0 and ok(hv[2], n, 0.5) and hv[3] == 0


# --- test_reset_all, circuit.rs:1971-1985 ---------------------------------
@kat("reset_all", 5, 5, 1024, [g("h", 0), g("z", 0), g("x", 4), g("h", 3), ("reset_all",),
                               ("measure_all", [0, 1, 2, 3, 4], "Z")])
def _(hv, h, cs, n):
    return hv[0] == n and all(c == 0 for c in hv[1:])


# --- Invariants fixed by the mathematics (no reference test) --------------
# A projective measurement repeated in the same basis repeats its outcome,
# which holds only if the basis change after the first measurement restores
# the measured eigenstate exactly (H for X; H then S for Y).
@kat("repeat_measure_x", 2, 2, 1024, [g("ry", 0, params=[0.7]), g("rz", 0, params=[1.9]),
                                      ("measure", 0, 0, "X"), ("measure", 0, 1, "X")])
def _(hv, h, cs, n):
    return hv[1] == 0 and hv[2] == 0 and hv[0] > 0 and hv[3] > 0


@kat("repeat_measure_y", 2, 2, 1024, [g("ry", 0, params=[0.7]), g("rz", 0, params=[1.9]),
                                      ("measure", 0, 0, "Y"), ("measure", 0, 1, "Y")])
def _(hv, h, cs, n):
    return hv[1] == 0 and hv[2] == 0 and hv[0] > 0 and hv[3] > 0


@kat("repeat_measure_all_y", 2, 4, 1024, [g("ry", 0, params=[0.7]), g("rx", 1, params=[2.2]),
                                          ("measure_all", [0, 1], "Y"), ("measure_all", [2, 3], "Y")])
def _(hv, h, cs, n):
    return all(c == 0 for k, c in enumerate(hv) if (k & 3) != (k >> 2)) and sum(1 for c in hv if c) >= 3


@kat("repeat_measure_all_x", 2, 4, 1024, [g("ry", 0, params=[0.7]), g("rx", 1, params=[2.2]),
                                          ("measure_all", [0, 1], "X"), ("measure_all", [2, 3], "X")])
def _(hv, h, cs, n):
    return all(c == 0 for k, c in enumerate(hv) if (k & 3) != (k >> 2)) and sum(1 for c in hv if c) >= 3


# A peek in the Y basis leaves the state unchanged: a Z measurement after it
# sees the distribution of the prepared state (|<1|psi>|^2 = sin^2(0.35)).
@kat("peek_y_restores", 1, 2, 4096, [g("ry", 0, params=[0.7]), ("peek", 0, 0, "Y"), ("measure", 0, 1, "Z")])
def _(hv, h, cs, n):
    import math
    return ok(hv[2] + hv[3], n, math.sin(0.35) ** 2)


@kat("peek_all_y_restores", 1, 2, 4096, [g("ry", 0, params=[0.7]), ("peek_all", [0], "Y"),
                                         ("measure", 0, 1, "Z")])
def _(hv, h, cs, n):
    import math
    return ok(hv[2] + hv[3], n, math.sin(0.35) ** 2)


# Y-basis outcome of the +i eigenstate S.H|0> is 0, of the -i eigenstate 1.
@kat("y_eigenstates", 2, 2, 256, [g("h", 0), g("s", 0), g("h", 1), g("sdg", 1), ("measure", 0, 0, "Y"),
                                  ("measure", 1, 1, "Y")])
def _(hv, h, cs, n):
    return hv == [0, 0, n, 0]


def histograms(cstate, nr_cbits):
    hv = [0] * (1 << nr_cbits)
    h = {}
    for v in cstate:
        v = int(v)
        hv[v] += 1
        h[v] = h.get(v, 0) + 1
    return hv, h
