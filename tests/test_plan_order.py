"""The GPU planner's gate order, checked on the CPU against the oracle.

The plain plan may apply a gate before earlier ones it commutes with (shared
bits that are diagonal in both), reorders gates into sweeps and register
phases, and applies SWAP / X / CX of local bits in a phase's write-back.
`Circuit.plan_order` reports the order the kernels apply the gates in
(computed on the host by the same planner code); applying the gates in that
order with the oracle must give the state the original order gives."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from oracle import oracle as orc  # noqa: E402
from q1tsim_amd import workloads  # noqa: E402
from q1tsim_amd.q1tsim import Circuit  # noqa: E402

NAMES = ["h", "x", "y", "rx", "ry", "rz", "s", "t", "u3", "v", "z", "cx", "cz", "ch", "crx", "crz", "cu1", "ccx",
         "ccz", "cy", "cu3", "ccrz", "swap", "i"]


def mixed(n, count, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = []
    for _ in range(count):
        name = NAMES[int(rng.integers(len(NAMES)))]
        kind, nctrl, npar = orc.GATES[name]
        k = (2 if kind == orc.SWAP else 1) + nctrl
        bits = [int(b) for b in rng.permutation(n)[:k]]
        ops.append((name, bits, [float(v) for v in rng.uniform(0, 2 * np.pi, npar)]))
    return ops


def swap_chains(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = [("swap", [j, n - 1 - j], []) for j in range(n // 2)]
    for _ in range(30):
        a, b, c = (int(x) for x in rng.permutation(n)[:3])
        ops += [("swap", [a, b], []), ("cx", [b, c], []), ("x", [a], []), ("swap", [b, c], []), ("t", [c], [])]
    return ops


def run_oracle(n, ops, amps):
    r = orc.RefState(n, 1, amplitudes=amps)
    for name, bits, params in ops:
        r.apply_gate(name, bits, params)
    return r.states[:, 0]


CASES = {
    "mixed": lambda n: mixed(n, 400, n),
    "layered": lambda n: workloads.random_layered(n, 12, n),
    "pairs": lambda n: workloads.random_pairs(n, 16, n),
    "qft": lambda n: workloads.qft(n),
    "swaps": lambda n: mixed(n, 60, 7 * n) + swap_chains(n, n) + mixed(n, 60, 9 * n),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("prec,search", [("c128", True), ("c128", False), ("c64", True)])
def test_plan_order_keeps_the_state(case, prec, search):
    n = 16
    ops = CASES[case](n)
    c = Circuit(n, 0)
    for name, bits, params in ops:
        c.add_gate(name, bits, params if params else None)
    order = c.plan_order(prec, search)
    live = [i for i, (name, _, _) in enumerate(ops) if name != "i"]
    assert sorted(order) == live          # every gate that runs, exactly once
    rng = np.random.Generator(np.random.PCG64(3))
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    ref = run_oracle(n, ops, a)
    got = run_oracle(n, [ops[i] for i in order], a)
    assert float(np.abs(got - ref).max()) <= 1e-12


def test_plan_order_reorders():
    """The planner does move gates (else the check above pins nothing)."""
    n = 16
    ops = workloads.random_layered(n, 12, 5)
    c = Circuit(n, 0)
    for name, bits, params in ops:
        c.add_gate(name, bits, params if params else None)
    order = c.plan_order("c128", True)
    assert order != sorted(order)


def test_plan_order_catches_a_wrong_order():
    """A planted swap of two non-commuting gates is caught by the oracle."""
    n = 16
    ops = [("h", [0], []), ("cx", [0, 1], []), ("h", [1], [])]
    a = np.zeros(1 << n, complex)
    a[0] = 1.0
    ref = run_oracle(n, ops, a)
    bad = run_oracle(n, [ops[0], ops[2], ops[1]], a)
    assert float(np.abs(bad - ref).max()) > 1e-3


def test_plan_order_rejects_non_gate_ops():
    c = Circuit(16, 1)
    c.add_gate("h", [0])
    c.measure(0, 0)
    with pytest.raises(Exception):
        c.plan_order("c128", True)
