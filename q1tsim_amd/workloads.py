"""Synthetic circuit recipes used by the benchmark and the parity tests.

Gate draws come from numpy's PCG64 generator seeded with the given seed, so
every consumer (GPU path, CPU oracle, golden fixtures) sees the same gates.
A circuit is a list of ``(name, qbits, params)`` with q1tsim's qubit
numbering (qubit 0 = most significant index bit).
"""
import math

import numpy as np

ONE_QUBIT = ("h", "x", "s", "t", "rx", "ry", "rz")
ROTATIONS = ("rx", "ry", "rz")


def random_layered(nr_qbits, depth, seed):
    """Layered random circuit: per layer one gate per qubit drawn uniformly
    from {H, X, S, T, RX, RY, RZ} (theta ~ U[0, 2pi)), then CX on nearest
    neighbours (0,1),(2,3),... on even layers and (1,2),(3,4),... on odd ones.
    n=18, depth 20 gives 360 one-qubit gates + 170 CX; n=30 gives 600 + 290."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = []
    for layer in range(depth):
        for q in range(nr_qbits):
            g = ONE_QUBIT[int(rng.integers(len(ONE_QUBIT)))]
            if g in ROTATIONS:
                ops.append((g, [q], [float(rng.uniform(0.0, 2.0 * math.pi))]))
            else:
                ops.append((g, [q], []))
        for q in range(layer % 2, nr_qbits - 1, 2):
            ops.append(("cx", [q, q + 1], []))
    return ops


def random_pairs(nr_qbits, depth, seed, two_qubit_frac=0.25):
    """Depth-`depth` circuit whose slots are two-qubit CX on random (not only
    neighbouring) pairs with probability `two_qubit_frac`, else a random
    one-qubit gate from the layered recipe's set."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = []
    for _ in range(depth):
        free = list(range(nr_qbits))
        rng.shuffle(free)
        while free:
            q = free.pop()
            if free and rng.uniform() < two_qubit_frac:
                r = free.pop()
                ops.append(("cx", [q, r], []))
            else:
                g = ONE_QUBIT[int(rng.integers(len(ONE_QUBIT)))]
                if g in ROTATIONS:
                    ops.append((g, [q], [float(rng.uniform(0.0, 2.0 * math.pi))]))
                else:
                    ops.append((g, [q], []))
    return ops


def qft(nr_qbits):
    """Quantum Fourier transform: H on qubit j, then CU1(2pi/2^(k-j+1))
    controlled by qubit k > j, then the swap network (j, n-1-j).  With
    qubit 0 the most significant bit this maps a -> sqrt(N) * ifft(a)."""
    ops = []
    for j in range(nr_qbits):
        ops.append(("h", [j], []))
        for k in range(j + 1, nr_qbits):
            ops.append(("cu1", [k, j], [2.0 * math.pi / (1 << (k - j + 1))]))
    for j in range(nr_qbits // 2):
        ops.append(("swap", [j, nr_qbits - 1 - j], []))
    return ops


_SELF_INVERSE = ("h", "x", "y", "z", "cx", "cy", "cz", "ch", "swap", "ccx", "ccz", "i")
_ADJOINT = {"s": "sdg", "sdg": "s", "t": "tdg", "tdg": "t", "cs": "csdg", "csdg": "cs", "ct": "ctdg",
            "ctdg": "ct", "v": "vdg", "vdg": "v", "cv": "cvdg", "cvdg": "cv"}
_NEGATED = ("rx", "ry", "rz", "u1", "crx", "cry", "crz", "cu1", "ccrx", "ccry", "ccrz")


def inverse(ops):
    """The inverse circuit: the gates in reverse order, each replaced by its
    adjoint (rotation angles negated; U3(t, p, l)^-1 = U3(-t, -l, -p))."""
    out = []
    for name, bits, params in reversed(ops):
        if name in _SELF_INVERSE:
            out.append((name, list(bits), []))
        elif name in _ADJOINT:
            out.append((_ADJOINT[name], list(bits), []))
        elif name in _NEGATED:
            out.append((name, list(bits), [-float(params[0])]))
        elif name in ("u3", "cu3"):
            t, p, l = params
            out.append((name, list(bits), [-float(t), -float(l), -float(p)]))
        else:
            raise ValueError("no inverse for gate %r" % (name,))
    return out


def branching(nr_qbits, depth, nr_meas, seed):
    """Layered random circuit (the recipe of `random_layered`) with `nr_meas`
    mid-circuit Z measurements at distinct random layer boundaries.
    Measurement m reads a random qubit into cbit nr_qbits + m; a classically
    conditioned X or RZ(theta) on a random qubit follows it (control
    [nr_qbits + m], target 1).  The circuit ends with measure_all into cbits
    0..nr_qbits-1.  Returns (nr_cbits, program) with program entries
    ("gate", name, qbits, params), ("measure", q, cbit, "Z"),
    ("cgate", control, target, name, qbits, params), ("measure_all", cbits, "Z")."""
    rng = np.random.Generator(np.random.PCG64(seed))
    layers = []
    for layer in range(depth):
        ops = []
        for q in range(nr_qbits):
            g = ONE_QUBIT[int(rng.integers(len(ONE_QUBIT)))]
            if g in ROTATIONS:
                ops.append(("gate", g, [q], [float(rng.uniform(0.0, 2.0 * math.pi))]))
            else:
                ops.append(("gate", g, [q], []))
        for q in range(layer % 2, nr_qbits - 1, 2):
            ops.append(("gate", "cx", [q, q + 1], []))
        layers.append(ops)
    nr_meas = min(nr_meas, max(depth - 1, 0))
    cuts = sorted(int(x) for x in rng.choice(np.arange(1, depth), size=nr_meas, replace=False)) if nr_meas else []
    prog = []
    m = 0
    for layer, ops in enumerate(layers):
        while m < len(cuts) and cuts[m] == layer:
            q = int(rng.integers(nr_qbits))
            prog.append(("measure", q, nr_qbits + m, "Z"))
            t = int(rng.integers(nr_qbits))
            if rng.uniform() < 0.5:
                prog.append(("cgate", [nr_qbits + m], 1, "x", [t], []))
            else:
                prog.append(("cgate", [nr_qbits + m], 1, "rz", [t], [float(rng.uniform(0.0, 2.0 * math.pi))]))
            m += 1
        prog.extend(ops)
    prog.append(("measure_all", list(range(nr_qbits)), "Z"))
    return nr_qbits + nr_meas, prog


def batch_workload(count, seed, qubits=(10, 22), depths=(10, 40), meas=(1, 3)):
    """`count` branching circuits with qubit counts, depths and measurement
    counts uniform in the given inclusive ranges.  Returns a list of
    (nr_qbits, nr_cbits, program, circuit_seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(rng.integers(qubits[0], qubits[1] + 1))
        d = int(rng.integers(depths[0], depths[1] + 1))
        k = int(rng.integers(meas[0], meas[1] + 1))
        cseed = int(rng.integers(1 << 62))
        nc, prog = branching(n, d, k, cseed)
        out.append((n, nc, prog, cseed))
    return out


def build_program(circ, program):
    """Add a program (as `branching` returns it, plus "peek", "peek_all",
    "reset" and "reset_all" entries) to a circuit object with q1tsim's
    add_gate / add_conditional_gate / measure / peek / reset API."""
    for op in program:
        kind = op[0]
        if kind == "gate":
            circ.add_gate(op[1], op[2], op[3] if op[3] else None)
        elif kind == "cgate":
            circ.add_conditional_gate(op[1], op[2], op[3], op[4], op[5] if op[5] else None)
        elif kind == "measure":
            circ.measure_basis(op[1], op[2], op[3])
        elif kind == "measure_all":
            circ.measure_all_basis(op[1], op[2])
        elif kind == "peek":
            circ.peek_basis(op[1], op[2], op[3])
        elif kind == "peek_all":
            circ.peek_all_basis(op[1], op[2])
        elif kind == "reset":
            circ.reset(op[1])
        elif kind == "reset_all":
            circ.reset_all()
        else:
            raise ValueError("unknown program op %r" % (kind,))
    return circ


def unfused_bytes(ops, state_bytes):
    """Minimal traffic of each gate run on its own: 2S for dense one-qubit
    gates, RZ and X; S for S, T, CX and Swap; S/2 for CU1 (S = state bytes)."""
    total = 0.0
    for name, _, _ in ops:
        if name in ("s", "t", "sdg", "tdg", "z", "cx", "swap", "cz"):
            total += state_bytes
        elif name in ("cu1", "cs", "ct", "csdg", "ctdg"):
            total += 0.5 * state_bytes
        else:
            total += 2.0 * state_bytes
    return total
