"""Independent dense statevector simulator used to pin the oracle.

Written from the mathematics only: every gate is its closed-form matrix
(the formulas of each gate's ``matrix()`` in the reference, e.g.
src/gates/rx.rs:64-70, src/gates/u3.rs:74-84, src/gates/controlled.rs:60-69),
and a k-qubit gate acts on the state viewed as an n-axis tensor, with axis q
holding qubit q (qubit 0 is the most significant index bit, vectorstate.rs:249
and gates.rs:53-80).  No permutation tables, no per-gate bodies: a plausible
mistake in the oracle's bit order, matrix layout or gate body disagrees with
this.
"""
import numpy as np

S2 = np.sqrt(0.5)


def _base(name, p):
    c = np.cos
    s = np.sin
    e = lambda t: np.exp(1j * t)  # noqa: E731
    if name == "i":
        return np.eye(2)
    if name == "h":
        return np.array([[S2, S2], [S2, -S2]])
    if name == "x":
        return np.array([[0, 1], [1, 0]])
    if name == "y":
        return np.array([[0, -1j], [1j, 0]])
    if name == "z":
        return np.array([[1, 0], [0, -1]])
    if name == "s":
        return np.array([[1, 0], [0, 1j]])
    if name == "sdg":
        return np.array([[1, 0], [0, -1j]])
    if name == "t":
        return np.array([[1, 0], [0, e(np.pi / 4)]])
    if name == "tdg":
        return np.array([[1, 0], [0, e(-np.pi / 4)]])
    if name == "rx":
        h = p[0] / 2
        return np.array([[c(h), -1j * s(h)], [-1j * s(h), c(h)]])
    if name == "ry":
        h = p[0] / 2
        return np.array([[c(h), -s(h)], [s(h), c(h)]])
    if name == "rz":
        h = p[0] / 2
        return np.array([[e(-h), 0], [0, e(h)]])
    if name == "u1":
        return np.array([[1, 0], [0, e(p[0])]])
    if name == "u2":
        phi, lam = p
        return S2 * np.array([[1, -e(lam)], [e(phi), e(phi + lam)]])
    if name == "u3":
        th, phi, lam = p
        return np.array([[c(th / 2), -e(lam) * s(th / 2)],
                         [e(phi) * s(th / 2), e(phi + lam) * c(th / 2)]])
    if name == "v":
        return 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])
    if name == "vdg":
        return 0.5 * np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]])
    if name == "swap":
        return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])
    raise KeyError(name)


def gate_matrix(name, params=()):
    """Closed-form matrix; 'c' / 'cc' prefixes add control layers."""
    name = name.lower()
    nctrl = 0
    base = name
    if name.startswith("cc"):
        base, nctrl = name[2:], 2
    elif name.startswith("c"):
        base, nctrl = name[1:], 1
    m = np.asarray(_base(base, params), dtype=np.complex128)
    for _ in range(nctrl):
        d = m.shape[0]
        big = np.eye(2 * d, dtype=np.complex128)
        big[d:, d:] = m
        m = big
    return m


def apply(state, mat, bits, n):
    """Apply a 2^k x 2^k matrix to qubits `bits` (first listed = local MSB)."""
    k = len(bits)
    t = state.reshape((2,) * n)
    m = np.asarray(mat, dtype=np.complex128).reshape((2,) * (2 * k))
    t = np.tensordot(m, t, axes=(list(range(k, 2 * k)), list(bits)))
    # tensordot puts the k output axes first; move them back to `bits`
    t = np.moveaxis(t, list(range(k)), list(bits))
    return np.ascontiguousarray(t).reshape(-1)


def run(n, gates, state=None):
    """gates: iterable of (name, bits, params)."""
    if state is None:
        state = np.zeros(1 << n, dtype=np.complex128)
        state[0] = 1.0
    for name, bits, params in gates:
        state = apply(state, gate_matrix(name, params), bits, n)
    return state
