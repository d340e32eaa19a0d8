"""CPU FP64 oracle bindings -- TEST INFRASTRUCTURE ONLY.

ctypes access to ``liboracle_q1t.so`` (the C restatement of q1tsim 0.5.0's
``VectorState`` in ``vectorstate_ref.c``) plus ``RefCircuit``, a restatement of
the reference's op interpreter ``Circuit::do_execute_with``
(src/circuit.rs:643-762) and its histogram outputs (src/circuit.rs:773-841)
driving that oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_q1t.so")

# orc_kind (vectorstate_ref.h)
I, H, X, Y, Z, S, SDG, T, TDG, RX, RY, RZ, U1, U2, U3, V, VDG, SWAP, MATRIX = range(19)

# gate name -> (kind, number of control layers, number of parameters); the
# names are those of ffi.rs:335-366 plus composite.rs:289-445.
GATES = {
    "i": (I, 0, 0), "h": (H, 0, 0), "x": (X, 0, 0), "y": (Y, 0, 0), "z": (Z, 0, 0),
    "s": (S, 0, 0), "sdg": (SDG, 0, 0), "t": (T, 0, 0), "tdg": (TDG, 0, 0),
    "rx": (RX, 0, 1), "ry": (RY, 0, 1), "rz": (RZ, 0, 1),
    "u1": (U1, 0, 1), "u2": (U2, 0, 2), "u3": (U3, 0, 3),
    "v": (V, 0, 0), "vdg": (VDG, 0, 0), "swap": (SWAP, 0, 0),
    "cx": (X, 1, 0), "cy": (Y, 1, 0), "cz": (Z, 1, 0), "ch": (H, 1, 0),
    "crx": (RX, 1, 1), "cry": (RY, 1, 1), "crz": (RZ, 1, 1),
    "cs": (S, 1, 0), "csdg": (SDG, 1, 0), "ct": (T, 1, 0), "ctdg": (TDG, 1, 0),
    "cu1": (U1, 1, 1), "cu2": (U2, 1, 2), "cu3": (U3, 1, 3),
    "cv": (V, 1, 0), "cvdg": (VDG, 1, 0),
    "ccx": (X, 2, 0), "ccz": (Z, 2, 0),
    "ccrx": (RX, 2, 1), "ccry": (RY, 2, 1), "ccrz": (RZ, 2, 1),
}


class OrcCplx(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class OrcGate(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("nctrl", ctypes.c_int),
                ("p", ctypes.c_double * 3), ("mat_bits", ctypes.c_size_t),
                ("mat", ctypes.c_void_p)]


class OrcRng(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64 * 4)]


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, dp, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
        sig = {
            "orc_rng_seed": (None, [vp, ctypes.c_uint64]),
            "orc_rng_next": (ctypes.c_uint64, [vp]),
            "orc_rng_uniform": (ctypes.c_double, [vp]),
            "orc_new": (vp, [sz, sz]),
            "orc_from_qubit_coefs": (vp, [dp, sz, sz]),
            "orc_from_amplitudes": (vp, [dp, sz, sz]),
            "orc_free": (None, [vp]),
            "orc_nr_bits": (sz, [vp]),
            "orc_nr_cols": (sz, [vp]),
            "orc_counts": (None, [vp, dp]),
            "orc_column": (None, [vp, sz, dp]),
            "orc_gate_nbits": (sz, [vp]),
            "orc_gate_matrix": (None, [vp, dp]),
            "orc_apply_gate": (i32, [vp, vp, dp, sz]),
            "orc_apply_unary_gate_all": (i32, [vp, vp]),
            "orc_apply_conditional_gate": (i32, [vp, dp, sz, vp, dp, sz]),
            "orc_marginal_w0": (i32, [vp, sz, dp]),
            "orc_measure_into": (i32, [vp, sz, sz, dp, sz, vp, i32]),
            "orc_measure_all_uniforms": (i32, [vp, dp, sz, dp, dp, sz, i32, i32]),
            "orc_measure_all_into": (i32, [vp, dp, sz, dp, sz, vp, i32]),
            "orc_reset": (i32, [vp, sz, vp]),
            "orc_reset_all": (None, [vp]),
            "orc_tree_sample": (sz, [dp, sz, ctypes.c_double, dp]),
            "orc_tree_sample_many": (None, [dp, sz, dp, sz, dp, dp]),
            "orc_weighted_index_seq": (sz, [dp, sz, ctypes.c_double]),
            "orc_reverse_bits": (ctypes.c_uint64, [ctypes.c_uint64, sz]),
            "orc_shuffle_bits": (ctypes.c_uint64, [ctypes.c_uint64, dp, sz]),
            "orc_collect_conditional_ranges": (sz, [dp, sz, dp, dp, dp, dp]),
            "orc_random_normal_state": (None, [dp, sz, ctypes.c_uint64]),
            "orc_canonical_weights": (None, [dp, sz, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sizes(vals):
    return np.ascontiguousarray(np.asarray(vals, dtype=np.uint64).astype(np.uintp))


class Rng:
    """xoshiro256** seeded with SplitMix64 (the stream the product also uses)."""

    def __init__(self, seed):
        self.st = OrcRng()
        lib().orc_rng_seed(ctypes.byref(self.st), ctypes.c_uint64(seed))

    def next_u64(self):
        return lib().orc_rng_next(ctypes.byref(self.st))

    def uniform(self):
        return lib().orc_rng_uniform(ctypes.byref(self.st))

    def uniforms(self, n):
        return np.array([self.uniform() for _ in range(n)], dtype=np.float64)


def make_gate(name, params=(), matrix=None):
    """Build an orc_gate for a gate name (or a raw matrix)."""
    g = OrcGate()
    keep = None
    if matrix is not None:
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
        k = int(np.log2(m.shape[0]))
        g.kind, g.nctrl, g.mat_bits = MATRIX, 0, k
        g.mat = m.ctypes.data
        keep = m
    else:
        kind, nctrl, npar = GATES[name.lower()]
        if len(params) != npar:
            raise ValueError("gate %s takes %d parameters" % (name, npar))
        g.kind, g.nctrl = kind, nctrl
        for j, p in enumerate(params):
            g.p[j] = float(p)
    g._keep = keep
    return g


def gate_matrix(name, params=(), matrix=None):
    g = make_gate(name, params, matrix)
    k = lib().orc_gate_nbits(ctypes.byref(g))
    out = np.zeros((1 << k, 1 << k), dtype=np.complex128)
    lib().orc_gate_matrix(ctypes.byref(g), _ptr(out))
    return out


class RefState:
    """The oracle VectorState (row-major 2^n x ncols, counts per column)."""

    ERRORS = {1: "InvalidNrBits", 2: "InvalidQBit", 3: "InvalidCBit", 4: "InvalidNrControlBits",
              5: "InvalidNrMeasurementBits", 6: "NotEnoughSpace", 7: "InvalidWeights"}

    def __init__(self, nr_bits, nr_shots, coefs=None, amplitudes=None):
        L = lib()
        self.nr_bits, self.nr_shots = nr_bits, nr_shots
        if coefs is not None:
            c = np.ascontiguousarray(np.asarray(coefs, dtype=np.complex128))
            self.h = L.orc_from_qubit_coefs(_ptr(c), nr_bits, nr_shots)
        elif amplitudes is not None:
            a = np.ascontiguousarray(np.asarray(amplitudes, dtype=np.complex128))
            self.h = L.orc_from_amplitudes(_ptr(a), nr_bits, nr_shots)
        else:
            self.h = L.orc_new(nr_bits, nr_shots)

    @classmethod
    def from_qubit_coefs(cls, coefs, nr_shots):
        return cls(len(coefs) // 2, nr_shots, coefs=coefs)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def _check(self, rc):
        if rc:
            raise RuntimeError(self.ERRORS.get(rc, "error %d" % rc))

    @property
    def ncols(self):
        return lib().orc_nr_cols(self.h)

    @property
    def counts(self):
        out = np.zeros(self.ncols, dtype=np.uintp)
        lib().orc_counts(self.h, _ptr(out))
        return [int(c) for c in out]

    @property
    def states(self):
        """The (2^n x ncols) complex128 matrix, as in the reference."""
        N = 1 << self.nr_bits
        m = np.zeros((N, self.ncols), dtype=np.complex128)
        col = np.zeros(N, dtype=np.complex128)
        for c in range(self.ncols):
            lib().orc_column(self.h, c, _ptr(col))
            m[:, c] = col
        return m

    def apply_gate(self, name, bits, params=(), matrix=None):
        g = make_gate(name, params, matrix)
        b = _sizes(bits)
        self._check(lib().orc_apply_gate(self.h, ctypes.byref(g), _ptr(b), len(bits)))

    def apply_unary_gate_all(self, name, params=()):
        g = make_gate(name, params)
        self._check(lib().orc_apply_unary_gate_all(self.h, ctypes.byref(g)))

    def apply_conditional_gate(self, control, name, bits, params=(), matrix=None):
        g = make_gate(name, params, matrix)
        c = np.ascontiguousarray(np.asarray(control, dtype=np.uint8))
        b = _sizes(bits)
        self._check(lib().orc_apply_conditional_gate(self.h, _ptr(c), len(control), ctypes.byref(g),
                                                     _ptr(b), len(bits)))

    def marginal_w0(self, qbit):
        out = np.zeros(self.ncols, dtype=np.float64)
        self._check(lib().orc_marginal_w0(self.h, qbit, _ptr(out)))
        return out

    def measure_into(self, qbit, cbit, res, rng, collapse=True):
        self._check(lib().orc_measure_into(self.h, qbit, cbit, _ptr(res), len(res),
                                           ctypes.byref(rng.st), 1 if collapse else 0))

    def measure_all_uniforms(self, cbits, u01, res, collapse=True, mode=0):
        cb = _sizes(cbits)
        u = np.ascontiguousarray(np.asarray(u01, dtype=np.float64))
        self._check(lib().orc_measure_all_uniforms(self.h, _ptr(cb), len(cbits), _ptr(u), _ptr(res),
                                                   len(res), 1 if collapse else 0, mode))

    def measure_all_into(self, cbits, res, rng, collapse=True):
        cb = _sizes(cbits)
        self._check(lib().orc_measure_all_into(self.h, _ptr(cb), len(cbits), _ptr(res), len(res),
                                               ctypes.byref(rng.st), 1 if collapse else 0))

    def reset(self, bit, rng):
        self._check(lib().orc_reset(self.h, bit, ctypes.byref(rng.st)))

    def reset_all(self):
        lib().orc_reset_all(self.h)


def tree_sample(w, u):
    w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    tot = ctypes.c_double(0.0)
    idx = lib().orc_tree_sample(_ptr(w), len(w), float(u), ctypes.byref(tot))
    return int(idx), tot.value


def tree_sample_many(w, us):
    """Canonical tree walk for many uniforms over one weight vector."""
    w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    us = np.ascontiguousarray(np.asarray(us, dtype=np.float64))
    out = np.zeros(len(us), dtype=np.uintp)
    tot = ctypes.c_double(0.0)
    lib().orc_tree_sample_many(_ptr(w), len(w), _ptr(us), len(us), _ptr(out), ctypes.byref(tot))
    return out.astype(np.uint64), tot.value


def canonical_weights(amps):
    """fma(re, re, im*im) per amplitude, as orc_norm_sqr computes it."""
    a = np.asarray(amps, dtype=np.complex128)
    out = np.zeros(len(a), dtype=np.float64)
    L = lib()
    # vectorised equivalent is not bit-exact (numpy has no fma); use the C loop
    c = np.ascontiguousarray(a)
    L.orc_canonical_weights(_ptr(c), len(c), _ptr(out))
    return out


def weighted_index_seq(w, u):
    w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    return int(lib().orc_weighted_index_seq(_ptr(w), len(w), float(u)))


def reverse_bits(idx, nr_bits):
    return int(lib().orc_reverse_bits(idx, nr_bits))


def shuffle_bits(idx, bits):
    b = _sizes(bits)
    return int(lib().orc_shuffle_bits(idx, _ptr(b), len(bits)))


def collect_conditional_ranges(counts, control):
    cnt = _sizes(counts)
    c = np.ascontiguousarray(np.asarray(control, dtype=np.uint8))
    cap = len(control) + len(counts)
    oc, ol = np.zeros(cap, np.uintp), np.zeros(cap, np.uintp)
    oa = np.zeros(cap, np.uint8)
    nr = lib().orc_collect_conditional_ranges(_ptr(cnt), len(counts), _ptr(c), _ptr(oc), _ptr(ol), _ptr(oa))
    return [(int(oc[k]), int(ol[k]), bool(oa[k])) for k in range(nr)]


def random_normal_state(nr_bits, seed):
    out = np.zeros(1 << nr_bits, dtype=np.complex128)
    lib().orc_random_normal_state(_ptr(out), nr_bits, ctypes.c_uint64(seed))
    return out


class RefCircuit:
    """Restatement of q1tsim's Circuit op loop (circuit.rs:643-762) on the oracle.

    Ops: ("gate", name, bits, params), ("cgate", control, target, name, bits,
    params), ("measure", q, c, basis), ("measure_all", cbits, basis),
    ("peek", q, c, basis), ("peek_all", cbits, basis), ("reset", q),
    ("reset_all",), ("barrier", bits).  Gate parameters are read at execute
    time (callables allowed), like Parameter::value (parameter.rs:49-57).
    """

    def __init__(self, nr_qbits, nr_cbits=0):
        self.nr_qbits, self.nr_cbits = nr_qbits, nr_cbits
        self.ops = []
        self.q_state = None
        self.c_state = None
        self.rng = None
        self.initial_amplitudes = None

    def add_gate(self, name, bits, params=()):
        self.ops.append(("gate", name, list(bits), list(params)))

    def add_conditional_gate(self, control, target, name, bits, params=()):
        self.ops.append(("cgate", list(control), int(target), name, list(bits), list(params)))

    def measure(self, q, c, basis="Z"):
        self.ops.append(("measure", q, c, basis.upper()))

    def measure_all(self, cbits, basis="Z"):
        self.ops.append(("measure_all", list(cbits), basis.upper()))

    def peek(self, q, c, basis="Z"):
        self.ops.append(("peek", q, c, basis.upper()))

    def peek_all(self, cbits, basis="Z"):
        self.ops.append(("peek_all", list(cbits), basis.upper()))

    def reset(self, q):
        self.ops.append(("reset", q))

    def reset_all(self):
        self.ops.append(("reset_all",))

    def execute(self, nr_shots, seed):
        if self.initial_amplitudes is not None:
            self.q_state = RefState(self.nr_qbits, nr_shots, amplitudes=self.initial_amplitudes)
        else:
            self.q_state = RefState(self.nr_qbits, nr_shots)
        self.c_state = np.zeros(nr_shots, dtype=np.uint64)
        self.rng = Rng(seed)
        self.reexecute()

    @staticmethod
    def _vals(params):
        return [p() if callable(p) else float(p) for p in params]

    def _basis_pre(self, basis, bits):
        st = self.q_state
        for b in bits:
            if basis == "X":
                st.apply_gate("h", [b])
            elif basis == "Y":
                st.apply_gate("sdg", [b])
                st.apply_gate("h", [b])

    def _basis_post(self, basis, bits):
        st = self.q_state
        for b in bits:
            if basis == "X":
                st.apply_gate("h", [b])
            elif basis == "Y":
                st.apply_gate("h", [b])
                st.apply_gate("s", [b])

    def _all_pre(self, basis):
        st = self.q_state
        if basis == "X":
            st.apply_unary_gate_all("h")
        elif basis == "Y":
            st.apply_unary_gate_all("sdg")
            st.apply_unary_gate_all("h")

    def _all_post(self, basis):
        st = self.q_state
        if basis == "X":
            st.apply_unary_gate_all("h")
        elif basis == "Y":
            st.apply_unary_gate_all("h")
            st.apply_unary_gate_all("s")

    def reexecute(self):
        st, cs, rng = self.q_state, self.c_state, self.rng
        if st is None:
            raise RuntimeError("NotExecuted")
        for op in self.ops:
            kind = op[0]
            if kind == "gate":
                st.apply_gate(op[1], op[2], self._vals(op[3]))
            elif kind == "cgate":
                control, target, name, bits, params = op[1:]
                words = np.zeros(len(cs), dtype=np.uint64)
                for idst, isrc in enumerate(control):
                    words |= ((cs >> np.uint64(isrc)) & np.uint64(1)) << np.uint64(idst)
                st.apply_conditional_gate(words == np.uint64(target), name, bits, self._vals(params))
            elif kind in ("measure", "peek"):
                q, c, basis = op[1:]
                self._basis_pre(basis, [q])
                st.measure_into(q, c, cs, rng, collapse=(kind == "measure"))
                self._basis_post(basis, [q])
            elif kind in ("measure_all", "peek_all"):
                cbits, basis = op[1:]
                self._all_pre(basis)
                st.measure_all_into(cbits, cs, rng, collapse=(kind == "measure_all"))
                self._all_post(basis)
            elif kind == "reset":
                st.reset(op[1], rng)
            elif kind == "reset_all":
                st.reset_all()

    def cstate(self):
        return self.c_state

    def histogram(self):
        """histogram() (circuit.rs:773-789): u64 key -> count."""
        keys, counts = np.unique(self.c_state, return_counts=True)
        return {int(k): int(c) for k, c in zip(keys, counts)}

    def histogram_vec(self):
        res = [0] * (1 << self.nr_cbits)
        for k in self.c_state:
            res[int(k)] += 1
        return res

    def histogram_string(self):
        """histogram_string() (circuit.rs:823-841): binary keys, cbit 0 last."""
        return {format(k, "0%db" % self.nr_cbits): c for k, c in self.histogram().items()}
