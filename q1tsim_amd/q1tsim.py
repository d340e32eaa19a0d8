"""Python front-end over libq1tsim.so, mirroring q1tsim's python/q1tsim.py.

Same class names, method names, argument meaning and error behaviour as the
reference wrapper (python/q1tsim.py:5-529), bound with ctypes instead of cffi
(cffi is not installed here); every circuit runs on the MI355X statevector.
Build-added methods (backend selection, seeds, amplitude read-out, run
statistics) are marked as such.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import Q1tError, lib, sizes, unpack_result


class RefParam(object):
    """Reference parameter (python/q1tsim.py:5-36): its value is read when the
    circuit executes, so assign() changes the next (re)execution."""

    def __init__(self, value):
        self.__val = ctypes.c_double(value)

    def __float__(self):
        return self.__val.value

    def assign(self, value):
        self.__val.value = value

    def pointer(self):
        return ctypes.pointer(self.__val)


def _params(values):
    arr = (_lib.Parameter * max(1, len(values)))()
    for i, v in enumerate(values):
        if isinstance(v, RefParam):
            arr[i].value = 0.0
            arr[i].value_ptr = v.pointer()
        else:
            arr[i].value = float(v)
            arr[i].value_ptr = None
    return arr


class Circuit(object):
    """A quantum circuit (python/q1tsim.py:38-529) on the GPU statevector."""

    def __init__(self, nr_qbits, nr_cbits=0):
        self.__sim = lib()
        self.__ptr = self.__sim.circuit_new(nr_qbits, nr_cbits)
        self.__keep = []      # RefParam objects must outlive the circuit

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_value, traceback):
        self.close()

    def close(self):
        if self.__ptr is not None:
            self.__sim.circuit_free(self.__ptr)
            self.__ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nr_qbits(self):
        return self.__sim.circuit_nr_qbits(self.__ptr)

    def nr_cbits(self):
        return self.__sim.circuit_nr_cbits(self.__ptr)

    def cstate(self):
        return unpack_result(self.__sim.circuit_cstate(self.__ptr))

    def add_gate(self, name, qbits, params=None):
        cname = bytes(name, "utf-8")
        q = sizes(qbits)
        if params is None:
            return unpack_result(self.__sim.circuit_add_gate(self.__ptr, cname, q, len(qbits), None, 0))
        self.__keep.extend(p for p in params if isinstance(p, RefParam))
        return unpack_result(self.__sim.circuit_add_gate(self.__ptr, cname, q, len(qbits), _params(params),
                                                         len(params)))

    def add_conditional_gate(self, control, target, name, qbits, params=None):
        """The bit at the position of the first index in control is the least
        significant bit of the control word."""
        cname = bytes(name, "utf-8")
        c, q = sizes(control), sizes(qbits)
        if params is None:
            return unpack_result(self.__sim.circuit_add_conditional_gate(
                self.__ptr, c, len(control), target, cname, q, len(qbits), None, 0))
        self.__keep.extend(p for p in params if isinstance(p, RefParam))
        return unpack_result(self.__sim.circuit_add_conditional_gate(
            self.__ptr, c, len(control), target, cname, q, len(qbits), _params(params), len(params)))

    # named gates (python/q1tsim.py:118-306)
    def ch(self, control, target):
        return self.add_gate("CH", [control, target])

    def crx(self, theta, control, target):
        return self.add_gate("CRX", [control, target], [theta])

    def cry(self, theta, control, target):
        return self.add_gate("CRY", [control, target], [theta])

    def crz(self, lmb, control, target):
        return self.add_gate("CRZ", [control, target], [lmb])

    def cx(self, control, target):
        return self.add_gate("CX", [control, target])

    def cy(self, control, target):
        return self.add_gate("CY", [control, target])

    def cz(self, control, target):
        return self.add_gate("CZ", [control, target])

    def h(self, qbit):
        return self.add_gate("H", [qbit])

    def i(self, qbit):
        return self.add_gate("I", [qbit])

    def rx(self, theta, qbit):
        return self.add_gate("RX", [qbit], [theta])

    def ry(self, theta, qbit):
        return self.add_gate("RY", [qbit], [theta])

    def rz(self, lmb, qbit):
        return self.add_gate("RZ", [qbit], [lmb])

    def s(self, qbit):
        return self.add_gate("S", [qbit])

    def sdg(self, qbit):
        return self.add_gate("Sdg", [qbit])

    def swap(self, qbit0, qbit1):
        return self.add_gate("Swap", [qbit0, qbit1])

    def t(self, qbit):
        return self.add_gate("T", [qbit])

    def tdg(self, qbit):
        return self.add_gate("Tdg", [qbit])

    def u1(self, lmb, qbit):
        return self.add_gate("U1", [qbit], [lmb])

    def u2(self, phi, lmb, qbit):
        return self.add_gate("U2", [qbit], [phi, lmb])

    def u3(self, theta, phi, lmb, qbit):
        return self.add_gate("U3", [qbit], [theta, phi, lmb])

    def v(self, qbit):
        return self.add_gate("V", [qbit])

    def vdg(self, qbit):
        return self.add_gate("Vdg", [qbit])

    def x(self, qbit):
        return self.add_gate("X", [qbit])

    def y(self, qbit):
        return self.add_gate("Y", [qbit])

    def z(self, qbit):
        return self.add_gate("Z", [qbit])

    # build-added names (composite.rs:289-445 gate table)
    def cu1(self, lmb, control, target):
        return self.add_gate("CU1", [control, target], [lmb])

    def ccx(self, c0, c1, target):
        return self.add_gate("CCX", [c0, c1, target])

    # measurement (python/q1tsim.py:308-452)
    def __peek_measure_basis(self, qbit, cbit, basis, collapse):
        return unpack_result(self.__sim.circuit_measure(self.__ptr, qbit, cbit, bytes(basis, "utf-8")[:1],
                                                        1 if collapse else 0))

    def __peek_measure_all_basis(self, cbits, basis, collapse):
        return unpack_result(self.__sim.circuit_measure_all(self.__ptr, sizes(cbits), len(cbits),
                                                            bytes(basis, "utf-8")[:1], 1 if collapse else 0))

    def measure_basis(self, qbit, cbit, basis):
        return self.__peek_measure_basis(qbit, cbit, basis, True)

    def measure_x(self, qbit, cbit):
        return self.measure_basis(qbit, cbit, "X")

    def measure_y(self, qbit, cbit):
        return self.measure_basis(qbit, cbit, "Y")

    def measure_z(self, qbit, cbit):
        return self.measure_basis(qbit, cbit, "Z")

    def measure(self, qbit, cbit):
        return self.measure_z(qbit, cbit)

    def measure_all_basis(self, cbits, basis):
        return self.__peek_measure_all_basis(cbits, basis, True)

    def measure_all(self, cbits):
        return self.measure_all_basis(cbits, "Z")

    def peek_basis(self, qbit, cbit, basis):
        return self.__peek_measure_basis(qbit, cbit, basis, False)

    def peek_x(self, qbit, cbit):
        return self.peek_basis(qbit, cbit, "X")

    def peek_y(self, qbit, cbit):
        return self.peek_basis(qbit, cbit, "Y")

    def peek_z(self, qbit, cbit):
        return self.peek_basis(qbit, cbit, "Z")

    def peek(self, qbit, cbit):
        return self.peek_z(qbit, cbit)

    def peek_all_basis(self, cbits, basis):
        return self.__peek_measure_all_basis(cbits, basis, False)

    def peek_all(self, cbits):
        return self.peek_all_basis(cbits, "Z")

    def reset(self, qbit):
        return unpack_result(self.__sim.circuit_reset(self.__ptr, qbit))

    def reset_all(self):
        return unpack_result(self.__sim.circuit_reset_all(self.__ptr))

    def execute(self, nr_shots):
        return unpack_result(self.__sim.circuit_execute(self.__ptr, nr_shots))

    def reexecute(self):
        return unpack_result(self.__sim.circuit_reexecute(self.__ptr))

    def histogram(self):
        """Histogram with binary string keys, cbit 0 as the last character."""
        return unpack_result(self.__sim.circuit_histogram(self.__ptr))

    def latex(self):
        return unpack_result(self.__sim.circuit_latex(self.__ptr))

    def open_qasm(self):
        return unpack_result(self.__sim.circuit_open_qasm(self.__ptr))

    def c_qasm(self):
        return unpack_result(self.__sim.circuit_c_qasm(self.__ptr))

    # ---- build-added ----
    def set_backend(self, precision="c128", fusion_k=0, tiling=True):
        prec = {"c128": 0, "c64": 1}[precision]
        _lib.check(self.__sim.q1t_circuit_set_backend(self.__ptr, prec, int(fusion_k), 1 if tiling else 0))

    def set_graph(self, on=True):
        """Flush graphs of the state: gate flushes of a reexecute replay the
        previous execute's, with reference parameters re-read."""
        _lib.check(self.__sim.q1t_circuit_set_graph(self.__ptr, 1 if on else 0))

    def set_jit(self, on=True, prefetch=False, factors=True):
        """Sweep kernels specialised to the state's sweep programs (modes as
        GpuVectorState.set_jit)."""
        from .gvs import _jit_mode
        _lib.check(self.__sim.q1t_circuit_set_jit(self.__ptr, _jit_mode(on, prefetch, factors)))

    def jit_compile(self, precision="c128", fusion_k=0, out_dir=None, prefetch=False, factors=True):
        """Plan this gate-only circuit on the host and compile the kernels
        specialised to its sweeps (no GPU needed): (kernels, seconds)."""
        from .gvs import _jit_mode
        prec = {"c128": 0, "c64": 1}[precision]
        n, sec = ctypes.c_size_t(0), ctypes.c_double(0.0)
        _lib.check(self.__sim.q1t_circuit_jit_compile(self.__ptr, prec, int(fusion_k),
                                                      _jit_mode(True, prefetch, factors),
                                                      out_dir.encode() if out_dir else None,
                                                      ctypes.byref(n), ctypes.byref(sec)))
        return int(n.value), float(sec.value)

    def graph_stats(self):
        """(graphs captured, replays launched) over the current state."""
        c, r = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _lib.check(self.__sim.q1t_circuit_graph_stats(self.__ptr, ctypes.byref(c), ctypes.byref(r)))
        return int(c.value), int(r.value)

    def seed(self, seed):
        _lib.check(self.__sim.q1t_circuit_seed(self.__ptr, int(seed)))

    def set_initial_amplitudes(self, amps):
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128))
        _lib.check(self.__sim.q1t_circuit_set_initial_amplitudes(
            self.__ptr, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a)))

    def set_initial_random(self, seed):
        _lib.check(self.__sim.q1t_circuit_set_initial_random(self.__ptr, int(seed)))

    def barrier(self, qbits):
        _lib.check(self.__sim.q1t_circuit_barrier(self.__ptr, sizes(qbits), len(qbits)))

    def nr_columns(self):
        n = ctypes.c_size_t(0)
        _lib.check(self.__sim.q1t_circuit_nr_columns(self.__ptr, ctypes.byref(n)))
        return n.value

    def counts(self):
        n = self.nr_columns()
        out = (ctypes.c_size_t * max(1, n))()
        _lib.check(self.__sim.q1t_circuit_counts(self.__ptr, out, n))
        return [int(out[i]) for i in range(n)]

    def amplitudes(self, col=0):
        n = self.nr_qbits()
        out = np.zeros(1 << n, dtype=np.complex128)
        _lib.check(self.__sim.q1t_circuit_read_amplitudes(
            self.__ptr, col, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1 << n))
        return out

    def stats(self):
        s = _lib.Stats()
        _lib.check(self.__sim.q1t_circuit_last_stats(self.__ptr, ctypes.byref(s)))
        return {"gate_seconds": s.gate_seconds, "bytes": s.bytes, "launches": s.launches, "sweeps": s.sweeps,
                "gates": s.gates, "blocks": s.blocks, "phases": s.phases,
                "mfma_flops": s.mfma_flops}

    def histogram_vec(self):
        n = 1 << self.nr_cbits()
        out = (ctypes.c_size_t * n)()
        _lib.check(self.__sim.q1t_circuit_histogram_vec(self.__ptr, out, n))
        return [int(out[i]) for i in range(n)]

    def plan_order(self, precision="c128", search=True):
        """Gate indices (in add order) in the order the plain GPU plan applies
        them; computed on the host, no GPU needed (gate-only circuits)."""
        prec = {"c128": 0, "c64": 1}[precision]
        cnt = ctypes.c_size_t(0)
        _lib.check(self.__sim.q1t_circuit_plan_order(self.__ptr, prec, 1 if search else 0, None, 0, ctypes.byref(cnt)))
        out = (ctypes.c_size_t * max(1, cnt.value))()
        _lib.check(self.__sim.q1t_circuit_plan_order(self.__ptr, prec, 1 if search else 0, out, cnt.value,
                                                     ctypes.byref(cnt)))
        return [int(out[i]) for i in range(cnt.value)]

    @property
    def handle(self):
        """The circuit_t* of this circuit (for batches)."""
        return self.__ptr


class Batch(object):
    """Many circuits executed together (build-added, include/q1tsim.h).

    Every circuit runs as ``circuit.execute(nr_shots)`` would run it, with its
    own seed and the same classical state; the gate runs of all circuits are
    flushed together between their measurements.
    """

    def __init__(self, precision="c128", fusion_k=0, mem_budget=0):
        self.__sim = _lib.lib()
        h = ctypes.c_void_p()
        prec = {"c128": 0, "c64": 1}[precision]
        _lib.check(self.__sim.q1t_batch_new(prec, int(fusion_k), int(mem_budget), ctypes.byref(h)))
        self.__ptr = h
        self.__circuits = []

    def close(self):
        if self.__ptr:
            self.__sim.q1t_batch_free(self.__ptr)
            self.__ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, circuit):
        _lib.check(self.__sim.q1t_batch_add(self.__ptr, circuit.handle))
        self.__circuits.append(circuit)

    def execute(self, nr_shots, keep_states=False):
        _lib.check(self.__sim.q1t_batch_execute(self.__ptr, int(nr_shots), 1 if keep_states else 0))

    def autotune(self, nr_shots, sample=64):
        """Time the first `sample` circuits under each compiled tile shape
        (and complex64 fusion width); keep the fastest.  Returns
        [(M, R, L, fusion_k, seconds)] for every configuration timed."""
        arr = (_lib.BatchConfig * 16)()
        n = ctypes.c_size_t(0)
        _lib.check(self.__sim.q1t_batch_autotune(self.__ptr, int(nr_shots), int(sample), arr, 16, ctypes.byref(n)))
        return [(arr[i].M, arr[i].R, arr[i].L, arr[i].fusion_k, arr[i].seconds) for i in range(min(n.value, 16))]

    def config(self):
        c = _lib.BatchConfig()
        _lib.check(self.__sim.q1t_batch_get_config(self.__ptr, ctypes.byref(c)))
        return (c.M, c.R, c.L, c.fusion_k)

    def set_config(self, M, R, L, fusion_k):
        c = _lib.BatchConfig(M, R, L, fusion_k, 0.0)
        _lib.check(self.__sim.q1t_batch_set_config(self.__ptr, ctypes.byref(c)))

    def set_tile_chunk(self, chunk):
        """Consecutive tiles a workgroup of a batched sweep takes at a time
        (1..1024); changes the schedule only."""
        _lib.check(self.__sim.q1t_batch_set_tile_chunk(self.__ptr, int(chunk)))

    def stats(self):
        s = _lib.BatchStats()
        _lib.check(self.__sim.q1t_batch_stats_get(self.__ptr, ctypes.byref(s)))
        g = s.gates
        return {"seconds": s.seconds, "waves": s.waves, "rounds": s.rounds, "sync_ops": s.sync_ops,
                "start_s": s.start_s, "queue_s": s.queue_s, "flush_s": s.flush_s, "measure_s": s.measure_s,
                "other_s": s.other_s, "column_launches": s.column_launches, "solo_ops": s.solo_ops,
                "plan_s": s.plan_s, "pack_s": s.pack_s,
                "bytes": g.bytes, "launches": g.launches, "sweeps": g.sweeps, "gates": g.gates,
                "blocks": g.blocks, "phases": g.phases}


__all__ = ["Circuit", "Batch", "RefParam", "Q1tError"]
