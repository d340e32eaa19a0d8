"""State-level Python wrapper of the GpuVectorState C ABI (include/q1tsim_gvs.h).

``GpuVectorState`` exposes the QuState operations (src/qustate.rs:5-89) on
the MI355X statevector; gates are given as FP64 matrices plus qubit indices,
as the trait hands them to a backend.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib, sizes

PREC = {"c128": 0, "c64": 1}


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u64ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


_matrix_cache = {}


def gate_matrix(name, params=()):
    """Gate::matrix() of a named gate from the library's gate table (FP64)."""
    key = (name.lower(), tuple(float(p) for p in params))
    m = _matrix_cache.get(key)
    if m is None:
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        out = np.zeros(2 * 64 * 64, dtype=np.float64)
        k = ctypes.c_size_t(0)
        check(lib().q1t_gate_matrix(name.encode(), _dptr(p) if len(p) else None, len(p), _dptr(out), len(out),
                                    ctypes.byref(k)))
        d = 1 << k.value
        m = out[:2 * d * d].view(np.complex128).reshape(d, d).copy()
        if len(_matrix_cache) < 4096:
            _matrix_cache[key] = m
    return m


# Small matrices passed to apply_gate again and again (gate_matrix hands out
# the same cached arrays) keep their address and size here, looked up by
# identity: numpy's ctypes interface costs a few microseconds a call, about
# half of the per-gate host time of a 900-gate circuit.  The entries hold
# their arrays, so an address stays valid; the C side copies the matrix at
# each call, so later changes to an array's values are seen.
_arg_cache = {}
_bits_cache = {}


def _matrix_arg(matrix):
    """(address, qubit count, object to keep alive during the call)."""
    hit = _arg_cache.get(id(matrix))
    if hit is not None and hit[0] is matrix:
        return hit[1], hit[2], None
    m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
    addr, k = m.ctypes.data, m.shape[0].bit_length() - 1
    if m is matrix and m.shape[0] <= 16:
        if len(_arg_cache) >= 4096:
            _arg_cache.clear()
        _arg_cache[id(matrix)] = (matrix, addr, k)
    return addr, k, m


def _bits_arg(bits):
    key = tuple(int(b) for b in bits)
    arr = _bits_cache.get(key)
    if arr is None:
        arr = sizes(key)
        if len(_bits_cache) >= 4096:
            _bits_cache.clear()
        _bits_cache[key] = arr
    return arr


def copy_bandwidth(nbytes=8 << 30, reps=5):
    """Device-to-device copy of `nbytes`, best of `reps` (HIP events):
    (ms, GB/s counting the bytes read and written)."""
    ms = ctypes.c_double(0.0)
    check(lib().q1t_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value, 2.0 * nbytes / (ms.value * 1e-3) / 1e9


def jit_stats():
    """(kernels compiled, launches of loaded kernels, compile seconds) of
    the specialised sweep kernels in this process."""
    c, h, sec = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_double(0.0)
    check(lib().q1t_jit_stats(ctypes.byref(c), ctypes.byref(h), ctypes.byref(sec)))
    return int(c.value), int(h.value), float(sec.value)


def set_device_cache_limit(nbytes=0):
    """Cap on released device blocks kept for reuse (0: 7/8 of device
    memory); lowering it frees the oldest cached blocks now."""
    check(lib().q1t_device_cache_limit(int(nbytes)))


def device_cache_bytes():
    out = ctypes.c_size_t(0)
    check(lib().q1t_device_cache_bytes(ctypes.byref(out)))
    return int(out.value)


def compare(a, b, col_a=0, col_b=0):
    """(infidelity, max-abs difference) between one column of each state (FP64 on device)."""
    out = np.zeros(5, dtype=np.float64)
    check(lib().q1t_gvs_compare(a.h, col_a, b.h, col_b, _dptr(out)))
    ov = complex(out[0], out[1])
    fid = abs(ov) ** 2 / (out[3] * out[4])
    return 1.0 - fid, float(out[2])


class Xoshiro:
    """The library's xoshiro256** stream (SplitMix64 seeding) as a q1t_rng."""

    def __init__(self, seed):
        self.h = lib().q1t_xoshiro_new(int(seed))
        fn = ctypes.cast(lib().q1t_xoshiro_next, ctypes.c_void_p).value
        self.rng = _lib.Rng(self.h, ctypes.CFUNCTYPE(ctypes.c_uint64, ctypes.c_void_p)(fn))

    def __del__(self):
        if getattr(self, "h", None):
            lib().q1t_xoshiro_free(self.h)
            self.h = None


def _jit_mode(on, prefetch, factors):
    """The C ABI's specialised-kernel mode: 0 off, 1 on, 2 on with prefetch,
    3 on without tile factors."""
    if not on:
        return 0
    if not factors:
        if prefetch:
            raise ValueError("prefetching kernels always fold tile factors")
        return 3
    return 2 if prefetch else 1


class GpuVectorState:
    def __init__(self, nr_bits, nr_shots, precision="c128"):
        h = ctypes.c_void_p()
        check(lib().q1t_gvs_new(nr_bits, nr_shots, PREC[precision], ctypes.byref(h)))
        self.h = h
        self.nr_bits, self.nr_shots = nr_bits, nr_shots

    def close(self):
        if getattr(self, "h", None):
            lib().q1t_gvs_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_qubit_coefs(cls, coefs, nr_shots, precision="c128"):
        s = cls(len(coefs) // 2, nr_shots, precision)
        c = np.ascontiguousarray(np.asarray(coefs, dtype=np.complex128))
        check(lib().q1t_gvs_from_qubit_coefs(s.h, _dptr(c), len(c)))
        return s

    @classmethod
    def from_amplitudes(cls, amps, nr_shots, precision="c128"):
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128))
        s = cls(int(np.log2(len(a))), nr_shots, precision)
        check(lib().q1t_gvs_set_amplitudes(s.h, _dptr(a), len(a)))
        return s

    def set_random_normal(self, seed):
        check(lib().q1t_gvs_set_random_normal(self.h, int(seed)))

    def set_plan(self, fusion_k=0, tiling=True):
        check(lib().q1t_gvs_set_plan(self.h, int(fusion_k), 1 if tiling else 0))

    def set_plan_search(self, mode=-1):
        """Local search of tiles and register phases: -1 by state size
        (512 MiB and more), 0 never, 1 always."""
        check(lib().q1t_gvs_set_plan_search(self.h, int(mode)))

    def set_graph(self, on=True):
        """Capture a flush's launches into a hipGraph and replay it while
        later flushes have the same launch signature (sweep programs are
        uploaded before each replay)."""
        check(lib().q1t_gvs_set_graph(self.h, int(on)))

    def set_jit(self, on=True, prefetch=False, factors=True):
        """Sweep kernels specialised to each sweep program (compiled with
        hipRTC on first use, cached per process by program structure);
        `prefetch`: the next tile is loaded while the current one runs;
        `factors`: diagonal ops on bits outside the tile folded into one
        factor per tile (without them the amplitudes equal the interpreter
        kernels' bit for bit)."""
        check(lib().q1t_gvs_set_jit(self.h, _jit_mode(on, prefetch, factors)))

    def graph_stats(self):
        """(graphs captured, replays launched) over the state's life."""
        c = ctypes.c_uint64(0)
        r = ctypes.c_uint64(0)
        check(lib().q1t_gvs_graph_stats(self.h, ctypes.byref(c), ctypes.byref(r)))
        return int(c.value), int(r.value)

    def set_geometry(self, M, R, L):
        check(lib().q1t_gvs_set_geometry(self.h, int(M), int(R), int(L)))

    @property
    def ncols(self):
        return lib().q1t_gvs_nr_cols(self.h)

    @property
    def counts(self):
        n = self.ncols
        out = (ctypes.c_size_t * max(1, n))()
        check(lib().q1t_gvs_counts(self.h, out, n))
        return [int(out[i]) for i in range(n)]

    def column(self, col):
        out = np.zeros(1 << self.nr_bits, dtype=np.complex128)
        check(lib().q1t_gvs_read_amplitudes(self.h, col, _dptr(out), len(out)))
        return out

    @property
    def states(self):
        return np.stack([self.column(c) for c in range(self.ncols)], axis=1)

    def apply_gate(self, matrix, bits, desc="gate"):
        addr, k, keep = _matrix_arg(matrix)
        check(lib().q1t_gvs_apply_gate(self.h, addr, k, _bits_arg(bits), len(bits), desc.encode()))
        del keep

    def apply_unary_gate_all(self, matrix, desc="gate"):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
        check(lib().q1t_gvs_apply_unary_gate_all(self.h, _dptr(m), desc.encode()))

    def apply_conditional_gate(self, control, matrix, bits, desc="gate"):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
        k = int(np.log2(m.shape[0]))
        c = np.ascontiguousarray(np.asarray(control, dtype=np.uint8))
        check(lib().q1t_gvs_apply_conditional_gate(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(c),
                                                   _dptr(m), k, sizes(bits), len(bits), desc.encode()))

    def marginal_w0(self, qbit):
        out = np.zeros(self.ncols, dtype=np.float64)
        check(lib().q1t_gvs_marginal_w0(self.h, qbit, _dptr(out), len(out)))
        return out

    def marginal_w0_noalloc(self, qbit):
        """marginal_w0 through the allocation-free entry point (at most 256
        columns, no pending gates: call flush() first)."""
        out = np.zeros(self.ncols, dtype=np.float64)
        check(lib().q1t_gvs_marginal_w0_noalloc(self.h, qbit, _dptr(out), len(out)))
        return out

    def measure_into(self, qbit, cbit, res, rng, collapse=True, basis="Z"):
        if basis != "Z":
            check(lib().q1t_gvs_measure_basis_into(self.h, qbit, cbit, basis.encode(), _u64ptr(res), len(res),
                                                   ctypes.byref(rng.rng), 1 if collapse else 0))
            return
        f = lib().q1t_gvs_measure_into if collapse else lib().q1t_gvs_peek_into
        check(f(self.h, qbit, cbit, _u64ptr(res), len(res), ctypes.byref(rng.rng)))

    def measure_with_n0(self, qbit, cbit, n0s, w0s, res, collapse=True):
        n0 = sizes(n0s)
        w0 = np.ascontiguousarray(np.asarray(w0s, dtype=np.float64))
        check(lib().q1t_gvs_measure_with_n0(self.h, qbit, cbit, n0, _dptr(w0), len(w0s), _u64ptr(res), len(res),
                                            1 if collapse else 0))

    def measure_all_into(self, cbits, res, rng, collapse=True):
        f = lib().q1t_gvs_measure_all_into if collapse else lib().q1t_gvs_peek_all_into
        check(f(self.h, sizes(cbits), len(cbits), _u64ptr(res), len(res), ctypes.byref(rng.rng)))

    def measure_all_uniforms(self, cbits, u01, res, collapse=True):
        u = np.ascontiguousarray(np.asarray(u01, dtype=np.float64))
        check(lib().q1t_gvs_measure_all_uniforms(self.h, sizes(cbits), len(cbits), _dptr(u), _u64ptr(res), len(res),
                                                 1 if collapse else 0))

    def reset(self, bit, rng):
        check(lib().q1t_gvs_reset(self.h, bit, ctypes.byref(rng.rng)))

    def reset_all(self):
        check(lib().q1t_gvs_reset_all(self.h))

    def flush(self):
        check(lib().q1t_gvs_flush(self.h))

    def stats(self):
        s = _lib.Stats()
        check(lib().q1t_gvs_last_stats(self.h, ctypes.byref(s)))
        return {"gate_seconds": s.gate_seconds, "bytes": s.bytes, "launches": s.launches, "sweeps": s.sweeps,
                "gates": s.gates, "blocks": s.blocks, "phases": s.phases,
                "mfma_flops": s.mfma_flops}

    def set_profiling(self, on=True):
        check(lib().q1t_gvs_set_profiling(self.h, 1 if on else 0))

    def kernel_profile(self):
        """{kernel name: (launches, total ms, algorithmic bytes)} from HIP events."""
        n = ctypes.c_size_t(0)
        check(lib().q1t_gvs_kernel_profile(self.h, None, 0, ctypes.byref(n)))
        arr = (_lib.KernelProfile * max(1, n.value))()
        check(lib().q1t_gvs_kernel_profile(self.h, ctypes.cast(arr, ctypes.c_void_p), n.value, ctypes.byref(n)))
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms), float(arr[i].bytes))
                for i in range(n.value)}

    def apply_named(self, name, bits, params=()):
        self.apply_gate(gate_matrix(name, params), bits, name)

    # ---- learned planner hooks ----
    PLAN_GREEDY, PLAN_LEARNED, PLAN_EXPLORE = 0, 1, 2

    def set_cost_model(self, cm, mode=1, seed=1):
        """Planner of fused complex64 flushes: mode 0 greedy fixed-width,
        1 ranked by `cm` (greedy while untrained), 2 seeded random choice."""
        check(lib().q1t_gvs_set_cost_model(self.h, cm.h if cm is not None else None, int(mode), int(seed)))

    def set_recording(self, on=True):
        check(lib().q1t_gvs_set_recording(self.h, 1 if on else 0))

    def take_samples(self, cap=1 << 20):
        """Recorded sweeps: (features n x 20 float32, tiles n, ms n)."""
        from .costmodel import NFEAT
        feat = np.zeros((cap, NFEAT), dtype=np.float32)
        tiles = np.zeros(cap, dtype=np.float64)
        ms = np.zeros(cap, dtype=np.float64)
        n = ctypes.c_size_t(0)
        check(lib().q1t_gvs_take_samples(self.h, feat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _dptr(tiles),
                                         _dptr(ms), cap, ctypes.byref(n)))
        k = n.value
        return feat[:k].copy(), tiles[:k].copy(), ms[:k].copy()
