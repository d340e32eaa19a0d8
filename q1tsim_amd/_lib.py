"""ctypes binding of libq1tsim.so (include/q1tsim.h, include/q1tsim_gvs.h).

The library is the product: there is no fallback.  If it has not been built,
importing raises with the build command to run.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libq1tsim.so")

RESULT_ERROR, RESULT_EMPTY, RESULT_STRING, RESULT_HISTOGRAM, RESULT_CSTATE = 0, 1, 2, 3, 5


class HistElem(ctypes.Structure):
    _fields_ = [("key", ctypes.c_char_p), ("count", ctypes.c_size_t)]


class Parameter(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double), ("value_ptr", ctypes.POINTER(ctypes.c_double))]


class Result(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("length", ctypes.c_size_t), ("size", ctypes.c_size_t),
                ("restype", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("gate_seconds", ctypes.c_double), ("bytes", ctypes.c_double), ("launches", ctypes.c_uint64),
                ("sweeps", ctypes.c_uint64), ("gates", ctypes.c_uint64), ("blocks", ctypes.c_uint64),
                ("phases", ctypes.c_uint64), ("mfma_flops", ctypes.c_double)]


class KernelProfile(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double),
                ("bytes", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("gates", Stats), ("waves", ctypes.c_uint64), ("rounds", ctypes.c_uint64),
                ("sync_ops", ctypes.c_uint64), ("seconds", ctypes.c_double), ("start_s", ctypes.c_double),
                ("queue_s", ctypes.c_double), ("flush_s", ctypes.c_double), ("measure_s", ctypes.c_double),
                ("other_s", ctypes.c_double), ("column_launches", ctypes.c_uint64), ("solo_ops", ctypes.c_uint64),
                ("plan_s", ctypes.c_double), ("pack_s", ctypes.c_double)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("R", ctypes.c_int), ("L", ctypes.c_int), ("fusion_k", ctypes.c_int),
                ("seconds", ctypes.c_double)]


class Rng(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("next_u64", ctypes.CFUNCTYPE(ctypes.c_uint64, ctypes.c_void_p))]


_vp, _sz, _u64, _i32, _d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
_ps = ctypes.POINTER(ctypes.c_size_t)
_pd = ctypes.POINTER(ctypes.c_double)
_pf = ctypes.POINTER(ctypes.c_float)
_pu64 = ctypes.POINTER(ctypes.c_uint64)
_pu8 = ctypes.POINTER(ctypes.c_uint8)
_cs = ctypes.c_char_p

SIGNATURES = {
    "result_free": (None, [Result]),
    "circuit_new": (_vp, [_sz, _sz]),
    "circuit_free": (None, [_vp]),
    "circuit_nr_qbits": (_sz, [_vp]),
    "circuit_nr_cbits": (_sz, [_vp]),
    "circuit_cstate": (Result, [_vp]),
    "circuit_add_gate": (Result, [_vp, _cs, _ps, _sz, ctypes.POINTER(Parameter), _sz]),
    "circuit_add_conditional_gate": (Result, [_vp, _ps, _sz, _u64, _cs, _ps, _sz, ctypes.POINTER(Parameter), _sz]),
    "circuit_reset": (Result, [_vp, _sz]),
    "circuit_reset_all": (Result, [_vp]),
    "circuit_measure": (Result, [_vp, _sz, _sz, ctypes.c_char, ctypes.c_uint8]),
    "circuit_measure_all": (Result, [_vp, _ps, _sz, ctypes.c_char, ctypes.c_uint8]),
    "circuit_execute": (Result, [_vp, _sz]),
    "circuit_reexecute": (Result, [_vp]),
    "circuit_histogram": (Result, [_vp]),
    "circuit_latex": (Result, [_vp]),
    "circuit_open_qasm": (Result, [_vp]),
    "circuit_c_qasm": (Result, [_vp]),
    "q1t_last_error": (_cs, []),
    "q1t_circuit_set_backend": (_i32, [_vp, _i32, _i32, _i32]),
    "q1t_circuit_seed": (_i32, [_vp, _u64]),
    "q1t_circuit_set_graph": (_i32, [_vp, _i32]),
    "q1t_circuit_graph_stats": (_i32, [_vp, _pu64, _pu64]),
    "q1t_circuit_set_jit": (_i32, [_vp, _i32]),
    "q1t_circuit_jit_compile": (_i32, [_vp, _i32, _i32, _i32, ctypes.c_char_p, _ps, _pd]),
    "q1t_jit_stats": (_i32, [_pu64, _pu64, _pd]),
    "q1t_circuit_set_initial_amplitudes": (_i32, [_vp, _pd, _sz]),
    "q1t_circuit_set_initial_random": (_i32, [_vp, _u64]),
    "q1t_circuit_barrier": (_i32, [_vp, _ps, _sz]),
    "q1t_circuit_nr_columns": (_i32, [_vp, _ps]),
    "q1t_circuit_counts": (_i32, [_vp, _ps, _sz]),
    "q1t_circuit_read_amplitudes": (_i32, [_vp, _sz, _pd, _sz]),
    "q1t_circuit_last_stats": (_i32, [_vp, ctypes.POINTER(Stats)]),
    "q1t_circuit_histogram_vec": (_i32, [_vp, _ps, _sz]),
    "q1t_circuit_plan_order": (_i32, [_vp, _i32, _i32, _ps, _sz, _ps]),
    "q1t_cm_new": (_i32, [_u64, ctypes.POINTER(_vp)]),
    "q1t_cm_free": (None, [_vp]),
    "q1t_cm_get_params": (_i32, [_vp, _pf, _sz]),
    "q1t_cm_set_params": (_i32, [_vp, _pf, _sz]),
    "q1t_cm_set_norm": (_i32, [_vp, _pf, _pf, _sz]),
    "q1t_cm_get_norm": (_i32, [_vp, _pf, _pf, _sz]),
    "q1t_cm_predict": (_i32, [_vp, _pf, _sz, _pf]),
    "q1t_cm_loss_grad": (_i32, [_vp, _pf, _pf, _sz, _pf, _sz, _pd]),
    "q1t_cm_adam_step": (_i32, [_vp, _pf, _pf, _sz, ctypes.c_float, _pd]),
    "q1t_cm_train": (_i32, [_vp, _pf, _pf, _sz, _i32, _sz, ctypes.c_float, _u64, _pd]),
    "q1t_cm_set_trained": (_i32, [_vp, _i32]),
    "q1t_gvs_set_cost_model": (_i32, [_vp, _vp, _i32, _u64]),
    "q1t_gvs_set_recording": (_i32, [_vp, _i32]),
    "q1t_gvs_take_samples": (_i32, [_vp, _pf, _pd, _pd, _sz, _ps]),
    "q1t_batch_new": (_i32, [_i32, _i32, _sz, ctypes.POINTER(_vp)]),
    "q1t_batch_free": (None, [_vp]),
    "q1t_batch_add": (_i32, [_vp, _vp]),
    "q1t_batch_execute": (_i32, [_vp, _sz, _i32]),
    "q1t_batch_stats_get": (_i32, [_vp, ctypes.POINTER(BatchStats)]),
    "q1t_batch_autotune": (_i32, [_vp, _sz, _sz, ctypes.POINTER(BatchConfig), _sz, _ps]),
    "q1t_batch_set_config": (_i32, [_vp, ctypes.POINTER(BatchConfig)]),
    "q1t_batch_set_tile_chunk": (_i32, [_vp, ctypes.c_uint]),
    "q1t_batch_get_config": (_i32, [_vp, ctypes.POINTER(BatchConfig)]),
    "q1t_gvs_new": (_i32, [_sz, _sz, _i32, ctypes.POINTER(_vp)]),
    "q1t_gvs_free": (None, [_vp]),
    "q1t_gvs_from_qubit_coefs": (_i32, [_vp, _pd, _sz]),
    "q1t_gvs_set_amplitudes": (_i32, [_vp, _pd, _sz]),
    "q1t_gvs_set_random_normal": (_i32, [_vp, _u64]),
    "q1t_gvs_nr_bits": (_sz, [_vp]),
    "q1t_gvs_nr_cols": (_sz, [_vp]),
    "q1t_gvs_counts": (_i32, [_vp, _ps, _sz]),
    "q1t_gvs_apply_gate": (_i32, [_vp, _vp, _sz, _ps, _sz, _cs]),
    "q1t_gvs_apply_unary_gate_all": (_i32, [_vp, _pd, _cs]),
    "q1t_gvs_apply_conditional_gate": (_i32, [_vp, _pu8, _sz, _pd, _sz, _ps, _sz, _cs]),
    "q1t_gvs_marginal_w0": (_i32, [_vp, _sz, _pd, _sz]),
    "q1t_gvs_marginal_w0_noalloc": (_i32, [_vp, _sz, _pd, _sz]),
    "q1t_gvs_measure_with_n0": (_i32, [_vp, _sz, _sz, _ps, _pd, _sz, _pu64, _sz, _i32]),
    "q1t_gvs_measure_into": (_i32, [_vp, _sz, _sz, _pu64, _sz, ctypes.POINTER(Rng)]),
    "q1t_gvs_peek_into": (_i32, [_vp, _sz, _sz, _pu64, _sz, ctypes.POINTER(Rng)]),
    "q1t_gvs_measure_basis_into": (_i32, [_vp, _sz, _sz, ctypes.c_char, _pu64, _sz, ctypes.POINTER(Rng), _i32]),
    "q1t_gvs_measure_all_into": (_i32, [_vp, _ps, _sz, _pu64, _sz, ctypes.POINTER(Rng)]),
    "q1t_gvs_peek_all_into": (_i32, [_vp, _ps, _sz, _pu64, _sz, ctypes.POINTER(Rng)]),
    "q1t_gvs_measure_all_uniforms": (_i32, [_vp, _ps, _sz, _pd, _pu64, _sz, _i32]),
    "q1t_gvs_reset": (_i32, [_vp, _sz, ctypes.POINTER(Rng)]),
    "q1t_gvs_reset_all": (_i32, [_vp]),
    "q1t_gvs_read_amplitudes": (_i32, [_vp, _sz, _pd, _sz]),
    "q1t_gvs_flush": (_i32, [_vp]),
    "q1t_gvs_set_plan": (_i32, [_vp, _i32, _i32]),
    "q1t_gvs_set_plan_search": (_i32, [_vp, _i32]),
    "q1t_gvs_set_graph": (_i32, [_vp, _i32]),
    "q1t_gvs_set_jit": (_i32, [_vp, _i32]),
    "q1t_gvs_graph_stats": (_i32, [_vp, _pu64, _pu64]),
    "q1t_gvs_set_geometry": (_i32, [_vp, _i32, _i32, _i32]),
    "q1t_gvs_last_stats": (_i32, [_vp, ctypes.POINTER(Stats)]),
    "q1t_gvs_set_profiling": (_i32, [_vp, _i32]),
    "q1t_gvs_kernel_profile": (_i32, [_vp, _vp, _sz, _ps]),
    "q1t_copy_bandwidth": (_i32, [_sz, _i32, _pd]),
    "q1t_device_cache_limit": (_i32, [_sz]),
    "q1t_device_cache_bytes": (_i32, [_ps]),
    "q1t_gvs_compare": (_i32, [_vp, _sz, _vp, _sz, _pd]),
    "q1t_gate_matrix": (_i32, [_cs, _pd, _sz, _pd, _sz, _ps]),
    "q1t_xoshiro_new": (_vp, [_u64]),
    "q1t_xoshiro_free": (None, [_vp]),
    "q1t_xoshiro_next": (_u64, [_vp]),
}

_lib = None


class Q1tError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libq1tsim.so; raise if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libq1tsim.so is not built; run `python -m q1tsim_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise Q1tError(rc, lib().q1t_last_error().decode("utf-8", "replace"))


def unpack_result(res):
    """Decode and free a result_t (the conventions of q1tsimffi.py:85-108)."""
    L = lib()
    try:
        if res.restype == RESULT_ERROR:
            raise Q1tError(-1, ctypes.cast(res.data, ctypes.c_char_p).value.decode("utf-8"))
        if res.restype == RESULT_EMPTY:
            return None
        if res.restype == RESULT_STRING:
            return ctypes.cast(res.data, ctypes.c_char_p).value.decode("utf-8")
        if res.restype == RESULT_HISTOGRAM:
            elems = ctypes.cast(res.data, ctypes.POINTER(HistElem))
            return {elems[i].key.decode("utf-8"): elems[i].count for i in range(res.length)}
        if res.restype == RESULT_CSTATE:
            ptr = ctypes.cast(res.data, ctypes.POINTER(ctypes.c_uint64))
            return [ptr[i] for i in range(res.length)]
        raise Q1tError(-1, "Unknown data type code: {}".format(res.restype))
    finally:
        L.result_free(res)


def sizes(vals):
    arr = (ctypes.c_size_t * max(1, len(vals)))(*[int(v) for v in vals])
    return arr
