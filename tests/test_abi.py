"""CPU tests of the C ABI: every declared symbol is exported, and the
host-side validation (errors of src/error.rs, ffi.rs null checks, the result
conventions) works without touching a GPU."""
import ctypes
import os
import re

import pytest

from q1tsim_amd import _lib
from q1tsim_amd.q1tsim import Circuit, Q1tError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("q1tsim.h", "q1tsim_gvs.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith(("circuit_", "result_", "q1t_")) and "(*" not in name:
                names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 55
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared symbol also has a ctypes signature in the binding
    assert not [n for n in names if n not in _lib.SIGNATURES]


def test_ffi_struct_layouts():
    # CResult {void*, size_t, size_t, u32} (ffi.rs:63-70), CHistElem, CParameter
    assert ctypes.sizeof(_lib.Result) == 32
    assert ctypes.sizeof(_lib.HistElem) == 16
    assert ctypes.sizeof(_lib.Parameter) == 16


def test_add_gate_errors():
    with Circuit(2, 2) as c:
        assert c.nr_qbits() == 2 and c.nr_cbits() == 2
        assert c.h(0) is None
        with pytest.raises(Q1tError, match='Unknown gate "foo"'):
            c.add_gate("foo", [0])
        with pytest.raises(Q1tError, match="Invalid index 5 for a quantum bit"):
            c.x(5)
        with pytest.raises(Q1tError, match='Expected 1 arguments to "RX" gate, got 0'):
            c.add_gate("rx", [0], [])
        with pytest.raises(Q1tError, match="Invalid index 3 for a classical bit"):
            c.measure(0, 3)
        with pytest.raises(Q1tError, match="Invalid index 7 for a classical bit"):
            c.add_conditional_gate([7], 1, "x", [0])
        with pytest.raises(Q1tError, match="Invalid index 4 for a classical bit"):
            c.measure_all([0, 4])
        with pytest.raises(Q1tError, match="Invalid measurement basis"):
            c.measure_basis(0, 0, "Q")
        # gate names are case-insensitive (ffi.rs:335 to_lowercase)
        assert c.add_gate("Cx", [0, 1]) is None


def test_not_executed_and_exports():
    with Circuit(2, 2) as c:
        with pytest.raises(Q1tError, match="has not been executed"):
            c.histogram()
        with pytest.raises(Q1tError, match="has not been run yet"):
            c.cstate()
        with pytest.raises(Q1tError, match="has not been executed"):
            c.reexecute()
        with pytest.raises(Q1tError, match="export is not provided"):
            c.open_qasm()


def test_null_pointers_give_error_results():
    L = _lib.lib()
    r = L.circuit_execute(None, 10)
    assert r.restype == _lib.RESULT_ERROR
    assert ctypes.cast(r.data, ctypes.c_char_p).value == b"Pointer to circuit is NULL"
    L.result_free(r)
    assert L.circuit_nr_qbits(None) == 0
    c = L.circuit_new(1, 1)
    r = L.circuit_add_gate(c, b"h", None, 1, None, 0)
    assert ctypes.cast(r.data, ctypes.c_char_p).value == b"Pointer to bit indices is NULL"
    L.result_free(r)
    L.circuit_free(c)
    L.circuit_free(None)


def test_xoshiro_stream_matches_oracle():
    from oracle import oracle as orc
    L = _lib.lib()
    h = L.q1t_xoshiro_new(12345)
    got = [L.q1t_xoshiro_next(h) for _ in range(8)]
    L.q1t_xoshiro_free(h)
    r = orc.Rng(12345)
    assert got == [r.next_u64() for _ in range(8)]


def test_no_gpu_fails_loudly():
    """Without a HIP device the state cannot be created: an error, never a CPU fallback."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with Circuit(2, 2) as c:
        c.h(0)
        with pytest.raises(Q1tError, match="HIP"):
            c.execute(10)


ALL_NAMES = [("i", 0), ("h", 0), ("x", 0), ("y", 0), ("z", 0), ("s", 0), ("sdg", 0), ("t", 0), ("tdg", 0),
             ("rx", 1), ("ry", 1), ("rz", 1), ("u1", 1), ("u2", 2), ("u3", 3), ("v", 0), ("vdg", 0),
             ("swap", 0), ("cx", 0), ("cy", 0), ("cz", 0), ("ch", 0), ("crx", 1), ("cry", 1), ("crz", 1),
             ("cs", 0), ("csdg", 0), ("ct", 0), ("ctdg", 0), ("cu1", 1), ("cu2", 2), ("cu3", 3), ("cv", 0),
             ("cvdg", 0), ("ccx", 0), ("ccz", 0), ("ccrx", 1), ("ccry", 1), ("ccrz", 1)]


def test_product_gate_table_matches_closed_forms():
    """The library's own gate table (the one Circuit.add_gate uses) against
    the independent closed forms of tests/golden/dense_sim.py, every name,
    several random parameter sets (distinct values, so a swapped parameter
    order fails)."""
    import numpy as np

    from q1tsim_amd.gvs import gate_matrix
    from tests.golden import dense_sim

    rng = np.random.Generator(np.random.PCG64(17))
    for name, npar in ALL_NAMES:
        for _ in range(4 if npar else 1):
            p = [float(x) for x in rng.uniform(-2 * np.pi, 2 * np.pi, size=npar)]
            got = gate_matrix(name, p)
            exp = dense_sim.gate_matrix(name, p)
            assert got.shape == exp.shape, name
            assert np.abs(got - exp).max() < 1e-15, (name, p)
            # upper-case names are accepted too (ffi.rs matches case-insensitively)
            assert np.array_equal(gate_matrix(name.upper(), p), got)


def test_product_gate_table_is_complete():
    """Every name of ffi.rs:335-366 plus the composite names the QFT needs."""
    from q1tsim_amd.gvs import gate_matrix
    for name, npar in ALL_NAMES:
        gate_matrix(name, [0.1 * (k + 1) for k in range(npar)])
    with pytest.raises(Q1tError, match='Unknown gate "cxx"'):
        gate_matrix("cxx", [])

