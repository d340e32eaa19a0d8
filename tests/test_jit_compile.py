"""Run-time specialised sweep kernels compiled on the host (hipRTC, no GPU):
every op form of the plain and fused programs compiles, and the code
objects hold no packed FP32 arithmetic (v_pk_fma/mul/add_f32) and no
scratch memory."""
import os
import re
import subprocess
import tempfile

import pytest

from q1tsim_amd import workloads
from q1tsim_amd.q1tsim import Circuit

OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"
READELF = "/opt/rocm/llvm/bin/llvm-readelf"


def mixed_circuit(n, seed):
    ops = workloads.random_pairs(n, 6, seed) + workloads.qft(min(n, 8))
    ops += [("y", [1], []), ("rx", [2], [0.3]), ("ccx", [0, 3, n - 1], []), ("cz", [n - 1, 2], []),
            ("swap", [0, n - 2], []), ("u3", [4], [0.1, 0.2, 0.3]), ("crz", [n - 3, 1], [0.7]),
            ("ch", [2, n - 1], []), ("tdg", [n - 1], []), ("cy", [n - 2, 5], [])]
    return ops


@pytest.mark.parametrize("precision,fusion_k,prefetch", [("c128", 0, False), ("c64", 0, False), ("c64", 4, False),
                                                         ("c64", 6, False), ("c64", 0, True)])
def test_specialised_kernels_compile(precision, fusion_k, prefetch):
    n = 16
    c = Circuit(n, 0)
    for name, bits, params in mixed_circuit(n, 3):
        c.add_gate(name, bits, params if params else None)
    with tempfile.TemporaryDirectory() as d:
        nk, secs = c.jit_compile(precision, fusion_k, d, prefetch=prefetch)
        assert nk >= 1
        for i in range(nk):
            co = os.path.join(d, "sweep_%d.co" % i)
            asm = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", co], capture_output=True, text=True).stdout
            assert ("q1t_jit_sweep_c64" if precision == "c64" else "q1t_jit_sweep_c128") in asm
            assert not re.search(r"v_pk_(fma|mul|add)_f32", asm), "packed FP32 in %s" % co
            notes = subprocess.run([READELF, "--notes", co], capture_output=True, text=True).stdout
            m = re.search(r"\.private_segment_fixed_size:\s+(\d+)", notes)
            assert m and int(m.group(1)) == 0, "scratch memory in %s" % co
            src = open(os.path.join(d, "sweep_%d.hip" % i)).read()
            if fusion_k:
                assert "block_mfma" in src or "dense_pairs" in src or "phase_op" in src
            assert ("pf0" in src) == prefetch


def test_jit_compile_rejects_non_gate_ops():
    c = Circuit(4, 1)
    c.h(0)
    c.measure(0, 0)
    with pytest.raises(Exception, match="non-gate ops"):
        c.jit_compile("c128")


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_tile_factors_only_when_asked(precision):
    """Diagonal gates on qubits outside the tile of a 20-qubit state (the
    tile holds the dense gates' bits and the lowest bits) become tile
    factors in the default specialised kernels and stay ops without them
    (mode 3)."""
    n = 20
    c = Circuit(n, 0)
    for q in range(8):
        c.add_gate("h", [q], None)
    for q in range(8, 14):
        c.add_gate("rz", [q], [0.1 * q])
        c.add_gate("t", [q], None)
    for q in range(8):
        c.add_gate("rx", [q], [0.2])
    with tempfile.TemporaryDirectory() as d:
        nk, _ = c.jit_compile(precision, 0, d)
        srcs = [open(os.path.join(d, "sweep_%d.hip" % i)).read() for i in range(nk)]
        assert sum(s.count("fac = cmul(fac") for s in srcs) == 12
    with tempfile.TemporaryDirectory() as d:
        nk, _ = c.jit_compile(precision, 0, d, factors=False)
        srcs = [open(os.path.join(d, "sweep_%d.hip" % i)).read() for i in range(nk)]
        assert not any("fac = cmul(fac" in s for s in srcs)
    with pytest.raises(ValueError):
        c.jit_compile(precision, 0, None, prefetch=True, factors=False)
