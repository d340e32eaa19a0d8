"""A flush that repeats the previous flush's gate structure reuses its plan
(states of 512 MiB and more, where the tile search runs): new angles must
still be applied, and the amplitudes must equal those of the same gates
planned afresh on a new state, bit for bit."""
import numpy as np
import pytest

from oracle import oracle as orc
from q1tsim_amd import workloads
from q1tsim_amd.gvs import GpuVectorState

pytestmark = pytest.mark.gpu

N = 25   # complex128: 512 MiB, the plan search's threshold


def with_angles(ops, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [(name, bits, [float(rng.uniform(0.0, 2.0 * np.pi)) for _ in params]) for name, bits, params in ops]


def run(state, ops):
    for name, bits, params in ops:
        state.apply_gate(orc.gate_matrix(name, params), bits)
    state.flush()


@pytest.mark.parametrize("jit", [False, True])
def test_repeated_structure_reuses_plan_with_new_angles(jit):
    base = workloads.random_layered(N, 4, 5)
    a, b = with_angles(base, 1), with_angles(base, 2)
    assert any(pa != pb for (_, _, pa), (_, _, pb) in zip(a, b))

    fresh = GpuVectorState(N, 1, "c128")
    fresh.set_jit(jit)
    fresh.set_random_normal(9)
    run(fresh, b)
    want = fresh.column(0)
    fresh.close()

    st = GpuVectorState(N, 1, "c128")
    st.set_jit(jit)
    st.set_random_normal(9)
    run(st, a)                  # plans and keeps the plan
    st.set_random_normal(9)
    run(st, b)                  # same structure, other angles: the kept plan
    got = st.column(0)
    st.close()
    assert np.array_equal(got, want)

    # and a different structure is planned anew
    other = workloads.random_layered(N, 4, 6)
    s1 = GpuVectorState(N, 1, "c128")
    s1.set_jit(jit)
    s1.set_random_normal(9)
    run(s1, a)
    s1.set_random_normal(9)
    run(s1, other)
    got2 = s1.column(0)
    s1.close()
    s2 = GpuVectorState(N, 1, "c128")
    s2.set_jit(jit)
    s2.set_random_normal(9)
    run(s2, other)
    want2 = s2.column(0)
    s2.close()
    assert np.array_equal(got2, want2)


def test_repeated_structure_fused_c64():
    # greedy 4-qubit fusion on a 512 MiB complex64 state: the kept plan's
    # blocks are composed from the new matrices
    n = 26
    base = workloads.random_layered(n, 4, 7)
    a, b = with_angles(base, 3), with_angles(base, 4)
    fresh = GpuVectorState(n, 1, "c64")
    fresh.set_plan(4, True)
    fresh.set_random_normal(2)
    run(fresh, b)
    want = fresh.column(0)
    fresh.close()
    st = GpuVectorState(n, 1, "c64")
    st.set_plan(4, True)
    st.set_random_normal(2)
    run(st, a)
    st.set_random_normal(2)
    run(st, b)
    got = st.column(0)
    st.close()
    assert np.array_equal(got, want)
