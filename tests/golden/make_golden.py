"""Write the golden amplitude fixtures under tests/golden/.

The expected outputs come from dense_sim.py (closed-form matrices applied as
tensor contractions), not from the oracle or the GPU path.  Re-run with
``python tests/golden/make_golden.py``; the fixtures are committed.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import dense_sim  # noqa: E402
from q1tsim_amd import workloads  # noqa: E402

ALL_GATES = [
    ("i", 1, 0), ("h", 1, 0), ("x", 1, 0), ("y", 1, 0), ("z", 1, 0), ("s", 1, 0), ("sdg", 1, 0),
    ("t", 1, 0), ("tdg", 1, 0), ("rx", 1, 1), ("ry", 1, 1), ("rz", 1, 1), ("u1", 1, 1),
    ("u2", 1, 2), ("u3", 1, 3), ("v", 1, 0), ("vdg", 1, 0), ("swap", 2, 0),
    ("cx", 2, 0), ("cy", 2, 0), ("cz", 2, 0), ("ch", 2, 0), ("crx", 2, 1), ("cry", 2, 1),
    ("crz", 2, 1), ("cs", 2, 0), ("csdg", 2, 0), ("ct", 2, 0), ("ctdg", 2, 0), ("cu1", 2, 1),
    ("cu2", 2, 2), ("cu3", 2, 3), ("cv", 2, 0), ("cvdg", 2, 0), ("ccx", 3, 0), ("ccz", 3, 0),
    ("ccrx", 3, 1), ("ccry", 3, 1), ("ccrz", 3, 1),
]


def random_state(n, rng):
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


def all_gates_circuit(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = []
    for _ in range(3):
        for name, k, npar in ALL_GATES:
            bits = [int(b) for b in rng.permutation(n)[:k]]
            params = [float(p) for p in rng.uniform(0, 2 * np.pi, size=npar)]
            ops.append((name, bits, params))
    return ops


def encode_ops(ops):
    names = np.array([o[0] for o in ops])
    bits = np.full((len(ops), 3), -1, dtype=np.int64)
    params = np.zeros((len(ops), 3), dtype=np.float64)
    for i, (_, b, p) in enumerate(ops):
        bits[i, :len(b)] = b
        params[i, :len(p)] = p
    return names, bits, params


def main():
    out = {}
    for n, depth, seed in [(10, 20, 7), (8, 12, 3)]:
        ops = workloads.random_layered(n, depth, seed)
        amps = dense_sim.run(n, ops)
        np.savez_compressed(os.path.join(HERE, "layered_n%d_d%d_s%d.npz" % (n, depth, seed)),
                            n=n, depth=depth, seed=seed, amplitudes=amps)
        out["layered_%d" % n] = float(np.linalg.norm(amps))

    rng = np.random.Generator(np.random.PCG64(11))
    n = 8
    a = random_state(n, rng)
    b = dense_sim.run(n, workloads.qft(n), state=a.copy())
    np.savez_compressed(os.path.join(HERE, "qft_n8.npz"), n=n, input=a, output=b)

    n = 6
    ops = all_gates_circuit(n, 5)
    a = random_state(n, np.random.Generator(np.random.PCG64(13)))
    b = dense_sim.run(n, ops, state=a.copy())
    names, bits, params = encode_ops(ops)
    np.savez_compressed(os.path.join(HERE, "all_gates_n6.npz"), n=n, input=a, output=b,
                        names=names, bits=bits, params=params)
    print("wrote fixtures", out)


if __name__ == "__main__":
    main()
