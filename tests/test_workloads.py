"""The synthetic circuit recipes: gate counts of the layered recipe, and the
inverse circuit the full-size round-trip tests rely on (checked with the
dense simulator, not the code under test)."""
import numpy as np

from q1tsim_amd import workloads
from tests.golden import dense_sim


def test_layered_counts():
    ops = workloads.random_layered(30, 20, 1)
    assert sum(len(b) == 1 for _, b, _ in ops) == 600
    assert sum(name == "cx" for name, _, _ in ops) == 290
    ops = workloads.random_layered(18, 20, 1)
    assert len(ops) == 530


def test_inverse_round_trip_dense():
    n = 6
    rng = np.random.default_rng(3)
    ops = workloads.random_pairs(n, 10, 2) + workloads.qft(n) + [
        ("u3", [1], [0.3, 1.1, -0.7]), ("cu3", [2, 4], [0.9, 0.2, 1.4]), ("ccrz", [0, 1, 2], [0.5]),
        ("v", [3], []), ("ct", [1, 5], []), ("u1", [0], [2.2]), ("sdg", [4], []), ("ccx", [5, 0, 3], [])]
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    b = dense_sim.run(n, ops + workloads.inverse(ops), state=a.copy())
    assert np.abs(a - b).max() < 1e-13
    # a planted sign mistake (a rotation not negated) does not round-trip
    bad = [(nm, b_, list(p)) for nm, b_, p in workloads.inverse(ops)]
    for k, (nm, _, p) in enumerate(bad):
        if nm == "u3":
            bad[k] = (nm, bad[k][1], [-p[0], p[1], p[2]])
            break
    c = dense_sim.run(n, ops + bad, state=a.copy())
    assert np.abs(a - c).max() > 1e-3
