"""Binomial acceptance band, restating src/stats.rs:10-33 (test-only helper).

``erf_inv`` comes from scipy instead of statrs 0.7; the band values of
src/stats.rs:40-59 pin it (tests/test_oracle.py::test_stats_bounds).
"""
import math

from scipy.special import erfinv


def get_bounds(nr_shots, p, tol):
    mu = nr_shots * p
    sigma = math.sqrt(nr_shots * p * (1.0 - p))
    quantile = mu + sigma * math.sqrt(2.0) * float(erfinv(2.0 * tol - 1.0))
    low = math.floor(quantile)
    high = math.ceil(mu + (mu - quantile))
    return low, high


def measurement_ok(count, nr_shots, p, tol):
    low, high = get_bounds(nr_shots, p, tol)
    return low < count < high


def chi2_ok(counts, probs, nr_shots, alpha=1e-6):
    """Pearson chi-square goodness of fit; bins with expected < 5 are pooled."""
    from scipy.stats import chi2
    exp = [p * nr_shots for p in probs]
    obs = list(counts)
    big = [(o, e) for o, e in zip(obs, exp) if e >= 5]
    small_o = sum(o for o, e in zip(obs, exp) if e < 5)
    small_e = sum(e for e in exp if e < 5)
    if small_e > 0:
        big.append((small_o, small_e))
    stat = sum((o - e) ** 2 / e for o, e in big)
    dof = max(len(big) - 1, 1)
    return chi2.sf(stat, dof) > alpha, stat, dof
