"""The max_features selection rule (tests/feature_rules.py) gives the reference's
distributions: Fixed(k) a uniform k-subset (Floyd), Probability(p) independent keeps with
probability p and one uniform position when none is kept."""
import itertools
import math

import numpy as np

from tests import feature_rules as R


def test_fixed_subsets_are_uniform():
    n, k, trials = 5, 2, 3000
    counts = {}
    for t in range(trials):
        pos, scale = R.feature_select(R.FIXED, k, 0.0, n, 7, t, 1, 3)
        assert len(pos) == k and len(set(pos)) == k and max(pos) < n and scale == 2.5
        counts[tuple(pos)] = counts.get(tuple(pos), 0) + 1
    subsets = list(itertools.combinations(range(n), k))
    assert set(counts) <= set(subsets)
    exp = trials / len(subsets)
    chi2 = sum((counts.get(s, 0) - exp) ** 2 / exp for s in subsets)
    assert chi2 < 9 + 6 * math.sqrt(18)  # 9 dof
    assert R.feature_select(R.FIXED, 8, 0.0, 5, 7, 0, 1, 0) == ([0, 1, 2, 3, 4], 1.0)


def test_probability_keeps_are_bernoulli_with_at_least_one():
    n, p, trials = 6, 0.3, 3000
    kept = np.zeros(n)
    forced = 0
    for t in range(trials):
        pos, scale = R.feature_select(R.PROBABILITY, 0, p, n, 11, t, 2, 1)
        assert 1 <= len(pos) <= n
        if abs(scale - 1 / np.float32(p)) > 1e-6:
            forced += 1
            assert len(pos) == 1 and abs(scale - (1 - 0.3 ** 6) / 0.3) < 1e-5
        kept[pos] += 1
    # P(none) = 0.7^6 = 0.118; marginal keep = p + P(none)/n
    assert abs(forced / trials - 0.7 ** 6) < 4 * math.sqrt(0.118 * 0.882 / trials)
    marg = p + 0.7 ** 6 / n
    assert np.all(np.abs(kept / trials - marg) < 4 * math.sqrt(marg * (1 - marg) / trials))
