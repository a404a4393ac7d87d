"""select_features' host step (_find_most_accessible_features) against the
reference's own loop form (snapatac2-python/python/snapatac2/preprocessing/
_basic.py:952-967): numpy's default argsort (so equal counts keep numpy's
order), the first position with a positive count found by a Python loop, then
the quantile cuts and the top total_features in descending order. Cases: no
zeros, leading zeros, all zeros (the loop ends on the last position), ties at
every cut, and an empty vector (the loop leaves its index unbound)."""
import numpy as np
import pytest

from snapatac2_amd.pp._basic import _find_most_accessible_features


def _reference_loop(feature_count, lo, hi, total):
    idx = np.argsort(feature_count)
    for i in range(idx.size):
        if feature_count[idx[i]] > 0:
            break
    idx = idx[i:]
    n = idx.size
    n_lower = int(lo * n)
    n_upper = int(hi * n)
    idx = idx[n_lower:n - n_upper]
    return idx[::-1][:total]


@pytest.mark.parametrize("seed", range(6))
def test_matches_reference_loop(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(200, 20000))
    c = rng.poisson(rng.uniform(0.3, 20.0), n).astype(np.float64)  # many ties and zeros
    if seed == 0:
        c[:] = 0.0
    if seed == 1:
        c[c == 0] = 1.0
    for lo, hi, total in ((0.005, 0.005, n // 3), (0.0, 0.0, n), (0.1, 0.2, 50)):
        assert np.array_equal(_find_most_accessible_features(c, lo, hi, total),
                              _reference_loop(c, lo, hi, total))


def test_empty_vector_fails_like_the_reference():
    with pytest.raises(UnboundLocalError):
        _reference_loop(np.zeros(0), 0.005, 0.005, 10)
    with pytest.raises(UnboundLocalError):
        _find_most_accessible_features(np.zeros(0), 0.005, 0.005, 10)
