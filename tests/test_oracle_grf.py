"""The CPU restatement of the lognormal random field against what defines it:
unit variance, zero mean and the squared-exponential correlation
exp(-r^2 / (2 l^2)) (white noise convolved with a Gaussian of std l/sqrt(2)).
CPU only."""
import numpy as np

import grf_oracle as G


def test_hash_uniform_range_and_determinism():
    u = G.unif(7, np.arange(100000, dtype=np.uint64))
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert np.array_equal(u, G.unif(7, np.arange(100000, dtype=np.uint64)))
    assert not np.array_equal(u, G.unif(8, np.arange(100000, dtype=np.uint64)))


def test_field_statistics_match_squared_exponential():
    n, ell = 96, 0.06
    h = 1.0 / n
    k = G.lognormal_cells((n, n, n), (h, h, h), 11, ell, 1.0)
    g = np.log(k).reshape(n, n, n)
    assert abs(g.mean()) < 0.1
    assert abs(g.var() - 1.0) < 0.15
    for lag in (2, 4, 6):
        r = lag * h
        c = np.mean(g[:, :, :-lag] * g[:, :, lag:]) / g.var()
        assert abs(c - np.exp(-r * r / (2 * ell * ell))) < 0.08, (lag, c)
