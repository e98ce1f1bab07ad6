"""Pins the oracle's binomial tail and quantile (SeqEval.h:199-216,
MIBFClassifier.hpp:930-939) against scipy.stats.binom, whose binomial
functions wrap Boost.Math, the library the reference calls.  Boost's
rounding of the quantile at exact ties is not pinned."""
import random

import numpy as np
import pytest
from scipy import stats

import oracle_lib as O


def test_pmf_and_sf_match_scipy():
    L = O.lib()
    rng = random.Random(0)
    for _ in range(200):
        n = rng.choice([1, 2, 10, 126, 1000, 1976, 49976])
        p = rng.choice([0.001, 0.01, 0.0237, 0.075, 0.1, 0.5, 0.9])
        x = rng.randint(0, n)
        want = stats.binom.pmf(x, n, p)
        got = L.orc_binom_pmf(x, n, p)
        assert got == pytest.approx(want, rel=1e-11, abs=1e-300)
        want = stats.binom.sf(x, n, p)
        got = L.orc_binom_sf(x, n, p)
        assert got == pytest.approx(want, rel=1e-9, abs=1e-15)


def test_quantile_definition_and_scipy():
    L = O.lib()
    rng = random.Random(1)
    for _ in range(300):
        n = rng.choice([1, 5, 126, 276, 1000, 5000])
        p = rng.choice([0.0237, 0.075, 0.05, 0.3])
        q = rng.choice([1e-10, 5e-13, 1e-3, 0.01, 0.2])
        x = int(L.orc_binom_isf_up(q, n, p))
        # smallest x with P(X > x) <= q
        assert stats.binom.sf(x, n, p) <= q * (1 + 1e-9)
        if x > 0:
            assert stats.binom.sf(x - 1, n, p) > q * (1 - 1e-9)
        assert x == int(stats.binom.isf(q, n, p))


def test_min_count_floor_of_one():
    L = O.lib()
    assert L.orc_calc_min_count(1, 1e-6, 0.5) == 1
    assert L.orc_calc_min_count(126, 0.075, 1e-10) == int(stats.binom.isf(1e-10, 126, 0.075))
