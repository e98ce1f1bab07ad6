"""Pins the oracle's sdust restatement (oracle/bbt_dust.cpp) against the
brute-force DUST definition (dust_model.py) on A/C/G/T sequences with planted
low-complexity stretches, for several thresholds and windows, plus the closed
forms: homopolymers and short tandem repeats are masked over their whole
length, high-complexity sequence is not."""
import random

import pytest

import dust_model as M
import oracle_lib as O


def planted(rng, n):
    s = [rng.choice("ACGT") for _ in range(n)]
    for _ in range(rng.randint(1, 3)):
        unit = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 4)))
        L = rng.randint(6, 40)
        a = rng.randrange(0, max(1, n - L))
        rep = (unit * (L // len(unit) + 1))[:L]
        for i, ch in enumerate(rep):
            if a + i < n and rng.random() > 0.1:
                s[a + i] = ch
    return "".join(s)


@pytest.mark.parametrize("T,W", [(20, 64), (20, 20), (10, 16), (40, 30), (20, 8)])
def test_sdust_matches_definition(T, W):
    rng = random.Random(T * 100 + W)
    for _ in range(25):
        s = planted(rng, rng.randint(20, 110))
        assert O.sdust(s, T, W) == M.intervals(s, T, W), s


def test_closed_forms():
    assert O.sdust("A" * 100) == [(0, 100)]
    assert O.sdust("AC" * 50) == [(0, 100)]
    assert O.sdust("ACG" * 30) == [(0, 90)]
    rng = random.Random(3)
    hi = "".join(rng.choice("ACGT") for _ in range(64))
    # a de Bruijn-like stretch: no triplet repeats, nothing to mask
    db = "AAACAAGAATACCACGACTAGCAGGAGTATCATGATTCCCGCCTCGGCGTCTGCTTGGGTGTTT"
    assert O.sdust(db) == []
    assert O.sdust("") == [] and O.sdust("ACG") == []
    # lower case is the same sequence
    assert O.sdust(("AC" * 50).lower()) == [(0, 100)]
    assert isinstance(O.sdust(hi), list)


def test_n_keeps_the_window():
    """sdust.c:156-160 resets only the word and the run length at an N, so the
    triplet window still holds the words before it and the next run's
    intervals are offset by their number: here 6 words (ACGTTGCA) shift the
    A-run at bases 9..48 to [15, 55).  With nothing before the N the run is
    masked where it lies."""
    assert O.sdust("ACGTTGCA" + "N" + "A" * 40) == [(15, 55)]
    assert O.sdust("N" + "A" * 40) == [(1, 41)]
    assert O.sdust("A" * 30 + "N" + "C" * 30)[0] == (0, 30)
