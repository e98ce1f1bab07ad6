"""Pins the oracle's SeqEval walks (Common/SeqEval.h) against:
  * Tests/SeqEvalTests.cpp (LENGTH mode known answers),
  * closed forms (all-hit, all-miss, reads shorter than k, t = 0),
  * an independent Python model (seqeval_model.py) fed the per-window
    contains() answers, comparing decision, contains() calls and lookups.
"""
import math
import random

import numpy as np
import pytest

import oracle_lib as O
import seqeval_model as M


def test_SeqEvalTests_length_kat():
    # Tests/SeqEvalTests.cpp:9-38 (passes in LENGTH mode, see SeqEval.h:295)
    f = O.Filter(8000, 2, 4)
    f.insert_seq("NNNNNAAAAA")
    assert O.eval_read(f, "AAAAANNNNN", O.params(O.LENGTH, 5))[0] is True
    assert O.eval_read(f, "AAAAANNNNN", O.params(O.LENGTH, 6))[0] is False


def full_filter(k, h, m=4096):
    f = O.Filter(m, h, k)
    f.bits[:] = 0xFF
    return f


def test_closed_forms_simple():
    rng = random.Random(0)
    k = 25
    for L in [25, 26, 60, 150, 151, 1000]:
        s = "".join(rng.choice("ACGT") for _ in range(L))
        n = L - k + 1
        for t in [0.0, 0.15, 0.5, 0.9, 1.0]:
            # every k-mer present: accept after the first j with 0.5+(j-1) >= t*n
            ok, probes, lookups = O.eval_read(full_filter(k, 3), s, O.params(O.SIMPLE, t))
            j = next((j for j in range(1, n + 1) if 0.5 + (j - 1) >= t * n), None)
            assert ok == (j is not None)
            assert probes == (j if j is not None else n)
            assert lookups == 3 * probes
            # no k-mer present: reject after floor((1-t)n) misses
            ok, probes, lookups = O.eval_read(O.Filter(4096, 3, k), s,
                                              O.params(O.SIMPLE, t))
            anti = math.floor((1 - t) * n)
            assert ok is False
            assert probes == (max(1, anti) if anti <= n else n)
            assert lookups == probes


def test_short_reads_are_rejected_without_probes():
    for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.15), (O.BINOMIAL, 1e-10),
                        (O.LENGTH, 30)]:
        for s in ["", "ACGT", "A" * 24]:
            ok, probes, _ = O.eval_read(full_filter(25, 3), s, O.params(scorer, thr))
            assert ok is False and probes == 0


def windows_and_hits(f, s):
    pos, hv = O.nthash(s, f.k, max(f.k, f.h))
    hits, looks = [], []
    for row in hv:
        r, lk = f.contains_hashes(row[: f.h])
        hits.append(r)
        looks.append(lk)
    return [int(p) for p in pos], hits, looks


def make_case(rng):
    k = rng.choice([5, 11, 25])
    h = rng.choice([1, 2, 3, 4])
    genome = "".join(rng.choice("ACGT") for _ in range(400))
    m = rng.choice([512, 1 << 13, 1 << 16, 1 << 18])
    f = O.Filter(m, h, k)
    f.insert_seq(genome)
    parts = []
    while sum(len(p) for p in parts) < rng.randint(k, 260):
        r = rng.random()
        if r < 0.5:
            a = rng.randrange(0, 350)
            parts.append(genome[a:a + rng.randint(1, 60)])
        elif r < 0.85:
            parts.append("".join(rng.choice("ACGT") for _ in range(rng.randint(1, 30))))
        else:
            parts.append("N" * rng.randint(1, 8))
    return f, "".join(parts)


@pytest.mark.parametrize("seed", range(6))
def test_model_agrees_bool_walks(seed):
    rng = random.Random(100 + seed)
    for _ in range(120):
        f, s = make_case(rng)
        pos, hits, looks = windows_and_hits(f, s)
        t = rng.choice([0.0, 0.05, 0.15, 0.3, 0.5, 0.8, 1.0])
        streak = rng.choice([0, 1, 2, 3, 5])
        for harmonic, scorer in [(False, O.SIMPLE), (True, O.HARMONIC)]:
            ok, probes, lookups = O.eval_read(f, s, O.params(scorer, t, streak))
            mok, w = M.simple(len(s), f.k, pos, hits, t, streak, harmonic)
            assert ok == mok
            assert probes == len(w.probes)
            assert lookups == sum(looks[i] for i in w.probes)
        mlen = rng.choice([2, f.k, f.k + 3, 40])
        ok, probes, lookups = O.eval_read(f, s, O.params(O.LENGTH, float(mlen)))
        mok, w = M.min_match_len(len(s), f.k, pos, hits, mlen)
        assert (ok, probes) == (mok, len(w.probes))
        min_fpr = rng.choice([1e-10, 1e-3, 0.05])
        if len(s) >= f.k:
            thres = int(O.lib().orc_calc_min_count(len(s) - f.k + 1, f.fpr, min_fpr))
        else:
            thres = 1
        ok, probes, lookups = O.eval_read(f, s, O.params(O.BINOMIAL, min_fpr, streak))
        mok, w = M.binomial(len(s), f.k, pos, hits, thres, streak)
        assert (ok, probes) == (mok, len(w.probes))


@pytest.mark.parametrize("seed", range(3))
def test_model_agrees_scores(seed):
    rng = random.Random(200 + seed)
    for _ in range(120):
        f, s = make_case(rng)
        pos, hits, looks = windows_and_hits(f, s)
        streak = rng.choice([0, 1, 3])
        for kind, scorer in [("simple", O.SIMPLE), ("harmonic", O.HARMONIC)]:
            sc, probes, _ = O.eval_score(f, s, O.params(scorer, 0.15, streak))
            msc, _, w = M.tiling_score(len(s), f.k, pos, hits, streak, kind)
            assert probes == len(w.probes)
            assert (math.isnan(sc) and math.isnan(msc)) or sc == msc
        sc, probes, _ = O.eval_score(f, s, O.params(O.LENGTH, 30.0))
        msc, w = M.min_match_len_score(len(s), f.k, pos, hits)
        assert sc == float(msc) and probes == len(w.probes)


def test_score_of_read_one_shorter_than_k_is_nan():
    # normalizeScore divides by seqLen - k + 1 == 0 (SeqEval.h:33-35)
    sc, probes, _ = O.eval_score(full_filter(25, 3), "A" * 24, O.params(O.SIMPLE))
    assert math.isnan(sc) and probes == 0
    sc, _, _ = O.eval_score(full_filter(25, 3), "A" * 10, O.params(O.SIMPLE))
    assert sc == 0.0


@pytest.mark.parametrize("seed", range(3))
def test_dust_walks_match_model(seed):
    """-D: a window starting in an sdust interval is a miss without a
    contains() call (SeqEval.h:53, :78 and the same test in every scorer)."""
    rng = random.Random(300 + seed)
    checked = 0
    for _ in range(120):
        f, s = make_case(rng)
        if rng.random() < 0.7:  # plant low-complexity stretches
            a = rng.randrange(0, max(1, len(s) - 1))
            s = s[:a] + rng.choice(["A", "AC", "GTT"]) * rng.randint(8, 25) + s[a:]
        T, W = rng.choice([(20, 64), (10, 32), (30, 64)])
        low = {p for a, b in O.sdust(s, T, W) for p in range(a, b)}
        pos, hits, looks = windows_and_hits(f, s)
        masked = [h and p not in low for p, h in zip(pos, hits)]
        t = rng.choice([0.05, 0.15, 0.5])
        streak = rng.choice([0, 1, 3])
        for harmonic, scorer in [(False, O.SIMPLE), (True, O.HARMONIC)]:
            ok, probes, lookups = O.eval_read(f, s, O.params(scorer, t, streak, True, T, W))
            mok, w = M.simple(len(s), f.k, pos, masked, t, streak, harmonic)
            called = [i for i in w.probes if pos[i] not in low]
            assert ok == mok
            assert probes == len(called)
            assert lookups == sum(looks[i] for i in called)
        mlen = rng.choice([f.k, f.k + 3, 40])
        ok, probes, _ = O.eval_read(f, s, O.params(O.LENGTH, float(mlen), 3, True, T, W))
        mok, w = M.min_match_len(len(s), f.k, pos, masked, mlen)
        assert (ok, probes) == (mok, len([i for i in w.probes if pos[i] not in low]))
        sc, probes, _ = O.eval_score(f, s, O.params(O.SIMPLE, 0.15, streak, True, T, W))
        msc, _, w = M.tiling_score(len(s), f.k, pos, masked, streak, "simple")
        assert (math.isnan(sc) and math.isnan(msc)) or sc == msc
        assert probes == len([i for i in w.probes if pos[i] not in low])
        checked += bool(low & set(pos))
    assert checked > 50
