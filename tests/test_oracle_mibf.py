"""Pins the oracle's MIBF restatement (oracle/bbt_mibf.cpp): build and query
against an independent Python model (mibf_model.py), order independence of
the build, brute-force rank, and the spaced-seed symmetry the reference's
Tests/ntHashTests.cpp asserts."""
import random

import numpy as np
import pytest

import mibf_model as M
import oracle_lib as O


def seeds_for(rng, span=15, weight=11):
    while True:
        s0 = ["1"] + [rng.choice("01") for _ in range(span - 2)] + ["1"]
        if s0.count("1") == weight:
            break
    s0 = "".join(s0)
    half = span // 2
    while True:
        left = ["1"] + [rng.choice("01") for _ in range(half - 1)]
        pal = left + ["0"] + left[::-1] if span % 2 else left + left[::-1]
        if pal.count("1") == weight or abs(pal.count("1") - weight) <= 1:
            break
    return [s0, "".join(pal), s0[::-1]]


def make_refs(rng, n=5, L=1500):
    refs = ["".join(rng.choice("ACGT") for _ in range(L)) for _ in range(n)]
    # shared segments between references
    for i in range(n):
        j = (i + 1) % n
        a = rng.randrange(0, L - 200)
        refs[i] = refs[i][:a] + refs[j][a:a + 150] + refs[i][a + 150:]
    return [(i + 1, s) for i, s in enumerate(refs)]


def test_build_matches_model_and_is_order_independent():
    rng = random.Random(1)
    seeds = seeds_for(rng)
    recs = make_refs(rng)
    mb = O.Mibf(recs, seeds)
    order, data = M.build(recs, seeds, mb.m)
    got_bits = np.flatnonzero(np.unpackbits(mb.bv.view(np.uint8), bitorder="little"))
    assert got_bits.tolist() == order
    assert mb.data[: mb.pop].tolist() == data
    shuffled = list(recs)
    rng.shuffle(shuffled)
    mb2 = O.Mibf(shuffled, seeds, m=mb.m)
    assert np.array_equal(mb2.data, mb.data)
    assert any(d & M.SAT for d in data), "no saturated slot in the test build"


def test_frame_probs_match_model():
    rng = random.Random(2)
    seeds = seeds_for(rng)
    recs = make_refs(rng)
    mb = O.Mibf(recs, seeds)
    for a in [0, 1, 2]:
        got = mb.frame_probs(a)
        want = M.frame_probs(mb.data[: mb.pop].tolist(), mb.pop, mb.m, 3, a, mb.nids)
        assert np.allclose(got, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("seed", range(3))
def test_query_matches_model(seed):
    rng = random.Random(10 + seed)
    seeds = seeds_for(rng)
    recs = make_refs(rng)
    mb = O.Mibf(recs, seeds, occupancy=rng.choice([0.3, 0.5]))
    order, data = mb.data, None
    order, data = M.build(recs, seeds, mb.m)
    reads = []
    for _ in range(60):
        if rng.random() < 0.7:
            _, ref = rng.choice(recs)
            a = rng.randrange(0, len(ref) - 150)
            r = list(ref[a:a + rng.choice([40, 100, 150])])
            for i in range(len(r)):
                if rng.random() < 0.02:
                    r[i] = rng.choice("ACGTN")
            reads.append("".join(r))
        else:
            reads.append("".join(rng.choice("ACGT") for _ in range(rng.choice([10, 120]))))
    for score, a, multi in [(1e-10, 2, 1.0), (1e-3, 1, 0.5), (1e-10, 0, 1.0)]:
        res, sat, ctr = mb.query(reads, score=score, allowed_miss=a, multi=multi)
        probs = M.frame_probs(data, mb.pop, mb.m, 3, a, mb.nids)
        for r, row, st in zip(reads, res, sat):
            want, wsat = M.query(order, data, mb.m, seeds, probs, mb.nids, score, a, multi,
                                 0, r)
            assert st == wsat
            assert row == [(c[0], c[1], c[2], c[3], c[4], c[5]) for c in want]
    assert sum(len(r) for r in res) > 0


def test_reads_classify_to_their_reference():
    rng = random.Random(5)
    seeds = seeds_for(rng, span=25, weight=20)
    recs = make_refs(rng, n=8, L=4000)
    mb = O.Mibf(recs, seeds)
    reads, truth = [], []
    for _ in range(100):
        rid, ref = rng.choice(recs)
        a = rng.randrange(0, len(ref) - 150)
        reads.append(ref[a:a + 150])
        truth.append(rid)
    res, _, _ = mb.query(reads)
    best = [row[0][0] if row else 0 for row in res]
    assert sum(b == t for b, t in zip(best, truth)) >= 90
