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
            assert row == [tuple(c) for c in want]
    assert sum(len(r) for r in res) > 0
    assert any(c[6] > 0 for row in res for c in row), "no non-saturated frame counted"


@pytest.mark.parametrize("best_agree", [False, True])
def test_pairs_and_best_agree_match_model(best_agree):
    """Pairs pool both mates' frames; frameCount follows the nested
    conditional of MIBFClassifier.hpp:804-805 (mate 2 only when mate 1 is
    shorter than k), which a short first mate exercises."""
    rng = random.Random(40 + best_agree)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=6, L=2000)
    mb = O.Mibf(recs, seeds)
    order, data = M.build(recs, seeds, mb.m)
    reads = []
    for i in range(80):
        _, ref = rng.choice(recs)
        a = rng.randrange(0, len(ref) - 300)
        l1 = rng.choice([10, 100, 150]) if i % 4 else 5
        l2 = rng.choice([8, 120, 150])
        reads.append(ref[a:a + l1])
        reads.append(ref[a + 150:a + 150 + l2])
    for score, a, multi in [(1e-10, 2, 1.0), (1e-3, 1, 0.3)]:
        res, sat, _ = mb.query(reads, score=score, allowed_miss=a, multi=multi, paired=True,
                               best_agree=best_agree)
        assert len(res) == len(reads) // 2
        probs = M.frame_probs(data, mb.pop, mb.m, 3, a, mb.nids)
        for j, (row, st) in enumerate(zip(res, sat)):
            want, wsat = M.query(order, data, mb.m, seeds, probs, mb.nids, score, a, multi, 0,
                                 reads[2 * j], reads[2 * j + 1], best_agree=best_agree)
            assert st == wsat and row == [tuple(c) for c in want], j
        assert sum(len(r) for r in res) > 0


def test_best_agree_drops_ambiguous_best_hit():
    """bestHitAgree empties exactly the results whose rank-first candidate
    trails another candidate in some statistic, and keeps all others."""
    rng = random.Random(42)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=6, L=2000)
    mb = O.Mibf(recs, seeds)
    reads = []
    for _ in range(200):
        _, ref = rng.choice(recs)
        a = rng.randrange(0, len(ref) - 150)
        reads.append(ref[a:a + rng.choice([40, 100, 150])])
    plain, _, _ = mb.query(reads, score=1e-3, allowed_miss=2, multi=0.0)
    agree, _, _ = mb.query(reads, score=1e-3, allowed_miss=2, multi=0.0, best_agree=True)
    dropped = 0
    for p, q in zip(plain, agree):
        ambiguous = bool(p) and any(c[f] > p[0][f] for c in p for f in range(1, 7))
        assert q == ([] if ambiguous else p)
        dropped += ambiguous
    assert dropped > 0


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


def test_kmer_mode_matches_model():
    """k-mer mode (no spaced seeds): h ntHash values per k-mer."""
    rng = random.Random(21)
    recs = make_refs(rng, n=4, L=1200)
    km = O.Kmer(25, 3)
    mb = O.Mibf(recs, km)
    order, data = M.build(recs, km, mb.m)
    got_bits = np.flatnonzero(np.unpackbits(mb.bv.view(np.uint8), bitorder="little"))
    assert got_bits.tolist() == order
    assert mb.data[: mb.pop].tolist() == data
    reads = []
    for _ in range(30):
        _, ref = rng.choice(recs)
        a = rng.randrange(0, len(ref) - 150)
        reads.append(ref[a:a + 150])
    res, sat, _ = mb.query(reads, score=1e-10, allowed_miss=0)
    probs = M.frame_probs(data, mb.pop, mb.m, 3, 0, mb.nids)
    for r, row, st in zip(reads, res, sat):
        want, wsat = M.query(order, data, mb.m, km, probs, mb.nids, 1e-10, 0, 1.0, 0, r)
        assert st == wsat and row == [tuple(c) for c in want]
    assert sum(len(r) for r in res) >= 25


def test_id_zero_slots_are_never_candidates():
    """Slots holding ID 0 (no build writes it; a loaded file may hold it)
    count toward their frame's distinct IDs and saturation but never become
    a candidate: oracle and model agree on planted ID-0 slots."""
    rng = random.Random(77)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=5, L=1500)
    mb = O.Mibf(recs, seeds)
    order, data = M.build(recs, seeds, mb.m)
    pick = random.Random(3).sample(range(mb.pop), mb.pop // 4)
    for j, i in enumerate(pick):
        v = 0 if j % 2 else 0x80000000
        mb.data[i] = v
        data[i] = v
    reads = []
    for _ in range(80):
        _, ref = rng.choice(recs)
        a = rng.randrange(0, len(ref) - 150)
        reads.append(ref[a:a + 150])
    for a in (0, 2):
        res, sat, _ = mb.query(reads, score=1e-10, allowed_miss=a, multi=1.0)
        probs = M.frame_probs(data, mb.pop, mb.m, 3, a, mb.nids)
        assert np.allclose(mb.frame_probs(a), probs, rtol=1e-12, atol=0)
        for r, row, st in zip(reads, res, sat):
            want, wsat = M.query(order, data, mb.m, seeds, probs, mb.nids, 1e-10, a, 1.0, 0, r)
            assert st == wsat
            assert row == [tuple(c) for c in want]
            assert all(c[0] >= 1 for c in row)
    assert sum(len(r) for r in res) > 0
