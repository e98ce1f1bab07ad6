"""The oracle's subtract-aware evalRead and progressive filter growth
(oracle/bbt_oracle.cpp: orc_eval_read_sub, orc_progressive), pinned by the
independent SeqEval model fed contains() answers of both filters, closed
forms, and the sequential recruitment order of BloomFilterGenerator's
progressive loop (BloomFilterGenerator.cpp:117-274)."""
import numpy as np
import pytest

import oracle_lib as O
import seqeval_model as M

K, H = 25, 3


def rand_seq(rng, n, alphabet="ACGT", p=None):
    return "".join(rng.choice(list(alphabet), n, p=p))


def windows(filt, sub, seq):
    pos, hv = O.nthash(seq, K, max(K, H))
    hits = [filt.contains_hashes(row[:H])[0] for row in hv]
    subs = [sub is not None and sub.contains_hashes(row[:H])[0] for row in hv]
    return [int(p) for p in pos], hits, subs


@pytest.mark.parametrize("scorer", [O.SIMPLE, O.LENGTH])
def test_eval_read_sub_matches_model(scorer):
    rng = np.random.default_rng(5)
    g = rand_seq(rng, 20_000)
    f = O.Filter(1 << 16, H, K)
    f.insert_seq(g[:12_000])
    sub = O.Filter(1 << 15, H, K)
    sub.insert_seq(g[6_000:9_000])
    agree = passed = 0
    for t in range(400):
        L = int(rng.integers(20, 300))
        a = int(rng.integers(0, len(g) - L))
        s = g[a:a + L] if t % 2 else rand_seq(rng, L, "ACGTN", [.24, .24, .24, .24, .04])
        if t % 5 == 0:
            s = s[:L // 2] + rand_seq(rng, L - L // 2)
        thr = [0.15, 0.5, 0.0][t % 3] if scorer == O.SIMPLE else [30, 60, 100][t % 3]
        p = O.params(scorer=scorer, threshold=thr)
        pos, hits, subs = windows(f, sub, s)
        if scorer == O.SIMPLE:
            want = M.simple(len(s), K, pos, hits, thr, sub=subs)[0]
        else:
            want = M.min_match_len(len(s), K, pos, hits, int(round(thr)), sub=subs)[0]
        got = O.eval_read_sub(f, sub, s, p)
        assert got == want, (t, L, thr)
        assert O.eval_read_sub(f, None, s, p) == O.eval_read(f, s, p)[0]
        agree += 1
        passed += got
    assert 0 < passed < agree


def test_subtract_equal_to_filter_scores_nothing():
    """With the filter itself as the subtract filter every hit scores zero: a
    read passes SIMPLE only at threshold 0 (the first hit meets 0 <= 0) and
    never reaches a positive match length."""
    rng = np.random.default_rng(6)
    g = rand_seq(rng, 5_000)
    f = O.Filter(1 << 15, H, K)
    f.insert_seq(g)
    s = g[100:250]
    assert O.eval_read_sub(f, None, s, O.params(threshold=0.15))
    assert not O.eval_read_sub(f, f, s, O.params(threshold=0.15))
    assert O.eval_read_sub(f, f, s, O.params(threshold=0.0))
    assert not O.eval_read_sub(f, f, s, O.params(scorer=O.LENGTH, threshold=30))


def prog_model(filt, pairs, score, expected, te, scorer=O.SIMPLE, inclusive=False,
               bait=None, bait_score=0.0, sub=None, iterations=1):
    """The progressive loop, one pair at a time, in Python."""
    p = O.params(scorer=scorer, threshold=score)
    pb = O.params(scorer=scorer, threshold=bait_score)
    tr = tg = 0

    def passes(s):
        return O.eval_read_sub(filt, sub, s, p) or (
            bait is not None and O.eval_read_sub(bait, sub, s, pb))

    def nk(s):
        return len(s) - K + 1 if len(s) > K else 0

    for _ in range(iterations):
        for s1, s2 in pairs:
            tr += 1
            if te >= expected:
                break
            if inclusive:
                tag = (nk(s1) > score and passes(s1)) or (nk(s2) > score and passes(s2))
            else:
                tag = passes(s1) and passes(s2)
            if tag:
                tg += 1
                te += filt.insert_seq(s1)[0] + filt.insert_seq(s2)[0]
    return te, tr, tg


def chain_pairs(rng, g, start, n, step=60, L=150):
    """Read pairs walking along g from start: each overlaps the previous."""
    out = []
    for i in range(n):
        a = start + i * step
        out.append((g[a:a + L], g[a + 20:a + 20 + L]))
    return out


@pytest.mark.parametrize("case", ["std", "inc", "bait", "sub", "length", "iter", "stop"])
def test_progressive_matches_model(case):
    rng = np.random.default_rng(11)
    g = rand_seq(rng, 40_000)
    seed = g[:3_000]
    pairs = chain_pairs(rng, g, 2_800, 150)
    noise = [(rand_seq(rng, 150), rand_seq(rng, 150)) for _ in range(60)]
    pairs = [x for pr in zip(pairs[:60], noise) for x in pr] + pairs[60:]
    kw = dict(score=0.15, expected=10 ** 9)
    if case == "inc":
        kw["inclusive"] = True
        pairs = [(s1, rand_seq(rng, 150)) for s1, _ in pairs]
    if case == "length":
        kw.update(scorer=O.LENGTH, score=60)
    if case == "iter":
        kw["iterations"] = 2
        pairs = pairs[::-1]
    if case == "stop":
        kw["expected"] = 3_000 + 2_000
    fs = []
    for _ in range(2):
        f = O.Filter(1 << 20, H, K)
        te = f.insert_seq(seed)[0]
        fs.append((f, te))
    if case == "bait":
        b = O.Filter(1 << 18, H, K)
        b.insert_seq(g[10_000:12_000])
        kw.update(bait=b, bait_score=0.15)
        pairs = pairs + chain_pairs(rng, g, 9_900, 20)
    if case == "sub":
        s = O.Filter(1 << 16, H, K)
        s.insert_seq(g[3_500:3_600])
        kw["sub"] = s
    (f1, te1), (f2, te2) = fs
    got = O.progressive(f1, pairs, total_entries=te1, **kw)
    want = prog_model(f2, pairs, te=te2, **kw)
    assert got == want
    assert np.array_equal(f1.bits, f2.bits)
    if case in ("std", "length", "inc"):
        assert got[2] >= 100  # the chain recruits pair after pair
    if case == "stop":
        assert got[0] >= kw["expected"] and got[1] < len(pairs)


def test_progressive_order_matters_and_threshold_stops():
    """Pairs along the genome recruit each other only in chain order, the
    one-thread order: reversed, nothing past the seed's reach is tagged; and
    a threshold already met reads one pair per iteration and tags nothing."""
    rng = np.random.default_rng(12)
    g = rand_seq(rng, 30_000)
    pairs = chain_pairs(rng, g, 1_900, 120)
    forward = O.Filter(1 << 20, H, K)
    te = forward.insert_seq(g[:2_000])[0]
    fwd = O.progressive(forward, pairs, 0.15, 10 ** 9, te)
    backward = O.Filter(1 << 20, H, K)
    backward.insert_seq(g[:2_000])
    bwd = O.progressive(backward, pairs[::-1], 0.15, 10 ** 9, te)
    assert fwd[2] == 120 and bwd[2] < 10
    done = O.Filter(1 << 20, H, K)
    te = done.insert_seq(g[:2_000])[0]
    assert O.progressive(done, pairs, 0.15, te, te, iterations=3) == (te, 3, 0)


def prog_model_single(filt, reads, score, expected, te, scorer=O.SIMPLE, sub=None,
                      iterations=1):
    """The single-end file-list loop (BloomFilterGenerator.cpp:1038-1170)."""
    p = O.params(scorer=scorer, threshold=score)
    tr = tg = 0
    if te >= expected:
        return te, tr, tg
    for _ in range(iterations):
        for s in reads:
            if te >= expected:
                break
            tr += 1
            nk = len(s) - K + 1 if len(s) > K else 0
            if nk > score and O.eval_read_sub(filt, sub, s, p):
                tg += 1
                te += filt.insert_seq(s)[0]
    return te, tr, tg


@pytest.mark.parametrize("expected", [10 ** 9, 4_500])
def test_progressive_single_end_matches_model(expected):
    rng = np.random.default_rng(13)
    g = rand_seq(rng, 30_000)
    reads = [g[1_900 + 50 * i:1_900 + 50 * i + 150] for i in range(200)]
    reads = [x for pr in zip(reads, [rand_seq(rng, 150) for _ in reads]) for x in pr]
    f1, f2 = O.Filter(1 << 20, H, K), O.Filter(1 << 20, H, K)
    te = f1.insert_seq(g[:2_000])[0]
    f2.insert_seq(g[:2_000])
    got = O.progressive(f1, reads, 0.15, expected, te, single=True, iterations=2)
    want = prog_model_single(f2, reads, 0.15, expected, te, iterations=2)
    assert got == want and np.array_equal(f1.bits, f2.bits)
    assert got[2] >= (200 if expected > 10 ** 6 else 1)
