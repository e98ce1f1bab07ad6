"""Pins the oracle's mode dispatch and ResultsManager against small exhaustive
combinatorics written from BioBloomClassifier.cpp:1145-1237 and
ResultsManager.hpp:40-237, and the summary text format."""
import itertools
import random

import numpy as np

import oracle_lib as O


def ref_class_single(hits, nf):
    cls = nf
    for h in hits:
        cls = h if cls == nf else nf + 1
    return cls


def ref_class_pair(h1, h2, nf, inclusive):
    s = sorted(set(h1) | set(h2)) if inclusive else sorted(set(h1) & set(h2))
    return ref_class_single(s, nf), s


def ref_summary(ids, recs, nrec):
    nf = len(ids)
    above = [0] * nf
    unique = [0] * nf
    multi = none = 0
    for hl in recs:
        for h in hl:
            above[h] += 1
        c = ref_class_single(hl, nf)
        if c == nf:
            none += 1
        elif c == nf + 1:
            multi += 1
        else:
            unique[c] += 1
    g = lambda x: format(x, ".6g")
    out = ["filter_id\thits\tmisses\tshared\trate_hit\trate_miss\trate_shared"]
    for i, name in enumerate(ids):
        out.append("\t".join([name, str(above[i]), str(nrec - above[i]),
                              str(above[i] - unique[i]), g(above[i] / nrec),
                              g((nrec - above[i]) / nrec),
                              g((above[i] - unique[i]) / nrec)]))
    out.append("\t".join(["multiMatch", str(multi), str(nrec - multi), "0",
                          g(multi / nrec), g((nrec - multi) / nrec), "0"]))
    out.append("\t".join(["noMatch", str(none), str(nrec - none), "0", g(none / nrec),
                          g((nrec - none) / nrec), "0"]))
    return "\n".join(out) + "\n"


def test_summary_exhaustive_small():
    ids = ["a", "b", "c"]
    nf = 3
    rows = list(itertools.product([0, 1], repeat=nf))
    hits = np.array(rows, np.uint8)
    txt = O.summary(ids, hits)
    recs = [[f for f in range(nf) if r[f]] for r in rows]
    assert txt == ref_summary(ids, recs, len(rows))
    pairs = list(itertools.product(rows, rows))
    ph = np.array([x for pr in pairs for x in pr], np.uint8)
    for inc in [False, True]:
        txt = O.summary(ids, ph, paired=True, inclusive=inc)
        recs = [ref_class_pair([f for f in range(nf) if a[f]],
                               [f for f in range(nf) if b[f]], nf, inc)[1]
                for a, b in pairs]
        assert txt == ref_summary(ids, recs, len(pairs))


def build_filters(rng, nf, k=15, glen=3000):
    genomes = ["".join(rng.choice("ACGT") for _ in range(glen)) for _ in range(nf)]
    fs = []
    for g in genomes:
        f = O.Filter(1 << 16, 3, k)
        f.insert_seq(g)
        fs.append(f)
    return genomes, fs


def sample_reads(rng, genomes, n, L=80):
    reads = []
    for _ in range(n):
        if rng.random() < 0.6:
            g = rng.choice(genomes + [genomes[0] + genomes[1]])
            a = rng.randrange(0, len(g) - L)
            r = list(g[a:a + L])
            for i in range(L):
                if rng.random() < 0.03:
                    r[i] = rng.choice("ACGTN")
            reads.append("".join(r))
        else:
            reads.append("".join(rng.choice("ACGT") for _ in range(rng.randint(5, 120))))
    return reads


def test_modes_against_single_read_calls():
    rng = random.Random(3)
    genomes, fs = build_filters(rng, 4)
    reads = sample_reads(rng, genomes, 200)
    buf, off = O.pack_reads(reads)
    for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.2), (O.LENGTH, 30.0),
                        (O.BINOMIAL, 1e-5)]:
        p = O.params(scorer, thr)
        ev = [[O.eval_read(f, r, p)[0] for f in fs] for r in reads]
        sc = [[O.eval_score(f, r, p)[0] for f in fs] for r in reads]
        hits, _, cls, _ = O.classify(fs, buf, off, mode=O.STD, p=p)
        assert hits.tolist() == [[int(x) for x in row] for row in ev]
        assert cls.tolist() == [ref_class_single([i for i in range(4) if row[i]], 4)
                                for row in ev]
        hits, _, _, _ = O.classify(fs, buf, off, mode=O.ORDERED, p=p)
        for row, e in zip(hits.tolist(), ev):
            first = next((i for i in range(4) if e[i]), None)
            assert row == [int(i == first) for i in range(4)]
        hits, scores, _, _ = O.classify(fs, buf, off, mode=O.BESTHIT, p=p,
                                        want_scores=True)
        for row, srow, s in zip(hits.tolist(), scores.tolist(), sc):
            mx = 0.0
            best = []
            for i, v in enumerate(s):
                if mx < v:
                    mx, best = v, [i]
                elif mx == v:
                    best.append(i)
            want = [int(i in best and mx > 0) for i in range(4)]
            assert row == want
            assert [a == b or (a != a and b != b) for a, b in zip(srow, s)] == [True] * 4
        hits, scores, _, _ = O.classify(fs, buf, off, mode=O.SCORES, p=p,
                                        want_scores=True)
        assert hits.tolist() == [[int(x) for x in row] for row in ev]


def test_paired_modes():
    rng = random.Random(4)
    genomes, fs = build_filters(rng, 3)
    reads = sample_reads(rng, genomes, 300)
    buf, off = O.pack_reads(reads)
    p = O.params()
    ev = [[O.eval_read(f, r, p)[0] for f in fs] for r in reads]
    for inc in [False, True]:
        hits, _, cls, _ = O.classify(fs, buf, off, paired=True, inclusive=inc, p=p)
        for j in range(len(reads) // 2):
            a, b = ev[2 * j], ev[2 * j + 1]
            c, _ = ref_class_pair([i for i in range(3) if a[i]],
                                  [i for i in range(3) if b[i]], 3, inc)
            assert cls[j] == c
        hits, _, cls, _ = O.classify(fs, buf, off, paired=True, inclusive=inc,
                                     mode=O.ORDERED, p=p)
        for j in range(len(reads) // 2):
            a, b = ev[2 * j], ev[2 * j + 1]
            first = next((i for i in range(3)
                          if (a[i] or b[i] if inc else a[i] and b[i])), None)
            want = [int(i == first) for i in range(3)]
            assert hits[2 * j].tolist() == want and hits[2 * j + 1].tolist() == want
            assert cls[j] == (3 if first is None else first)


def test_thread_count_does_not_change_results():
    rng = random.Random(5)
    genomes, fs = build_filters(rng, 3)
    reads = sample_reads(rng, genomes, 500)
    buf, off = O.pack_reads(reads)
    a = O.classify(fs, buf, off, nthreads=1)
    b = O.classify(fs, buf, off, nthreads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])
    assert np.array_equal(a[3], b[3])
