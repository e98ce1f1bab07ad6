"""The device walk (biobloom_amd/csrc/bbt_device.h) compiled for the CPU and
checked against the oracle: decisions, scores (bit-exact doubles), contains()
calls and bit lookups, for every scorer, plus the exact 64-bit modulo."""
import ctypes as C
import os
import random
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
NATIVE = os.path.join(HERE, "native")
LIB = os.path.join(NATIVE, "libwalker_host.so")


def build():
    src = os.path.join(NATIVE, "walker_host.cpp")
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(
            os.path.getmtime(src),
            os.path.getmtime(os.path.join(ROOT, "biobloom_amd", "csrc", "bbt_device.h"))):
        return
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "biobloom_amd", "csrc"), "-o", LIB, src],
                   check=True)


@pytest.fixture(scope="module")
def W():
    build()
    L = C.CDLL(LIB)
    p = C.c_void_p
    L.walker_host_eval.restype = C.c_int
    L.walker_host_eval.argtypes = [p, C.c_uint64, C.c_uint32, C.c_uint32, p, C.c_uint32,
                                   p, p, C.c_uint64, C.c_uint32, C.c_int, C.c_double,
                                   C.c_uint32, p, p, p, p]
    L.walker_host_wave_eval.restype = C.c_int
    L.walker_host_wave_eval.argtypes = L.walker_host_eval.argtypes
    L.walker_host_fastmod.argtypes = [C.c_uint64, p, C.c_uint64, p]
    L.walker_host_bases.argtypes = [p, p]
    L.walker_host_set_flat.argtypes = [C.c_uint]
    L.walker_host_set_ilv.argtypes = [C.c_uint]
    L.walker_host_roll.argtypes = [C.c_uint, p]
    return L


def padded(buf):
    n = (len(buf) + 15) // 8 * 8
    out = np.zeros(n, np.uint8)
    out[: len(buf)] = buf
    return out


def run_walker(W, f, reads, scorer, exh, thr, streak, binom=None, wave=False, dust=None):
    buf, off = O.pack_reads(reads)
    buf = padded(buf)
    n = len(reads)
    hits = np.zeros(n, np.uint8)
    scores = np.zeros(n, np.float64)
    ctr = np.zeros(3, np.uint64)
    b = binom if binom is not None else np.zeros(1, np.uint32)
    fn = W.walker_host_wave_eval if wave else W.walker_host_eval
    rc = fn(f.bits.ctypes.data, f.m, f.h, f.k, b.ctypes.data,
                            len(b) if binom is not None else 0, buf.ctypes.data,
                            off.ctypes.data, n, scorer, int(exh), thr, streak,
                            hits.ctypes.data, scores.ctypes.data, ctr.ctypes.data,
                            None if dust is None else dust.ctypes.data)
    assert rc == 0
    return hits, scores, ctr


def test_fastmod_exact(W):
    rng = random.Random(0)
    ds = [1, 2, 3, 7, 64, 127, 128, 1000, 27387648, 109550784, 624697664,
          16_430_000_000, (1 << 33) + 64, (1 << 40) - 64, (1 << 63) - 1, (1 << 63),
          (1 << 64) - 1]
    ns = [0, 1, 2, (1 << 64) - 1, (1 << 63), (1 << 32) - 1, 1 << 32]
    ns += [rng.getrandbits(64) for _ in range(2000)]
    arr = np.array(ns, np.uint64)
    for d in ds + [rng.getrandbits(rng.randint(1, 64)) | 1 for _ in range(200)]:
        out = np.zeros_like(arr)
        W.walker_host_fastmod(d, arr.ctypes.data, len(arr), out.ctypes.data)
        assert [int(x) for x in out] == [n % d for n in ns], d


def test_bases_every_byte(W):
    """Only A, C, G, T in either case are bases (ntHashIterator restarts at
    any other byte), coded A 0, C 1, T 2, G 3."""
    valid = np.zeros(256, np.uint8)
    code = np.zeros(256, np.uint8)
    W.walker_host_bases(valid.ctypes.data_as(C.c_void_p), code.ctypes.data_as(C.c_void_p))
    want = {ord(b): i for i, b in enumerate("ACTG")}
    want.update({ord(b.lower()): i for b, i in list(zip("ACTG", range(4)))})
    assert [c for c in range(256) if valid[c]] == sorted(want)
    assert all(int(code[c]) == v for c, v in want.items())


def rotl(v, s):
    s %= 64
    return ((v << s) | (v >> (64 - s))) & (2**64 - 1) if s else v


def test_roll_table_rolls(W):
    """Rolling with the table's terms equals ntHash computed from scratch on
    every window (forward: XOR of rotl(seed(b_j), k-1-j); reverse complement:
    XOR of rotl(seed(comp b_j), j)), for k below, at and above 64."""
    seeds = {"A": 0x3c8bfbb395c60474, "C": 0x3193c18562a02b4c,
             "G": 0x20323ed082572324, "T": 0x295549f54be24456}
    comp = {"A": "T", "C": "G", "G": "C", "T": "A"}
    code = {"A": 0, "C": 1, "T": 2, "G": 3}
    rng = random.Random(5)
    for k in (1, 5, 25, 31, 63, 64, 65, 96):
        t = np.zeros(40, np.uint64)
        W.walker_host_roll(k, t.ctypes.data_as(C.c_void_p))
        s = "".join(rng.choice("ACGT") for _ in range(k + 40))
        fwd = rc = 0
        for i, b in enumerate(s):
            e = (code[s[i - k]] << 2 | code[b]) if i >= k else 16 + code[b]
            fwd = rotl(fwd, 1) ^ int(t[2 * e])
            rc = rotl(rc ^ int(t[2 * e + 1]), 63)
            if i >= k - 1:
                w = s[i - k + 1:i + 1]
                f0 = r0 = 0
                for j, x in enumerate(w):
                    f0 ^= rotl(seeds[x], k - 1 - j)
                    r0 ^= rotl(seeds[comp[x]], j)
                assert (fwd, rc) == (f0, r0), (k, i)


def make_filter_and_reads(rng, k, h, m, nreads=150):
    genome = "".join(rng.choice("ACGT") for _ in range(3000))
    f = O.Filter(m, h, k)
    f.insert_seq(genome)
    f.genome = genome
    reads = []
    for _ in range(nreads):
        L = rng.choice([0, 1, k - 1, k, k + 1, 60, 150, 151, 400])
        if rng.random() < 0.5 and L < 2900:
            a = rng.randrange(0, 3000 - L)
            r = list(genome[a:a + L])
            for i in range(L):
                x = rng.random()
                if x < 0.02:
                    r[i] = rng.choice("ACGT")
                elif x < 0.03:
                    r[i] = "N"
            s = "".join(r)
        else:
            s = "".join(rng.choice("ACGTNacgt") if rng.random() < 0.05 else
                        rng.choice("ACGT") for _ in range(L))
        reads.append(s)
    return f, reads


@pytest.fixture(params=[0, 1, 4, 101, 104],
                ids=["advance", "step1", "step4", "lookahead1", "lookahead4"])
def flat(W, request):
    """lane walk form: advance()/consume(), or step()/take() with a call's
    bits fetched one or four at a time, or the lookahead order (the next
    window rolled and its bits fetched before the pending call's are taken)"""
    W.walker_host_set_flat(request.param)
    yield request.param
    W.walker_host_set_flat(0)


@pytest.mark.parametrize("wave", [False, True])
@pytest.mark.parametrize("k,h,m", [(25, 3, 1 << 16), (25, 4, 1 << 14), (11, 2, 4096),
                                   (5, 1, 1 << 12), (31, 3, 1 << 18), (32, 5, 1 << 17)])
def test_walker_matches_oracle(W, k, h, m, wave, flat):
    if wave and flat:
        pytest.skip("the lane walk form does not apply to whole-wave walks")
    rng = random.Random(k * 1000 + h)
    f, reads = make_filter_and_reads(rng, k, h, m)
    if wave:
        long = []
        for _ in range(12):
            a = rng.randrange(0, 1500)
            seg = list(f.genome[a:a + rng.randint(300, 1400)])
            for i in range(len(seg)):
                if rng.random() < 0.05:
                    seg[i] = rng.choice("ACGTN")
            long.append("".join(seg))
        reads = reads + long
    for scorer, thr in [(O.SIMPLE, 0.15), (O.SIMPLE, 0.0), (O.SIMPLE, 0.5),
                        (O.HARMONIC, 0.15), (O.LENGTH, 30.0), (O.BINOMIAL, 1e-5)]:
        for streak in [3, 0, 1]:
            p = O.params(scorer, thr, streak)
            binom = None
            if scorer == O.BINOMIAL:
                top = max(len(r) for r in reads)
                binom = np.array([1] + [O.lib().orc_calc_min_count(fl, f.fpr, thr)
                                        for fl in range(1, top + 1)], np.uint32)
            hits, _, ctr = run_walker(W, f, reads, scorer, False, thr, streak, binom, wave)
            want, probes, lookups = [], 0, 0
            for r in reads:
                ok, pr, lk = O.eval_read(f, r, p)
                want.append(int(ok))
                probes += pr
                lookups += lk
            assert hits.tolist() == want
            assert (int(ctr[0]), int(ctr[1])) == (probes, lookups)
            if wave:
                assert int(ctr[2]) >= lookups  # speculative probes only add
            if scorer == O.BINOMIAL:
                continue
            _, scores, ctr = run_walker(W, f, reads, scorer, True, thr, streak, None, wave)
            ws, probes = [], 0
            for r in reads:
                sc, pr, _ = O.eval_score(f, r, p)
                ws.append(sc)
                probes += pr
            got = scores.view(np.uint64).tolist()
            exp = np.array(ws, np.float64).view(np.uint64).tolist()
            assert got == exp
            assert int(ctr[0]) == probes


def dust_bitmap(reads, T, W):
    """bit off[read] + p set for the oracle's sdust intervals (clipped to the read)"""
    _, off = O.pack_reads(reads)
    words = np.zeros(int(off[-1]) // 32 + 2, np.uint32)
    for r, s in enumerate(reads):
        for a, b in O.sdust(s, T, W):
            for p in range(a, min(b, len(s))):
                g = int(off[r]) + p
                words[g >> 5] |= np.uint32(1 << (g & 31))
    return words


@pytest.mark.parametrize("wave", [False, True])
def test_walker_dust_matches_oracle(W, wave, flat):
    if wave and flat:
        pytest.skip("the lane walk form does not apply to whole-wave walks")
    """-D: the device walk (lane and whole-wave) with the sdust bitmap equals
    the oracle's dust walks, every scorer, counts included."""
    rng = random.Random(808 + wave)
    f, reads = make_filter_and_reads(rng, 25, 3, 1 << 15)
    planted = []
    for r in reads:
        a = rng.randrange(0, max(1, len(r)))
        planted.append(r[:a] + rng.choice(["A", "CA", "GGT", "ACGT"]) * rng.randint(6, 40) + r[a:])
    reads = reads + planted
    if wave:
        reads += [f.genome[:600] + "AC" * 100 + f.genome[600:1100]]
    for T, Wd in [(20, 64), (10, 32)]:
        bm = dust_bitmap(reads, T, Wd)
        for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.15), (O.LENGTH, 30.0)]:
            for streak in [3, 0]:
                p = O.params(scorer, thr, streak, True, T, Wd)
                hits, _, ctr = run_walker(W, f, reads, scorer, False, thr, streak, None, wave, bm)
                want, probes, lookups = [], 0, 0
                for r in reads:
                    ok, pr, lk = O.eval_read(f, r, p)
                    want.append(int(ok))
                    probes += pr
                    lookups += lk
                assert hits.tolist() == want
                assert (int(ctr[0]), int(ctr[1])) == (probes, lookups)
                _, scores, ctr = run_walker(W, f, reads, scorer, True, thr, streak, None, wave, bm)
                ws = [O.eval_score(f, r, p)[0] for r in reads]
                got = scores.view(np.uint64).tolist()
                assert got == np.array(ws, np.float64).view(np.uint64).tolist()


@pytest.mark.parametrize("wave", [False, True])
@pytest.mark.parametrize("m,h", [(8192, 2), (12288, 2), (6000, 1)])
def test_walker_sparse_hits_matches_oracle(W, m, h, wave, flat):
    if wave and flat:
        pytest.skip("the lane walk form does not apply to whole-wave walks")
    """Isolated hits among misses, the pattern the whole-wave walk replays in one
    step for SIMPLE and BINOMIAL when streakThreshold >= 2: a dense filter gives
    false positives at 20-50% of windows, and genome reads mutated every few
    bases give runs of single hits.  Decisions and probe/lookup counts equal the
    oracle's for streak thresholds 2..6 and thresholds that accept or reject
    on the first, middle or last hits of a read."""
    rng = random.Random(m + h)
    k = 21
    genome = "".join(rng.choice("ACGT") for _ in range(3000))
    f = O.Filter(m, h, k)
    f.insert_seq(genome)
    reads = []
    for i in range(90):
        L = rng.choice([30, 64, 100, 150, 300, 700, 1300])
        if i % 3 == 0:
            reads.append("".join(rng.choice("ACGT") for _ in range(L)))
            continue
        a = rng.randrange(0, 3000 - L)
        s = list(genome[a:a + L])
        step = rng.choice([k - 1, k, k + 2, 2 * k])
        for p in range(rng.randrange(step), L, step):
            s[p] = "ACGT"[("ACGT".index(s[p]) + 1) % 4]
            if rng.random() < 0.1:
                s[p] = "N"
        reads.append("".join(s))
    top = max(len(r) for r in reads)
    for scorer, thr in [(O.SIMPLE, 0.05), (O.SIMPLE, 0.15), (O.SIMPLE, 0.3),
                        (O.SIMPLE, 0.6), (O.BINOMIAL, 1e-2), (O.BINOMIAL, 1e-6)]:
        binom = None
        if scorer == O.BINOMIAL:
            binom = np.array([1] + [O.lib().orc_calc_min_count(fl, f.fpr, thr)
                                    for fl in range(1, top + 1)], np.uint32)
        for streak in [2, 3, 4, 6]:
            p = O.params(scorer, thr, streak)
            hits, _, ctr = run_walker(W, f, reads, scorer, False, thr, streak, binom, wave)
            want, probes, lookups = [], 0, 0
            for r in reads:
                ok, pr, lk = O.eval_read(f, r, p)
                want.append(int(ok))
                probes += pr
                lookups += lk
            assert hits.tolist() == want, (scorer, thr, streak)
            assert (int(ctr[0]), int(ctr[1])) == (probes, lookups), (scorer, thr, streak)
            assert 0 < sum(want) < len(want) or thr in (0.05, 0.6, 1e-2, 1e-6)


@pytest.mark.parametrize("sh", [1, 2, 4, 6])
def test_walker_interleaved_layout(W, sh):
    """A filter read through an interleaved copy (its bytes every 2^sh bytes,
    the other slots unrelated): every lane-walk form and the whole-wave walk
    equal the oracle, counts included."""
    rng = random.Random(40 + sh)
    f, reads = make_filter_and_reads(rng, 25, 3, 1 << 15)
    W.walker_host_set_ilv(sh)
    try:
        for mode in (0, 4, 104):
            W.walker_host_set_flat(mode)
            for wave in (False, True):
                for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.15), (O.LENGTH, 30.0)]:
                    p = O.params(scorer, thr, 3)
                    hits, _, ctr = run_walker(W, f, reads, scorer, False, thr, 3, None, wave)
                    want, probes, lookups = [], 0, 0
                    for r in reads:
                        ok, pr, lk = O.eval_read(f, r, p)
                        want.append(int(ok))
                        probes += pr
                        lookups += lk
                    assert hits.tolist() == want, (mode, wave, scorer)
                    assert (int(ctr[0]), int(ctr[1])) == (probes, lookups)
    finally:
        W.walker_host_set_ilv(0)
        W.walker_host_set_flat(0)


@pytest.mark.parametrize("G", [3, 4, 7, 10, 16])
def test_rows_walk_matches_oracle(W, G):
    """One read against every filter of an interleaved group (SIMPLE evalRead,
    the read hashed once, one group load per bit position, an automaton per
    filter): every filter's decision and the contains() calls and bit lookups
    summed over the filters equal the oracle's, streak thresholds 0-5, with
    and without -D."""
    rng = random.Random(70 + G)
    k, h, m = 21, 3, 1 << 16
    genome = "".join(rng.choice("ACGT") for _ in range(6000))
    filters = []
    for g in range(G):
        f = O.Filter(m, h, k)
        a = rng.randrange(0, 4000)
        f.insert_seq(genome[a:a + 1800])
        filters.append(f)
    reads = []
    for i in range(150):
        L = rng.choice([0, 10, 21, 22, 60, 150, 151, 300, 700])
        if i % 3 and L < 5000:
            a = rng.randrange(0, 6000 - L)
            r = list(genome[a:a + L])
            for q in range(L):
                x = rng.random()
                if x < 0.02:
                    r[q] = rng.choice("ACGT")
                elif x < 0.03:
                    r[q] = "N"
            reads.append("".join(r))
        else:
            reads.append("".join(rng.choice("ACGTN") if rng.random() < 0.03 else rng.choice("ACGT")
                                 for _ in range(L)))
    reads.append(genome[100:130] + "AC" * 40 + genome[200:300])
    W.walker_host_rows.restype = C.c_int
    W.walker_host_rows.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                   C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_uint32,
                                   C.c_void_p, C.c_void_p, C.c_void_p]
    buf, off = O.pack_reads(reads)
    buf = padded(buf)
    ptrs = (C.c_void_p * G)(*[f.bits.ctypes.data for f in filters])
    for dust in (None, (20, 32)):
        bm = dust_bitmap(reads, *dust) if dust else None
        for thr, streak in [(0.15, 3), (0.15, 0), (0.5, 1), (0.05, 2), (0.9, 5), (0.0, 3)]:
            hits = np.zeros(len(reads) * G, np.uint8)
            ctr = np.zeros(3, np.uint64)
            rc = W.walker_host_rows(ptrs, G, m, h, k, buf.ctypes.data, off.ctypes.data,
                                    len(reads), thr, streak, hits.ctypes.data, ctr.ctypes.data,
                                    None if bm is None else bm.ctypes.data)
            assert rc == 0
            p = O.params(O.SIMPLE, thr, streak, bool(dust), *(dust or (20, 64)))
            want, probes, lookups = [], 0, 0
            for r in reads:
                for f in filters:
                    ok, pr, lk = O.eval_read(f, r, p)
                    want.append(int(ok))
                    probes += pr
                    lookups += lk
            assert hits.tolist() == want, (thr, streak, dust)
            assert (int(ctr[0]), int(ctr[1])) == (probes, lookups), (thr, streak, dust)
            assert 0 < sum(want) < len(want) or thr in (0.0, 0.9)
