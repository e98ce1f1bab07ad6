"""Pins the oracle's ntHash / spaced-seed iterators.

The upstream hashing code (btl_bloomfilter) is absent, so these pins come from
the mathematics of the published algorithm and from the reference's own
Tests/ntHashTests.cpp symmetry assertion:
  * rolling value == closed form XOR_j rol(seed(s_j), k-1-j) computed here in
    Python big integers, for every window of random sequences with Ns;
  * canonical(x) == canonical(revcomp(x));
  * iterator positions == brute-force list of N-free windows;
  * multi-hash value h_i = t ^ (t >> 27), t = h0 * (i ^ k*seed) mod 2^64;
  * spaced seeds: masked closed form and Tests/ntHashTests.cpp:32-40.
"""
import random

import numpy as np
import pytest

import oracle_lib as O

M64 = (1 << 64) - 1
SEED = {"A": 0x3c8bfbb395c60474, "C": 0x3193c18562a02b4c,
        "G": 0x20323ed082572324, "T": 0x295549f54be24456}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rol(v, s):
    s %= 64
    return ((v << s) | (v >> (64 - s))) & M64 if s else v


def closed_fwd(w):
    k = len(w)
    v = 0
    for j, c in enumerate(w.upper()):
        v ^= rol(SEED[c], k - 1 - j)
    return v


def revcomp(s):
    return "".join(COMP[c] for c in reversed(s.upper()))


def rand_seq(rng, n, pn=0.0):
    return "".join("N" if rng.random() < pn else rng.choice("ACGT") for _ in range(n))


def valid_windows(s, k):
    return [p for p in range(len(s) - k + 1)
            if all(c in "ACGTacgt" for c in s[p:p + k])]


def test_closed_form_matches_rolling():
    rng = random.Random(11)
    for trial in range(40):
        k = rng.choice([1, 2, 4, 9, 20, 25, 31, 32, 33, 63, 64, 65])
        s = rand_seq(rng, rng.randint(0, 200), pn=rng.choice([0, 0.01, 0.05, 0.2]))
        pos, hv = O.nthash(s, k, 1)
        assert list(pos) == valid_windows(s, k)
        for p, h0 in zip(pos, hv[:, 0]):
            w = s[p:p + k]
            f = closed_fwd(w)
            r = closed_fwd(revcomp(w))  # rc strand == forward hash of revcomp
            assert int(h0) == min(f, r)
            f2, r2 = O.nthash_window(w, k)
            assert (f2, r2) == (f, r)


def test_lowercase_and_other_bytes():
    s = "acgtACGTnNxacgtacgtRYacgtACGT"
    k = 4
    pos, hv = O.nthash(s, k, 1)
    assert list(pos) == valid_windows(s, k)
    up_pos, up_hv = O.nthash(s.upper(), k, 1)
    assert list(pos) == list(up_pos) and np.array_equal(hv, up_hv)


def test_canonical_reverse_complement_symmetry():
    rng = random.Random(5)
    for _ in range(30):
        k = rng.choice([5, 25, 31])
        s = rand_seq(rng, rng.randint(k, 120))
        _, a = O.nthash(s, k, 3)
        _, b = O.nthash(revcomp(s), k, 3)
        # window p of s is the reverse complement of window n-1-p of rc(s)
        assert np.array_equal(a, b[::-1])


def test_multihash_formula():
    rng = random.Random(3)
    mseed = 0x90b45d39fb6da1fa
    for _ in range(50):
        h0 = rng.getrandbits(64)
        k = rng.randint(1, 64)
        h = rng.randint(1, 12)
        out = O.multihash(h0, k, h)
        assert out[0] == h0
        for i in range(1, h):
            t = (h0 * (i ^ ((k * mseed) & M64))) & M64
            assert out[i] == t ^ (t >> 27)


def test_short_and_empty_sequences():
    for s in ["", "A", "ACG"]:
        pos, _ = O.nthash(s, 4, 2)
        assert len(pos) == 0
    pos, _ = O.nthash("ACGT", 4, 2)
    assert list(pos) == [0]
    pos, _ = O.nthash("NNNN", 4, 2)
    assert len(pos) == 0


def masked(w, seed):
    k = len(w)
    f = r = 0
    rcw = revcomp(w)
    for j in range(k):
        if seed[j] == "1":
            f ^= rol(SEED[w[j]], k - 1 - j)
    # the reverse strand sees the mirrored seed over revcomp(w)
    for j in range(k):
        if seed[k - 1 - j] == "1":
            r ^= rol(SEED[rcw[j]], k - 1 - j)
    return min(f, r)


def test_spaced_seed_closed_form():
    rng = random.Random(7)
    for _ in range(20):
        k = rng.choice([9, 15, 25])
        seeds = ["".join(rng.choice("01") for _ in range(k)) for _ in range(3)]
        s = rand_seq(rng, rng.randint(k, 90), pn=0.02)
        pos, hv = O.sthash(s, seeds, k)
        assert list(pos) == valid_windows(s, k)
        for p, row in zip(pos, hv):
            for j, sd in enumerate(seeds):
                assert int(row[j]) == masked(s[p:p + k], sd)


def test_ntHashTests_symmetry():
    # Tests/ntHashTests.cpp:22-40
    ss = ["111100001", "100001111"]
    fw, rv = "TCAAATCTAA", "TTAGATTTGA"
    _, a = O.sthash(fw, ss, 9)
    _, b = O.sthash(rv, ss, 9)
    assert a[0, 0] != b[0, 1]
    assert a[0, 1] != b[0, 0]
    assert a[1, 0] == b[0, 1]
    assert a[1, 1] == b[0, 0]


def test_all_ones_seed_equals_plain_hash():
    rng = random.Random(9)
    s = rand_seq(rng, 100, pn=0.03)
    pos, a = O.sthash(s, ["1" * 25], 25)
    pos2, b = O.nthash(s, 25, 1)
    assert np.array_equal(pos, pos2) and np.array_equal(a[:, 0], b[:, 0])
