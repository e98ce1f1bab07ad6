"""Pins the oracle's Bloom filter: the reference's membership assertions
(Tests/BloomFilterTests.cpp:44-59, Tests/BloomFilterMakerTests.cpp:65-66
shape), brute-force contains with Python integer modulo, and the sizing
formulas of Common/BloomFilterInfo.h:57-76 against the figures the
reference's documentation and configs state."""
import math
import random

import numpy as np

import oracle_lib as O


def test_reference_membership_asserts():
    # Tests/BloomFilterTests.cpp:44-59: 1e9 bits, 5 hashes, k = 20
    f = O.Filter(1000000000, 5, 20)
    ins = ["ATCGGGTCATCAACCAATAT", "ATCGGGTCATCAACCAATAC",
           "ATCGGGTCATCAACCAATAG", "ATCGGGTCATCAACCAATAA"]
    for s in ins:
        _, hv = O.nthash(s, 20, 5)
        f.insert_hashes(hv[0])
    for s in ins:
        _, hv = O.nthash(s, 20, 5)
        assert f.contains_hashes(hv[0])[0]
    for s in ["ATCGGGTCATCAACCAATTA", "ATCGGGTCATCAACCAATTC"]:
        _, hv = O.nthash(s, 20, 5)
        assert not f.contains_hashes(hv[0])[0]


def brute_contains(bits, m, hv):
    looked = 0
    for h in hv:
        p = int(h) % m
        looked += 1
        if not (int(bits[p // 8]) >> (p % 8)) & 1:
            return False, looked
    return True, looked


def test_contains_brute_force_and_lookup_count():
    rng = random.Random(1)
    for m in [64, 128, 1000 * 64, 2 ** 20 + 64, 27387648]:
        f = O.Filter(m, 4, 25)
        f.bits[:] = np.frombuffer(bytes(rng.getrandbits(8) for _ in range(f.bits.size)),
                                  np.uint8)
        for _ in range(300):
            hv = np.array([rng.getrandbits(64) for _ in range(4)], np.uint64)
            assert f.contains_hashes(hv) == brute_contains(f.bits, m, hv)


def test_insert_sets_exactly_h_bits_mod_m():
    rng = random.Random(2)
    m = 1 << 20
    f = O.Filter(m, 3, 25)
    hv = np.array([rng.getrandbits(64) for _ in range(3)], np.uint64)
    assert f.insert_hashes(hv) == 0
    assert f.insert_hashes(hv) == 1  # insertAndCheck: already present
    want = {int(h) % m for h in hv}
    got = set(np.flatnonzero(np.unpackbits(f.bits, bitorder="little")).tolist())
    assert got == want


def test_insert_seq_counts():
    rng = random.Random(4)
    s = "".join(rng.choice("ACGT") for _ in range(2000))
    f = O.Filter(1 << 24, 3, 25)
    new, red = f.insert_seq(s)
    assert new + red == len(s) - 24
    new2, red2 = f.insert_seq(s)
    assert new2 == 0 and red2 == len(s) - 24
    pos, hv = O.nthash(s, 25, 3)
    assert all(f.contains_hashes(row)[0] for row in hv)


def test_sizing_matches_documented_geometry():
    L = O.lib()
    # BloomFilterInfo::calcOptimalHashNum: h = unsigned(-ln p / ln 2)
    assert L.orc_calc_optimal_hashnum(0.075) == 3
    assert L.orc_calc_optimal_hashnum(0.05) == 4
    assert L.orc_calc_optimal_hashnum(0.0078125) == 7
    # filter bits for the documented workloads (k = 25, one record each)
    assert L.orc_calc_optimal_size(5_000_000 - 24, 0.075, 3) == 27_387_648
    assert L.orc_calc_optimal_size(20_000_000 - 24, 0.075, 3) == 109_550_784
    assert L.orc_calc_optimal_size(100_000_000 - 24, 0.05, 4) == 624_697_664
    for n, p, h in [(1000, 0.075, 3), (12345, 0.01, 6)]:
        v = int(-n * h / math.log(1 - p ** (1 / h)))
        assert L.orc_calc_optimal_size(n, p, h) == v + 64 - v % 64
        assert L.orc_calc_optimal_size(n, p, h) % 64 == 0


def test_insert_subtract_pinned_per_window():
    """loadFilterSubtract restated: the result equals inserting, window by
    window, exactly the hashes the subtract filter does not contain (hashes
    and contains() pinned separately); an empty subtract filter changes
    nothing, one holding the same sequence removes every k-mer."""
    rng = np.random.default_rng(3)
    k, h = 25, 3
    seq = "".join(rng.choice(list("ACGTN"), 3000, p=[.248, .248, .248, .248, .008]))
    other = "".join(rng.choice(list("ACGT"), 2000))
    sub = O.Filter(4096 * 8, h, k)
    sub.insert_seq(seq[:1200] + "N" + other)
    got = O.Filter(8192 * 8, h, k)
    new, red, rem = got.insert_seq_subtract(seq, sub)
    want = O.Filter(8192 * 8, h, k)
    pos, hv = O.nthash(seq, k, h)
    w_new = w_red = w_rem = 0
    for row in hv:
        if sub.contains_hashes(row)[0]:
            w_rem += 1
            continue
        if want.insert_hashes(row):
            w_red += 1
        else:
            w_new += 1
    assert np.array_equal(got.bits, want.bits)
    assert (new, red, rem) == (w_new, w_red, w_rem)
    assert rem >= int((pos <= 1200 - k).sum()) > 0 and new + red + rem == len(hv)
    # an empty subtract filter is plain insertion
    plain, none = O.Filter(8192 * 8, h, k), O.Filter(64, h, k)
    a = O.Filter(8192 * 8, h, k)
    assert a.insert_seq_subtract(seq, none)[2] == 0
    plain.insert_seq(seq)
    assert np.array_equal(a.bits, plain.bits)
    # a subtract filter holding the sequence leaves the output empty
    full = O.Filter(8192 * 8, h, k)
    full.insert_seq(seq)
    z = O.Filter(4096 * 8, h, k)
    assert z.insert_seq_subtract(seq, full) == (0, 0, len(hv))
    assert z.popcount() == 0
