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
