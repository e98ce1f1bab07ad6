"""The device sdust pass (biobloom_amd/csrc/bbt_dust_core.h, the -D pre-pass)
compiled for the CPU, against the oracle's sdust (oracle/bbt_dust.cpp, which
follows Common/sdust.c:75-165): the device-memory kernel's list capacity
always, and the LDS kernel's capacity whenever it does not hand the read on."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def L():
    path = os.path.join(HERE, "native", "libdust_host.so")
    if not os.path.exists(path):
        pytest.skip("tests/native/libdust_host.so not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.dust_host.restype = C.c_int
    lib.dust_host.argtypes = [C.c_char_p, C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_void_p]
    lib.dust_host_fast_pcap.restype = C.c_uint32
    lib.dust_host_fast_max_w.restype = C.c_int
    return lib


def device_mask(L, seq, T, W, pcap):
    m = np.zeros(len(seq), np.uint8)
    ok = L.dust_host(seq, len(seq), T, W, pcap, m.ctypes.data)
    return bool(ok), m


def oracle_mask(seq, T, W):
    m = np.zeros(len(seq), np.uint8)
    for a, b in O.sdust(seq, T, W):
        m[a:min(b, len(seq))] = 1
    return m


def sequences(seed, n):
    """Random reads with planted short-unit repeats, point changes and Ns."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    out = []
    for _ in range(n):
        s = bytearray(acgt[rng.integers(0, 4, int(rng.integers(20, 400)))].tobytes())
        for _ in range(int(rng.integers(0, 4))):
            a = int(rng.integers(0, len(s)))
            unit = np.frombuffer(b"ACGTN", np.uint8)[rng.integers(0, 5, int(rng.integers(1, 5)))]
            s[a:a] = unit.tobytes() * int(rng.integers(3, 30))
        for _ in range(int(rng.integers(0, 5))):
            s[int(rng.integers(0, len(s)))] = int(acgt[rng.integers(0, 4)])
        out.append(bytes(s))
    return out


@pytest.mark.parametrize("T,W", [(20, 64), (10, 32), (20, 100), (15, 7)])
def test_device_dust_equals_oracle(L, T, W):
    fast = W <= L.dust_host_fast_max_w()
    handed_on = 0
    for s in sequences(7 + W, 1500):
        ok, m = device_mask(L, s, T, W, 0)
        assert ok
        assert np.array_equal(m, oracle_mask(s, T, W)), (T, W, s)
        if fast:
            ok16, m16 = device_mask(L, s, T, W, L.dust_host_fast_pcap())
            if ok16:
                assert np.array_equal(m16, m), (T, W, s)
            else:
                handed_on += 1
    if fast:
        # most reads stay in LDS; with W >= 32 some planted runs are handed on
        assert handed_on < 1500 and (handed_on > 0 or W < 32)


def test_lds_capacity_cases(L):
    """A homopolymer, a dinucleotide and a trinucleotide repeat need far more
    list entries than an LDS lane keeps (the LDS kernel hands them on);
    uniform random reads need none past it."""
    T, W = 20, 64
    pc = L.dust_host_fast_pcap()
    for s in (b"A" * 300, b"AC" * 150, b"CAG" * 100):
        assert not device_mask(L, s, T, W, pc)[0]
        ok, m = device_mask(L, s, T, W, 0)
        assert ok and np.array_equal(m, oracle_mask(s, T, W)) and m.all()
    rng = np.random.default_rng(3)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    for _ in range(500):
        s = acgt[rng.integers(0, 4, 150)].tobytes()
        ok, m = device_mask(L, s, T, W, pc)
        assert ok and np.array_equal(m, oracle_mask(s, T, W))


def test_list_bound_after_n(L):
    """After an N the window keeps its words while the interval start stays at
    the run's first word, so the list can hold more entries than distinct
    (start, finish) pairs of one window; the device-memory kernel's capacity
    2 (W - 2)^2 + 64 still covers the longest list found on N-rich
    low-complexity sequences (about (W - 2)^2)."""
    rng = np.random.default_rng(5)
    for W in (5, 7, 10, 16, 32, 64):
        bound = 2 * (W - 2) ** 2
        worst = 0
        for it in range(300):
            alph = np.frombuffer([b"AN", b"ACN", b"AAAN", b"ACGTN", b"AAAAAAAN"][it % 5], np.uint8)
            s = alph[rng.integers(0, len(alph), int(rng.integers(10, 300)))].tobytes()
            T = int(rng.integers(1, 40))
            ok, m = device_mask(L, s, T, W, 0)
            assert ok and np.array_equal(m, oracle_mask(s, T, W)), (W, T, s)
            lo, hi = 0, bound + 64
            while hi - lo > 1:
                mid = (lo + hi) // 2
                if device_mask(L, s, T, W, mid)[0]:
                    hi = mid
                else:
                    lo = mid
            worst = max(worst, hi)
        assert worst <= bound
    # a read from the random-sequence tests that needs more than (W-1)(W-2)/2
    s = (b"CTACTTACTTGTTTATGTTATGCTTCAGCCATGCAATGGGGCCATCCGCAAAAATCCGCGGAGCCTAGC"
         + b"AC" * 62 + b"N" + b"TNGN" * 24 + b"AC" * 21 + b"AGAGATTTTTCATTTGTACTGCTG")
    need_more = not device_mask(L, s, 20, 64, 62 * 63 // 2 + 64)[0]
    ok, m = device_mask(L, s, 20, 64, 0)
    assert need_more and ok and np.array_equal(m, oracle_mask(s, 20, 64))
