"""ctypes binding of the CPU oracle (oracle/libbbt_oracle.so).

Test infrastructure: only the tests, the smoke check and the benchmark's
CPU-baseline leg load this module, as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "libbbt_oracle.so")

SIMPLE, LENGTH, HARMONIC, BINOMIAL = 0, 1, 2, 3
STD, ORDERED, BESTHIT, SCORES = 0, 1, 2, 3


class OrcFilter(C.Structure):
    _fields_ = [("bits", C.c_void_p), ("m", C.c_uint64), ("h", C.c_uint32),
                ("k", C.c_uint32), ("fpr", C.c_double)]


class OrcParams(C.Structure):
    _fields_ = [("scorer", C.c_uint32), ("streak", C.c_uint32),
                ("threshold", C.c_double), ("dust", C.c_uint32), ("dust_t", C.c_uint32),
                ("dust_w", C.c_uint32), ("pad", C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = C.CDLL(LIB_PATH)
        u64p = C.POINTER(C.c_uint64)
        L.orc_nthash.restype = C.c_int64
        L.orc_nthash.argtypes = [C.c_char_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                 u64p, u64p, C.c_uint64]
        L.orc_nthash_window.argtypes = [C.c_char_p, C.c_uint32, u64p, u64p]
        L.orc_multihash.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, u64p]
        L.orc_sthash.restype = C.c_int64
        L.orc_sthash.argtypes = [C.c_char_p, C.c_uint64, C.c_char_p, C.c_uint32,
                                 C.c_uint32, u64p, u64p, C.c_uint64]
        L.orc_bf_contains.restype = C.c_int
        L.orc_bf_contains.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, u64p,
                                      C.POINTER(C.c_uint32)]
        L.orc_bf_insert.restype = C.c_int
        L.orc_bf_insert.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, u64p]
        L.orc_bf_insert_seq.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32,
                                        C.c_uint32, C.c_char_p, C.c_uint64, u64p,
                                        u64p]
        L.orc_bf_insert_seq_subtract.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32,
                                                 C.c_uint32, C.c_void_p, C.c_uint64,
                                                 C.c_char_p, C.c_uint64, u64p, u64p, u64p]
        L.orc_popcount.restype = C.c_uint64
        L.orc_popcount.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_calc_optimal_size.restype = C.c_uint64
        L.orc_calc_optimal_size.argtypes = [C.c_uint64, C.c_double, C.c_uint32]
        L.orc_calc_optimal_hashnum.restype = C.c_uint32
        L.orc_calc_optimal_hashnum.argtypes = [C.c_double]
        L.orc_binom_pmf.restype = C.c_double
        L.orc_binom_pmf.argtypes = [C.c_uint64, C.c_uint64, C.c_double]
        L.orc_binom_sf.restype = C.c_double
        L.orc_binom_sf.argtypes = [C.c_int64, C.c_uint64, C.c_double]
        L.orc_binom_isf_up.restype = C.c_uint64
        L.orc_binom_isf_up.argtypes = [C.c_double, C.c_uint64, C.c_double]
        L.orc_calc_min_count.restype = C.c_uint64
        L.orc_calc_min_count.argtypes = [C.c_uint64, C.c_double, C.c_double]
        L.orc_eval_read_sub.restype = C.c_int
        L.orc_eval_read_sub.argtypes = [C.POINTER(OrcFilter), C.POINTER(OrcFilter), C.c_char_p,
                                        C.c_uint64, C.POINTER(OrcParams)]
        L.orc_progressive.restype = C.c_int
        L.orc_progressive.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                      C.POINTER(OrcFilter), C.POINTER(OrcFilter), C.c_char_p,
                                      u64p, C.c_uint64, C.POINTER(OrcProgParams), u64p, u64p,
                                      u64p]
        L.orc_eval_read.restype = C.c_int
        L.orc_eval_read.argtypes = [C.POINTER(OrcFilter), C.c_char_p, C.c_uint64,
                                    C.POINTER(OrcParams), u64p, u64p]
        L.orc_eval_score.restype = C.c_double
        L.orc_eval_score.argtypes = [C.POINTER(OrcFilter), C.c_char_p, C.c_uint64,
                                     C.POINTER(OrcParams), u64p, u64p]
        L.orc_classify.restype = C.c_int
        L.orc_classify.argtypes = [C.POINTER(OrcFilter), C.c_uint32, C.c_void_p,
                                   u64p, C.c_uint64, C.c_int, C.c_int, C.c_uint32,
                                   C.POINTER(OrcParams), C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_void_p, u64p]
        L.orc_summary.restype = C.c_uint64
        L.orc_summary.argtypes = [C.POINTER(C.c_char_p), C.c_uint32, C.c_void_p,
                                  C.c_uint64, C.c_int, C.c_int, C.c_char_p,
                                  C.c_uint64]
        _lib = L
    return _lib


def _u64(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def nthash(seq, k, h):
    """(positions, hashes[n, h]) for every N-free window of seq."""
    if isinstance(seq, str):
        seq = seq.encode()
    cap = max(1, len(seq))
    pos = np.zeros(cap, np.uint64)
    hv = np.zeros(cap * h, np.uint64)
    n = lib().orc_nthash(seq, len(seq), k, h, _u64(pos), _u64(hv), cap)
    assert n >= 0
    return pos[:n].copy(), hv[: n * h].reshape(n, h).copy()


def nthash_window(win, k):
    if isinstance(win, str):
        win = win.encode()
    f, r = C.c_uint64(), C.c_uint64()
    lib().orc_nthash_window(win, k, C.byref(f), C.byref(r))
    return f.value, r.value


def multihash(h0, k, h):
    out = np.zeros(h, np.uint64)
    lib().orc_multihash(h0, k, h, _u64(out))
    return [int(x) for x in out]


def sthash(seq, seeds, k):
    if isinstance(seq, str):
        seq = seq.encode()
    cap = max(1, len(seq))
    pos = np.zeros(cap, np.uint64)
    hv = np.zeros(cap * len(seeds), np.uint64)
    n = lib().orc_sthash(seq, len(seq), "".join(seeds).encode(), len(seeds), k,
                         _u64(pos), _u64(hv), cap)
    assert n >= 0
    return pos[:n].copy(), hv[: n * len(seeds)].reshape(n, len(seeds)).copy()


class Filter:
    """A Bloom bit array owned by numpy, in the oracle's layout."""

    def __init__(self, m, h, k, bits=None):
        self.m, self.h, self.k = int(m), int(h), int(k)
        nbytes = (self.m + 7) // 8
        self.bits = np.zeros(nbytes, np.uint8) if bits is None else bits
        self._fpr = None

    def insert_seq(self, seq):
        if isinstance(seq, str):
            seq = seq.encode()
        new, red = C.c_uint64(0), C.c_uint64(0)
        lib().orc_bf_insert_seq(self.bits.ctypes.data, self.m, self.h, self.k, seq,
                                len(seq), C.byref(new), C.byref(red))
        self._fpr = None
        return new.value, red.value

    def insert_seq_subtract(self, seq, sub):
        """loadFilterSubtract: insert the k-mers of seq that filter sub does
        not contain -> (new, redundant, removed)."""
        if isinstance(seq, str):
            seq = seq.encode()
        new, red, rem = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        lib().orc_bf_insert_seq_subtract(self.bits.ctypes.data, self.m, self.h, self.k,
                                         sub.bits.ctypes.data, sub.m, seq, len(seq),
                                         C.byref(new), C.byref(red), C.byref(rem))
        self._fpr = None
        return new.value, red.value, rem.value

    def insert_hashes(self, hv):
        hv = np.ascontiguousarray(hv, np.uint64)
        self._fpr = None
        return lib().orc_bf_insert(self.bits.ctypes.data, self.m, self.h, _u64(hv))

    def contains_hashes(self, hv):
        hv = np.ascontiguousarray(hv, np.uint64)
        lk = C.c_uint32(0)
        r = lib().orc_bf_contains(self.bits.ctypes.data, self.m, self.h, _u64(hv),
                                  C.byref(lk))
        return bool(r), lk.value

    def popcount(self):
        return int(lib().orc_popcount(self.bits.ctypes.data, self.bits.size))

    @property
    def fpr(self):
        if self._fpr is None:
            self._fpr = (self.popcount() / self.m) ** self.h
        return self._fpr

    def struct(self):
        return OrcFilter(self.bits.ctypes.data, self.m, self.h, self.k, self.fpr)


def params(scorer=SIMPLE, threshold=0.15, streak=3, dust=False, dust_t=20, dust_w=64):
    return OrcParams(scorer, streak, threshold, int(dust), dust_t, dust_w, 0)


def eval_read(filt, seq, p):
    if isinstance(seq, str):
        seq = seq.encode()
    pr, lk = C.c_uint64(0), C.c_uint64(0)
    fs = filt.struct()
    r = lib().orc_eval_read(C.byref(fs), seq, len(seq), C.byref(p), C.byref(pr),
                            C.byref(lk))
    return bool(r), pr.value, lk.value


def eval_read_sub(filt, sub, seq, p):
    """evalRead with a subtract filter (None: no subtract)."""
    if isinstance(seq, str):
        seq = seq.encode()
    fs = filt.struct()
    ss = sub.struct() if sub is not None else None
    return bool(lib().orc_eval_read_sub(C.byref(fs), C.byref(ss) if ss is not None else None,
                                        seq, len(seq), C.byref(p)))


class OrcProgParams(C.Structure):
    _fields_ = [("score", C.c_double), ("bait", C.c_double), ("scorer", C.c_uint32),
                ("inclusive", C.c_uint32), ("streak", C.c_uint32), ("iterations", C.c_uint32),
                ("single", C.c_uint32), ("pad", C.c_uint32), ("expected", C.c_uint64)]


def progressive(filt, pairs, score, expected, total_entries, scorer=SIMPLE, inclusive=False,
                bait=None, bait_score=0.0, sub=None, streak=3, iterations=1, single=False):
    """Grows filt (in place) over pairs [(seq1, seq2), ...] (or single reads
    [seq, ...] when single) as the progressive builder does on one thread
    -> (total_entries, total_reads, tagged)."""
    buf, off = pack_reads(list(pairs) if single else [s for pr in pairs for s in pr])
    pp = OrcProgParams(score, bait_score, scorer, int(inclusive), streak, iterations,
                       int(single), 0, expected)
    te, tr, tg = (np.array([v], np.uint64) for v in (total_entries, 0, 0))
    bs = bait.struct() if bait is not None else None
    ss = sub.struct() if sub is not None else None
    rc = lib().orc_progressive(filt.bits.ctypes.data, filt.m, filt.h, filt.k,
                               C.byref(bs) if bs is not None else None,
                               C.byref(ss) if ss is not None else None,
                               buf.ctypes.data_as(C.c_char_p), _u64(off),
                               len(pairs), C.byref(pp), _u64(te), _u64(tr), _u64(tg))
    assert rc == 0
    filt._fpr = None
    return int(te[0]), int(tr[0]), int(tg[0])


def eval_score(filt, seq, p):
    if isinstance(seq, str):
        seq = seq.encode()
    pr, lk = C.c_uint64(0), C.c_uint64(0)
    fs = filt.struct()
    r = lib().orc_eval_score(C.byref(fs), seq, len(seq), C.byref(p), C.byref(pr),
                             C.byref(lk))
    return r, pr.value, lk.value


def kseq_seqs(data):
    """orc_kseq_seqs: the sequences kseq_read gives for FASTA/FASTQ text ->
    (uint8 buffer, uint64 offsets, kseq status after the last record)."""
    L = lib()
    L.orc_kseq_seqs.restype = C.c_int64
    L.orc_kseq_seqs.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                C.c_uint64, C.POINTER(C.c_int64)]
    data = bytes(data)
    cap = len(data) + 8
    buf = np.zeros(cap, np.uint8)
    off = np.zeros(len(data) // 2 + 2, np.uint64)
    st = C.c_int64(0)
    n = L.orc_kseq_seqs(data, len(data), buf.ctypes.data, cap, off.ctypes.data, len(off),
                        C.byref(st))
    assert n >= 0
    return buf[: int(off[n])], off[: n + 1].copy(), int(st.value)


def pack_reads(reads):
    """Concatenate reads (bytes/str) into (uint8 buffer, uint64 offsets)."""
    bs = [r.encode() if isinstance(r, str) else bytes(r) for r in reads]
    off = np.zeros(len(bs) + 1, np.uint64)
    off[1:] = np.cumsum([len(b) for b in bs], dtype=np.uint64)
    buf = np.frombuffer(b"".join(bs) + b"\0", np.uint8).copy()
    return buf, off


def classify(filters, buf, off, paired=False, inclusive=False, mode=STD, p=None,
             nthreads=0, want_scores=False):
    """Oracle classification of a packed batch -> (hits, scores, cls, counters)."""
    p = p or params()
    nreads = len(off) - 1
    nf = len(filters)
    arr = (OrcFilter * nf)(*[f.struct() for f in filters])
    hits = np.zeros(nreads * nf, np.uint8)
    scores = np.zeros(nreads * nf, np.float64) if want_scores else None
    nrec = nreads // 2 if paired else nreads
    cls = np.zeros(max(1, nrec), np.uint32)
    counters = np.zeros(2, np.uint64)
    rc = lib().orc_classify(arr, nf, buf.ctypes.data, _u64(off), nreads, int(paired),
                            int(inclusive), mode, C.byref(p), nthreads,
                            hits.ctypes.data,
                            scores.ctypes.data if scores is not None else None,
                            cls.ctypes.data, _u64(counters))
    assert rc == 0
    return (hits.reshape(nreads, nf), None if scores is None else scores.reshape(nreads, nf),
            cls[:nrec], counters)


def summary(filter_ids, hits, paired=False, inclusive=False):
    nreads, nf = hits.shape
    ids = (C.c_char_p * nf)(*[s.encode() for s in filter_ids])
    h = np.ascontiguousarray(hits, np.uint8)
    n = lib().orc_summary(ids, nf, h.ctypes.data, nreads, int(paired), int(inclusive),
                          None, 0)
    buf = C.create_string_buffer(int(n) + 1)
    lib().orc_summary(ids, nf, h.ctypes.data, nreads, int(paired), int(inclusive), buf,
                      n + 1)
    return buf.value.decode()


# ---------------------------------------------------------------- dust

def sdust(seq, T=20, W=64):
    """Merged low-complexity intervals [(start, finish)] of the oracle's sdust."""
    L = lib()
    L.orc_sdust.restype = C.c_int64
    L.orc_sdust.argtypes = [C.c_char_p, C.c_uint64, C.c_int, C.c_int, C.c_void_p,
                            C.c_void_p, C.c_uint64]
    b = seq.encode() if isinstance(seq, str) else bytes(seq)
    cap = len(b) + 1
    st = np.zeros(cap, np.uint32)
    fi = np.zeros(cap, np.uint32)
    n = L.orc_sdust(b, len(b), T, W, st.ctypes.data, fi.ctypes.data, cap)
    assert n >= 0
    return list(zip(st[:n].tolist(), fi[:n].tolist()))


# ---------------------------------------------------------------- MIBF

class MibfHit(C.Structure):
    _fields_ = [("id", C.c_uint32), ("count", C.c_uint32), ("nonsat", C.c_uint32),
                ("total", C.c_uint32), ("total_nonsat", C.c_uint32), ("solid", C.c_uint32),
                ("nonsat_frame", C.c_uint32)]


_mibf_ready = False


def _mibf_sigs():
    global _mibf_ready
    if _mibf_ready:
        return lib()
    L = lib()
    vp, u64p = C.c_void_p, C.POINTER(C.c_uint64)
    L.orc_mibf_calc_size.restype = C.c_uint64
    L.orc_mibf_calc_size.argtypes = [C.c_uint64, C.c_uint32, C.c_double]
    L.orc_mibf_set_bits.argtypes = [C.c_char_p, C.c_uint64, C.c_char_p, C.c_uint32,
                                    C.c_uint32, C.c_uint64, vp]
    L.orc_mibf_popcount.restype = C.c_uint64
    L.orc_mibf_popcount.argtypes = [vp, C.c_uint64]
    L.orc_mibf_assign.argtypes = [vp, vp, vp, C.c_uint64, C.c_char_p, C.c_uint32,
                                  C.c_uint32, C.c_uint64, vp, vp]
    L.orc_mibf_frame_probs.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.c_uint32, vp]
    L.orc_mibf_query.restype = C.c_int
    L.orc_mibf_query.argtypes = [vp, vp, C.c_uint64, C.c_char_p, C.c_uint32, C.c_uint32,
                                 vp, C.c_uint32, C.c_double, C.c_uint32, C.c_double,
                                 C.c_uint32, C.c_int, C.c_int, vp, vp, C.c_uint64,
                                 C.c_uint32, vp, vp, vp, vp]
    _mibf_ready = True
    return L


class Kmer:
    """k-mer mode of a MIBF: h ntHash values per k-mer instead of spaced seeds."""

    def __init__(self, k, h):
        self.k, self.h = k, h

    def __len__(self):
        return self.h


def seed_shape(seeds):
    """(number of hashes per frame, frame span, seed bytes or None)"""
    if isinstance(seeds, Kmer):
        return seeds.h, seeds.k, None
    return len(seeds), len(seeds[0]), "".join(seeds).encode()


class Mibf:
    """Oracle MIBF built from (id, sequence) records (ids 1..n); seeds is a
    list of '0'/'1' strings or a Kmer."""

    def __init__(self, records, seeds, occupancy=0.5, m=None):
        L = _mibf_sigs()
        self.seeds = seeds
        self.S, self.k, sd = seed_shape(seeds)
        entries = sum(max(0, len(s) - self.k + 1) for _, s in records)
        self.m = int(m or L.orc_mibf_calc_size(entries, self.S, occupancy))
        self.bv = np.zeros(self.m // 64, np.uint64)
        for _, s in records:
            b = s.encode() if isinstance(s, str) else bytes(s)
            L.orc_mibf_set_bits(b, len(b), sd, self.S, self.k, self.m, self.bv.ctypes.data)
        self.pop = int(L.orc_mibf_popcount(self.bv.ctypes.data, self.m))
        buf, off = pack_reads([s for _, s in records])
        ids = np.array([i for i, _ in records], np.uint32)
        self.data = np.zeros(max(1, self.pop), np.uint32)
        L.orc_mibf_assign(buf.ctypes.data, off.ctypes.data, ids.ctypes.data, len(records), sd,
                          self.S, self.k, self.m, self.bv.ctypes.data, self.data.ctypes.data)
        self.nids = int(ids.max()) if len(ids) else 0

    @classmethod
    def from_arrays(cls, bv, data, m, seeds, nids):
        """Wraps an existing bit vector (m/64 words) and ID array."""
        self = cls.__new__(cls)
        self.seeds = seeds
        self.S, self.k, _ = seed_shape(seeds)
        self.m = int(m)
        self.bv = np.ascontiguousarray(bv, np.uint64)
        self.pop = int(_mibf_sigs().orc_mibf_popcount(self.bv.ctypes.data, self.m))
        self.data = np.ascontiguousarray(data, np.uint32)
        self.nids = int(nids)
        return self

    def frame_probs(self, allowed_miss):
        out = np.zeros(self.nids + 1, np.float64)
        _mibf_sigs().orc_mibf_frame_probs(self.data.ctypes.data, self.pop, self.m, self.S,
                                          allowed_miss, self.nids, out.ctypes.data)
        return out

    def query(self, reads, score=1e-10, allowed_miss=None, multi=1.0, min_nonsat=0,
              max_hits=16, paired=False, best_agree=False):
        """Per record (read, or pair of reads 2j/2j+1): hit tuples
        (id, count, nonsat, total, total_nonsat, solid, nonsat_frame) in rank
        order; satCount per record; counters."""
        L = _mibf_sigs()
        a = self.S - 1 if allowed_miss is None else allowed_miss
        probs = self.frame_probs(a)
        buf, off = pack_reads(reads)
        n = len(reads) // 2 if paired else len(reads)
        hits = (MibfHit * max(1, n * max_hits))()
        nh = np.zeros(max(1, n), np.uint32)
        sat = np.zeros(max(1, n), np.uint32)
        ctr = np.zeros(3, np.uint64)
        L.orc_mibf_query(self.bv.ctypes.data, self.data.ctypes.data, self.m,
                         seed_shape(self.seeds)[2], self.S, self.k, probs.ctypes.data,
                         self.nids, score, a, multi, min_nonsat, int(paired), int(best_agree),
                         buf.ctypes.data, off.ctypes.data, len(reads), max_hits, hits,
                         nh.ctypes.data, sat.ctypes.data, ctr.ctypes.data)
        res = []
        for r in range(n):
            row = []
            for j in range(min(int(nh[r]), max_hits)):
                h = hits[r * max_hits + j]
                row.append((h.id, h.count, h.nonsat, h.total, h.total_nonsat, h.solid,
                            h.nonsat_frame))
            res.append(row)
        return res, sat[:n].tolist(), ctr
