"""The front ends' FASTA/FASTQ readers (biobloom_amd/host/seq_ingest.cpp) on
the CPU: hand-derived kseq_read cases (Common/kseq.h:93-219) pin the
sequential and the block-parallel reader; on generated files (CRLF, empty
lines, multi-line records, '@' opening quality lines, garbage, truncation,
gzip) both equal an independent Python restatement, across thread counts and
block sizes that force many splits and unfinished records at block ends."""
import ctypes as C
import gzip
import os
import subprocess

import numpy as np
import pytest

import kseq_model

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
NATIVE = os.path.join(HERE, "native")
LIB = os.path.join(NATIVE, "libingest_host.so")
SRCS = [os.path.join(NATIVE, "ingest_host.cpp"),
        os.path.join(ROOT, "biobloom_amd", "host", "seq_ingest.cpp"),
        os.path.join(ROOT, "biobloom_amd", "host", "seq_ingest.h"),
        os.path.join(ROOT, "biobloom_amd", "host", "deflate_rt.h"),
        os.path.join(ROOT, "biobloom_amd", "host", "read_pack.h")]


def build():
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(map(os.path.getmtime, SRCS)):
        return
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread",
                    "-I", os.path.join(ROOT, "biobloom_amd", "host"), "-o", LIB,
                    SRCS[0], SRCS[1], "-lz", "-ldl"], check=True)


@pytest.fixture(scope="module")
def L():
    build()
    lib = C.CDLL(LIB)
    lib.ingest_dump.restype = C.c_long
    lib.ingest_dump.argtypes = [C.c_char_p, C.c_int, C.c_uint, C.c_ulong,
                                C.POINTER(C.c_void_p), C.POINTER(C.c_ulong),
                                C.POINTER(C.c_long), C.POINTER(C.c_ulong)]
    return lib


def dump(L, path, mode, threads=4, block=1 << 16):
    """(records, final status, shares re-parsed) from one pass over path."""
    ptr, ln, st, rp = C.c_void_p(), C.c_ulong(0), C.c_long(0), C.c_ulong(0)
    n = L.ingest_dump(str(path).encode(), mode, threads, block, C.byref(ptr), C.byref(ln),
                      C.byref(st), C.byref(rp))
    assert n >= 0
    raw = C.string_at(ptr, ln.value) if ln.value else b""
    recs = []
    for item in raw.split(b"\x1e")[:-1]:
        f = item.split(b"\x1f")
        recs.append((f[0], f[1], f[2], f[3], f[4] == b"q"))
    assert len(recs) == n
    return recs, st.value, rp.value


A, Q = False, True
CASES = [
    (b"@r1 c1\nACGT\n+\nIIII\n@r2\nAC\n+r2\nII\n",
     [(b"r1", b"c1", b"ACGT", b"IIII", Q), (b"r2", b"", b"AC", b"II", Q)], -1),
    (b">s1 desc here\nACG\n\nTT\n>s2\n\nGG\n",
     [(b"s1", b"desc here", b"ACGTT", b"", A), (b"s2", b"", b"GG", b"", A)], -1),
    (b"@r1 c\r\nACGT\r\n+\r\nIIII\r\n", [(b"r1", b"c", b"ACGT", b"IIII", Q)], -1),
    (b"@r1\r\nAC\r\n+\r\nII\r\n", [(b"r1", b"", b"AC", b"II", Q)], -1),
    (b"@r1\nACGT\n+\n@III\n@r2\nA\n+\nI\n",
     [(b"r1", b"", b"ACGT", b"@III", Q), (b"r2", b"", b"A", b"I", Q)], -1),
    (b"xx\nyy>s1\nAC\n", [(b"s1", b"", b"AC", b"", A)], -1),
    (b"@r0\nA\n+\nI\n@r1\nACGT\n+\nII", [(b"r0", b"", b"A", b"I", Q)], -2),
    (b"@r1\nACGT\n+", [], -2),
    (b"@r1\nACGTAC\n+r1 x\nIII\nIII\n", [(b"r1", b"", b"ACGTAC", b"IIIIII", Q)], -1),
    (b"@r1\nAC\n+\nIII\n@r2\nA\n+\nI\n", [], -2),
    (b">s1\nACGT", [(b"s1", b"", b"ACGT", b"", A)], -1),
    (b">s1\n>s2\nAC\n", [(b"s1", b"", b"", b"", A), (b"s2", b"", b"AC", b"", A)], -1),
    (b">s1\nAC>GT\n", [(b"s1", b"", b"AC>GT", b"", A)], -1),
    (b">s1\nAC\n@x\nGG\n", [(b"s1", b"", b"AC", b"", A), (b"x", b"", b"GG", b"", A)], -1),
    # a first sequence line of a lone '\r' keeps it: the sequence is 1 long
    (b">s1\n\r\nAC\n", [(b"s1", b"", b"\rAC", b"", A)], -1),
    (b"@r1\tsome comment\nA\n+\nI\n", [(b"r1", b"some comment", b"A", b"I", Q)], -1),
    (b"@r1 \r\nA\n+\nI\n", [(b"r1", b"\r", b"A", b"I", Q)], -1),
    (b"", [], -1),
    (b"no header at all\n", [], -1),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_hand_derived_cases(L, tmp_path, case):
    data, want, status = CASES[case]
    p = tmp_path / "in.fq"
    p.write_bytes(data)
    assert kseq_model.read_all(data) == (want, status)
    for mode in (0, 1, 2):
        recs, st, _ = dump(L, p, mode)
        assert (recs, st) == (want, status)
    pz = tmp_path / "in.fq.gz"
    pz.write_bytes(gzip.compress(data))
    for mode in (0, 1, 2):
        assert dump(L, pz, mode)[:2] == (want, status)


def random_file(rng, n, strict=False):
    """n records of mixed shapes (strict: 4-line FASTQ only)."""
    out = []
    for i in range(n):
        L = int(rng.integers(0, 400))
        seq = bytes(rng.choice(list(b"ACGTN"), L))
        eol = b"\r\n" if (not strict and rng.random() < 0.1) else b"\n"
        com = b" c%d" % i if rng.random() < 0.5 else b""
        if strict or rng.random() < 0.7:
            qual = bytes(rng.choice(list(b"@+IJ#"), L))
            if not strict and rng.random() < 0.2 and L > 2:
                cut = int(rng.integers(1, L))
                sq = seq[:cut] + eol + seq[cut:]
                qual_lines = qual[:cut] + eol + qual[cut:]
            else:
                sq, qual_lines = seq, qual
            out.append(b"@r%d" % i + com + eol + sq + eol + b"+" + eol + qual_lines + eol)
        else:
            lines = [seq[j:j + 60] for j in range(0, len(seq), 60)] or [b""]
            body = eol.join(lines) + eol
            if rng.random() < 0.2:
                body += eol
            out.append(b">s%d" % i + com + eol + body)
        if not strict and rng.random() < 0.02:
            out.append(b"junk line\n")
    return b"".join(out)


@pytest.mark.parametrize("seed", range(4))
def test_generated_files_equal_model(L, tmp_path, seed):
    rng = np.random.default_rng(seed)
    data = random_file(rng, 1500)
    if seed % 3 == 2:
        data = data[:int(len(data) * 0.97)]  # a record cut anywhere
    want, status = kseq_model.read_all(data)
    p = tmp_path / "in.txt"
    p.write_bytes(data)
    pz = tmp_path / "in.txt.gz"
    pz.write_bytes(gzip.compress(data))
    assert dump(L, p, 0)[:2] == (want, status)
    reparsed = 0
    for threads, block in [(1, 1 << 16), (4, 1 << 16), (8, 100_003), (16, 1 << 20)]:
        for path, mode in ((p, 1), (pz, 1), (p, 2), (pz, 2)):
            recs, st, rp = dump(L, path, mode, threads, block)
            assert (recs, st) == (want, status), (threads, block, path, mode)
            reparsed += rp
    # multi-line records and '@'-led quality lines defeat some guesses
    assert reparsed > 0


def test_pipe_input(L, tmp_path):
    """A named pipe (not mappable) carrying gzip or plain text is read through
    zlib and yields the same records."""
    import threading
    rng = np.random.default_rng(4)
    data = random_file(rng, 800)
    want, status = kseq_model.read_all(data)
    for payload in (data, gzip.compress(data)):
        fifo = tmp_path / "fifo"
        if fifo.exists():
            fifo.unlink()
        os.mkfifo(fifo)
        t = threading.Thread(target=lambda: fifo.write_bytes(payload))
        t.start()
        recs, st, _ = dump(L, fifo, 2, 4, 1 << 16)
        t.join()
        assert (recs, st) == (want, status)


def test_strict_fastq_splits_without_reparse(L, tmp_path):
    """Plain 4-line FASTQ: every guessed share start is a record start, so no
    share is parsed twice; records equal the sequential reader's."""
    rng = np.random.default_rng(9)
    data = random_file(rng, 40_000, strict=True)
    p = tmp_path / "big.fq"
    p.write_bytes(data)
    ref, st0, _ = dump(L, p, 0)
    assert st0 == -1 and len(ref) == 40_000
    for threads, block, mode in [(8, 1 << 18, 1), (16, 1 << 20, 2)]:
        recs, st, rp = dump(L, p, mode, threads, block)
        assert st == -1 and recs == ref
        assert rp == 0


def bgzf(data, chunk=65280, level=6):
    """data as BGZF: gzip members of at most `chunk` bytes of text, each with
    the 'BC' extra subfield holding the member size - 1, then the empty
    end-of-file member that bgzip writes."""
    import struct
    import zlib
    out = []
    for i in range(0, len(data), chunk):
        part = data[i:i + chunk]
        c = zlib.compressobj(level, zlib.DEFLATED, -15)
        payload = c.compress(part) + c.flush()
        out.append(b"\x1f\x8b\x08\x04\0\0\0\0\0\xff" + struct.pack("<H", 6) + b"BC" +
                   struct.pack("<HH", 2, len(payload) + 25) + payload +
                   struct.pack("<II", zlib.crc32(part), len(part)))
    out.append(b"\x1f\x8b\x08\x04\0\0\0\0\0\xff\x06\0BC\x02\0\x1b\0\x03\0\0\0\0\0\0\0\0\0")
    return b"".join(out)


@pytest.mark.parametrize("zlib_only", [False, True])
@pytest.mark.parametrize("chunk", [700, 9000, 65280])
def test_bgzf_equal_model(L, tmp_path, monkeypatch, chunk, zlib_only):
    """A BGZF file is inflated member by member on threads (with libdeflate
    where the system has it, and with zlib): records and status equal the
    sequential model for every reader mode, thread count and block size
    (blocks of 64 KiB hold parts of members)."""
    monkeypatch.setenv("BBT_NO_LIBDEFLATE", "1" if zlib_only else "0")
    rng = np.random.default_rng(chunk)
    data = random_file(rng, 1500)
    want, status = kseq_model.read_all(data)
    p = tmp_path / "in.bgz"
    p.write_bytes(bgzf(data, chunk))
    assert dump(L, p, 0)[:2] == (want, status)
    for threads, block in [(1, 1 << 16), (4, 1 << 16), (8, 100_003), (16, 1 << 20)]:
        for mode in (1, 2):
            assert dump(L, p, mode, threads, block)[:2] == (want, status), (threads, block, mode)


def test_bgzf_then_gzip_member(L, tmp_path):
    """BGZF members followed by an ordinary gzip member: the reader goes on
    with sequential zlib from that member, and the records are those of the
    concatenated text."""
    rng = np.random.default_rng(41)
    a, b = random_file(rng, 600, strict=True), random_file(rng, 600, strict=True)
    want, status = kseq_model.read_all(a + b)
    p = tmp_path / "mix.gz"
    p.write_bytes(bgzf(a, 5000)[:-28] + gzip.compress(b))
    for mode, threads in [(0, 1), (1, 4), (2, 8)]:
        assert dump(L, p, mode, threads)[:2] == (want, status)


@pytest.mark.parametrize("zlib_only", [False, True])
def test_bgzf_corrupt_member(L, tmp_path, monkeypatch, zlib_only):
    """A member whose contents fail their CRC ends the input with an error
    status; the records before it are the model's."""
    monkeypatch.setenv("BBT_NO_LIBDEFLATE", "1" if zlib_only else "0")
    rng = np.random.default_rng(43)
    data = random_file(rng, 1200, strict=True)
    want, _ = kseq_model.read_all(data)
    z = bytearray(bgzf(data, 4000))
    z[len(z) // 2] ^= 0x55
    p = tmp_path / "bad.bgz"
    p.write_bytes(bytes(z))
    for mode, threads in [(1, 1), (2, 4), (2, 16)]:
        recs, st, _ = dump(L, p, mode, threads)
        assert st < -1
        assert 0 < len(recs) < len(want)
        # the text ends inside a record: the last one may be cut short
        assert recs[:-1] == want[:len(recs) - 1]
        assert recs[-1][0] == want[len(recs) - 1][0]


def test_bgzf_cut_short_equals_sequential(L, tmp_path):
    """A BGZF file cut anywhere (inside a header, a payload or a trailer, or
    between members): the block-parallel reader gives the sequential
    reader's (gzread's) records and status."""
    rng = np.random.default_rng(47)
    data = random_file(rng, 400, strict=True)
    z = bgzf(data, 3000)
    p = tmp_path / "cut.bgz"
    for cut in sorted(set(int(x) for x in rng.integers(1, len(z), 40)) | {18, len(z) - 28}):
        p.write_bytes(z[:cut])
        want = dump(L, p, 0)[:2]
        for mode, threads in [(1, 4), (2, 16)]:
            assert dump(L, p, mode, threads)[:2] == want, (cut, mode)
    # bytes after the last member that do not start a gzip member: gzread
    # ignores them and the input ends normally, with every record
    p.write_bytes(z)
    full = dump(L, p, 0)[:2]
    assert full[1] == -1 and len(full[0]) == 400
    for pad in (b"\0" * 64, b"\x1f", b"garbage\n", b"\x1f\x00" + b"\0" * 30):
        for body in (z, z[:-28]):  # with and without bgzip's empty last member
            p.write_bytes(body + pad)
            want = dump(L, p, 0)[:2]
            assert want == full, pad
            for mode, threads in [(1, 4), (2, 16)]:
                assert dump(L, p, mode, threads)[:2] == want, (pad, mode)


def pack_model(reads):
    """The BBT_READS_2BIT layout (include/bbt_hip.h) restated with numpy:
    code (b >> 1) & 3 per base, four per byte; then one N-mask bit per base."""
    buf = np.frombuffer(b"".join(reads), np.uint8)
    n = len(buf)
    cb = (n + 31) // 32 * 8
    padded = np.full(cb * 4, ord("A"), np.uint8)
    padded[:n] = buf
    codes = (padded >> 1) & 3
    packed = (codes.reshape(-1, 4) << np.array([0, 2, 4, 6], np.uint8)).sum(1).astype(np.uint8)
    up = padded & 0xDF
    bad = ~np.isin(up, np.frombuffer(b"ACGT", np.uint8))
    mask = np.packbits(bad, bitorder="little")
    return cb, np.concatenate([packed, mask])


def unpack_model(packed, cb, n):
    """What the device's unpacking gives back: 'A' 'C' 'T' 'G' by code, 'N'
    for masked bases."""
    codes = np.unpackbits(packed[:cb], bitorder="little").reshape(-1, 2)
    codes = codes[:, 0] | (codes[:, 1] << 1)
    mask = np.unpackbits(packed[cb:cb + cb // 2], bitorder="little").astype(bool)
    out = np.frombuffer(b"ACTG", np.uint8)[codes]
    out[mask] = ord("N")
    return out[:n]


@pytest.mark.parametrize("simd", [0, 1], ids=["swar", "avx2"])
@pytest.mark.parametrize("threads", [1, 3, 16])
def test_pack_2bit_matches_model(L, threads, simd):
    """The front ends' packer (SWAR and AVX2 forms) equals the layout's
    restatement on reads of every length 0..80 (unaligned starts), upper and
    lower case, N and other bytes; unpacking gives the text with case folded
    and non-ACGT as N."""
    L.pack_set_simd.restype = C.c_int
    L.pack_set_simd.argtypes = [C.c_int]
    if simd and not L.pack_set_simd(-1):
        pytest.skip("no AVX2 on this host")
    L.pack_set_simd(simd)
    L.pack_2bit_reads.restype = C.c_ulong
    L.pack_2bit_reads.argtypes = [C.c_char_p, C.POINTER(C.c_ulong), C.c_ulong, C.c_void_p,
                                  C.c_uint]
    L.packed_bytes_for.restype = C.c_ulong
    rng = np.random.default_rng(threads)
    alphabet = np.frombuffer(b"ACGTacgtNnRY-.*\0", np.uint8)
    reads = []
    for i in range(3000):
        ln = i % 81 if i < 1000 else int(rng.integers(0, 400))
        p = np.full(16, 0.0)
        p[:8] = 0.118
        p[8:] = 0.056 / 8
        reads.append(bytes(rng.choice(alphabet, ln, p=p / p.sum())))
    off = np.zeros(len(reads) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in reads])
    seq = b"".join(reads)
    n = int(off[-1])
    out = np.zeros(L.packed_bytes_for(n), np.uint8)
    cb = L.pack_2bit_reads(seq, off.ctypes.data_as(C.POINTER(C.c_ulong)), len(reads),
                           out.ctypes.data, threads)
    want_cb, want = pack_model(reads)
    assert cb == want_cb and np.array_equal(out, want)
    text = np.frombuffer(seq, np.uint8)
    folded = np.where(np.isin(text & 0xDF, np.frombuffer(b"ACGT", np.uint8)), text & 0xDF,
                      ord("N")).astype(np.uint8)
    assert np.array_equal(unpack_model(out, cb, n), folded)
    L.pack_set_simd(-1)
