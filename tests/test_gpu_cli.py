"""End to end through the restated CLIs: biobloommaker writes .bf/.txt from
FASTA, biobloomcategorizer classifies FASTQ (single and paired end, plain
and batched kernels, --fa/--fq partitions, --status), and the outputs are
compared byte for byte with the oracle on the same filters and reads."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAKER = os.path.join(ROOT, "biobloom_amd", "biobloommaker")
CAT = os.path.join(ROOT, "biobloom_amd", "biobloomcategorizer")


def write_fasta(path, recs):
    with open(path, "w") as f:
        for name, seq in recs:
            f.write(f">{name} some comment\n")
            for i in range(0, len(seq), 80):
                f.write(seq[i:i + 80] + "\n")


def write_fastq(path, recs):
    with open(path, "w") as f:
        for name, seq in recs:
            f.write(f"@{name} c\n{seq}\n+\n{'I' * len(seq)}\n")


def load_bf(path):
    raw = open(path, "rb").read()
    magic, hlen, _, m, h, k = struct.unpack("<8sIIQII", raw[:32])
    assert magic == b"BlOOMFXX" and hlen == 56
    f = O.Filter(m, h, k, np.frombuffer(raw[56:], np.uint8).copy())
    return f


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    d = tmp_path_factory.mktemp("cli")
    rng = np.random.default_rng(7)
    genomes = ["".join(rng.choice(list("ACGT"), 60_000)) for _ in range(3)]
    for i, g in enumerate(genomes):
        write_fasta(d / f"g{i}.fa", [(f"chr{i}a", g[:30000]), (f"chr{i}b", g[30000:])])
        print(f"[gpu] biobloommaker g{i}", flush=True)
        subprocess.run([MAKER, "-p", f"g{i}", "-o", str(d), "-k", "25", "-f", "0.075",
                        str(d / f"g{i}.fa")], check=True, timeout=120)
    reads = []
    for j in range(3000):
        if j % 3 == 0:
            s = "".join(rng.choice(list("ACGTN"), 150, p=[.249, .249, .249, .249, .004]))
        else:
            g = genomes[j % 3]
            a = rng.integers(0, len(g) - 150)
            s = g[a:a + 150]
        reads.append((f"r{j}", s))
    write_fastq(d / "r1.fq", reads[0::2])
    write_fastq(d / "r2.fq", reads[1::2])
    write_fastq(d / "all.fq", reads)
    filters = [load_bf(str(d / f"g{i}.bf")) for i in range(3)]
    return d, reads, filters


def run_cat(d, args, prefix):
    print(f"[gpu] biobloomcategorizer {' '.join(args)}", flush=True)
    bfs = " ".join(str(d / f"g{i}.bf") for i in range(3))
    subprocess.run([CAT, "-p", str(d / prefix), "-f", bfs] + args, check=True, timeout=120)
    return open(d / f"{prefix}_summary.tsv").read()


def test_maker_matches_oracle_filter(data):
    d, _, filters = data
    for i in range(3):
        txt = open(d / f"g{i}.txt").read()
        assert f"filter_id=g{i}" in txt and "kmer_size=25" in txt
        assert "desired_false_positve_rate=0.075" in txt
        seqs = open(d / f"g{i}.fa").read().split(">")[1:]
        of = O.Filter(filters[i].m, filters[i].h, filters[i].k)
        for rec in seqs:
            of.insert_seq("".join(rec.split("\n")[1:]))
        assert np.array_equal(of.bits, filters[i].bits)
        assert filters[i].m == O.lib().orc_calc_optimal_size(60_000 - 48, 0.075, 3)


@pytest.mark.parametrize("extra", [[], ["--plain"], ["--ascii_reads"]])
def test_single_end_summary_and_status(data, extra):
    """Reads go to the device 2-bit packed by default, as text with
    --ascii_reads; both kernel paths."""
    d, reads, filters = data
    got = run_cat(d, extra + ["--status", str(d / "all.fq")], "se")
    buf, off = O.pack_reads([s for _, s in reads])
    hits, _, cls, _ = O.classify(filters, buf, off)
    assert got == O.summary(["g0", "g1", "g2"], hits)
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    status = open(d / "se_readStatus.tsv").read().splitlines()
    assert status == [f"{n}\t{names[c]}" for (n, _), c in zip(reads, cls)]


def test_paired_end_and_inclusive(data):
    d, reads, filters = data
    buf, off = O.pack_reads([s for _, s in reads])
    for inc in [False, True]:
        got = run_cat(d, (["-i"] if inc else []) + ["-e", str(d / "r1.fq"), str(d / "r2.fq")],
                      "pe")
        hits, _, _, _ = O.classify(filters, buf, off, paired=True, inclusive=inc)
        assert got == O.summary(["g0", "g1", "g2"], hits, paired=True, inclusive=inc)


def test_partition_files(data):
    d, reads, filters = data
    run_cat(d, ["--fq", str(d / "all.fq")], "part")
    buf, off = O.pack_reads([s for _, s in reads])
    _, _, cls, _ = O.classify(filters, buf, off)
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    for ci, n in enumerate(names):
        want = "".join(f"@{rn} c\n{s}\n+\n{'I' * len(s)}\n"
                       for (rn, s), c in zip(reads, cls) if c == ci)
        assert open(d / f"part_{n}.fq").read() == want


def test_missing_filter_file_exits_nonzero(data, tmp_path):
    d, _, _ = data
    r = subprocess.run([CAT, "-p", str(tmp_path / "x"), "-f", str(tmp_path / "nope.bf"),
                        str(d / "all.fq")], capture_output=True, timeout=60)
    assert r.returncode != 0 and b"cannot be opened" in r.stderr


def test_dust_option(data, tmp_path):
    """-D -T -W reach the device sdust pre-pass: the status of every read
    equals the oracle's dust walks on reads with low-complexity stretches."""
    d, reads, filters = data
    rng = np.random.default_rng(31)
    dreads = []
    for j, (n, s) in enumerate(reads[:600]):
        a = int(rng.integers(0, len(s)))
        dreads.append((n, s[:a] + ["A", "CA", "GTT"][j % 3] * int(rng.integers(10, 40)) + s[a:]))
    write_fastq(tmp_path / "dust.fq", dreads)
    buf, off = O.pack_reads([s for _, s in dreads])
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    for T, W in [(20, 64), (10, 40)]:
        print(f"[gpu] biobloomcategorizer -D -T {T} -W {W}", flush=True)
        bfs = " ".join(str(d / f"g{i}.bf") for i in range(3))
        subprocess.run([CAT, "-p", str(tmp_path / "dust"), "-f", bfs, "-D", "-T", str(T),
                        "-W", str(W), "--status", str(tmp_path / "dust.fq")],
                       check=True, timeout=120)
        hits, _, cls, _ = O.classify(filters, buf, off, p=O.params(dust=True, dust_t=T,
                                                                    dust_w=W))
        status = open(tmp_path / "dust_readStatus.tsv").read().splitlines()
        assert status == [f"{n}\t{names[c]}" for (n, _), c in zip(dreads, cls)]
        assert open(tmp_path / "dust_summary.tsv").read() == O.summary(["g0", "g1", "g2"], hits)


def smart_pair_model(items):
    """The pairing rule of single-file paired input (BioBloomClassifier.cpp:
    497-551): a name ending in '1' or '2' loses its last two characters; the
    first record of a name waits, the next one with that name completes the
    pair (earlier record first); pairs come out in the order they complete and
    records left waiting are dropped."""
    waiting, pairs = {}, []
    for name, seq in items:
        key = name[:-2] if len(name) >= 2 and name[-1] in "12" else name
        if key in waiting:
            pairs.append((key, waiting.pop(key), seq))
        else:
            waiting[key] = seq
    return pairs


@pytest.mark.parametrize("batch", [100_000, 97])
def test_smart_pair_single_file(data, tmp_path, batch):
    """-e with one file: mates found by name anywhere in the file (either mate
    first), unpaired records dropped, names printed without the /1 /2 suffix;
    summary and --fq partitions equal the oracle on the completed pairs."""
    d, reads, filters = data
    rng = np.random.default_rng(11)
    items = []
    for i in range(1200):
        items.append((f"p{i}/1", reads[2 * i][1]))
        items.append((f"p{i}/2", reads[2 * i + 1][1]))
    for j in range(40):
        items.append((f"lone{j}/1", reads[2400 + j][1]))
    items.append(("x", reads[2500][1]))
    items.append(("x", reads[2501][1]))
    order = rng.permutation(len(items))
    items = [items[i] for i in order]
    write_fastq(tmp_path / "smart.fq", items)
    pairs = smart_pair_model(items)
    assert len(pairs) == 1201
    bfs = " ".join(str(d / f"g{i}.bf") for i in range(3))
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    for inc in [False, True]:
        args = [CAT, "-p", str(tmp_path / "sp"), "-f", bfs, "-e", "--fq",
                f"--batch_reads={batch}"] + (["-i"] if inc else []) + [str(tmp_path / "smart.fq")]
        print(f"[gpu] biobloomcategorizer smart pair inclusive={inc} batch={batch}", flush=True)
        subprocess.run(args, check=True, timeout=120)
        buf, off = O.pack_reads([s for p in pairs for s in p[1:]])
        hits, _, cls, _ = O.classify(filters, buf, off, paired=True, inclusive=inc)
        assert open(tmp_path / "sp_summary.tsv").read() == O.summary(
            ["g0", "g1", "g2"], hits, paired=True, inclusive=inc)
        for ci, n in enumerate(names):
            for m in (1, 2):
                want = "".join(f"@{key} c\n{p[m]}\n+\n{'I' * len(p[m])}\n"
                               for p, c in zip(pairs, cls) for key in [p[0]] if c == ci)
                assert open(tmp_path / f"sp_{n}_{m}.fq").read() == want


def test_maker_subtract(data, tmp_path):
    """biobloommaker -s: k-mers the subtract filter contains are left out;
    bits and the removed count equal the oracle's loadFilterSubtract, and a
    subtract filter with another hash number is refused."""
    d, _, filters = data
    recs = []
    for i in (0, 1):
        body = open(d / f"g{i}.fa").read().split(">")[1:]
        recs.append(("".join(body[0].split("\n")[1:])))
    write_fasta(tmp_path / "mix.fa", [("a", recs[0][:20000]), ("b", recs[1][:20000])])
    print("[gpu] biobloommaker -s g0.bf", flush=True)
    r = subprocess.run([MAKER, "-p", "mix", "-o", str(tmp_path), "-k", "25", "-f", "0.075",
                        "-s", str(d / "g0.bf"), str(tmp_path / "mix.fa")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = load_bf(str(tmp_path / "mix.bf"))
    want = O.Filter(got.m, got.h, got.k)
    removed = 0
    for s in (recs[0][:20000], recs[1][:20000]):
        removed += want.insert_seq_subtract(s, filters[0])[2]
    assert np.array_equal(got.bits, want.bits)
    assert f"Total Number of K-mers not added: {removed}" in r.stderr
    assert removed >= 20000 - 24
    r = subprocess.run([MAKER, "-p", "bad", "-o", str(tmp_path), "-k", "25", "-g", "5",
                        "-s", str(d / "g0.bf"), str(tmp_path / "mix.fa")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "hash number" in r.stderr


@pytest.mark.parametrize("fmt", ["fq.gz", "fa_crlf_multiline"])
def test_input_formats_through_parallel_ingest(data, tmp_path, fmt):
    """gzip FASTQ and CRLF multi-line FASTA go through the block-parallel
    reader (small blocks, 4 threads): status and summary equal the oracle on
    the same sequences."""
    import gzip
    d, reads, filters = data
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    if fmt == "fq.gz":
        path = tmp_path / "in.fq.gz"
        text = "".join(f"@{n} c\n{s}\n+\n{'I' * len(s)}\n" for n, s in reads)
        path.write_bytes(gzip.compress(text.encode()))
    else:
        path = tmp_path / "in.fa"
        parts = []
        for n, s in reads:
            body = "\r\n".join(s[i:i + 60] for i in range(0, len(s), 60))
            parts.append(f">{n} some comment\r\n{body}\r\n")
        path.write_bytes("".join(parts).encode())
    bfs = " ".join(str(d / f"g{i}.bf") for i in range(3))
    print(f"[gpu] biobloomcategorizer ingest {fmt}", flush=True)
    subprocess.run([CAT, "-p", str(tmp_path / "fmt"), "-f", bfs, "--status", "-t", "4",
                    "--ingest_block=65536", "--batch_reads=700", str(path)],
                   check=True, timeout=120)
    buf, off = O.pack_reads([s for _, s in reads])
    hits, _, cls, _ = O.classify(filters, buf, off)
    status = open(tmp_path / "fmt_readStatus.tsv").read().splitlines()
    assert status == [f"{n}\t{names[c]}" for (n, _), c in zip(reads, cls)]
    assert open(tmp_path / "fmt_summary.tsv").read() == O.summary(["g0", "g1", "g2"], hits)


def test_long_reads_in_small_blocks(data, tmp_path):
    """20 kbp reads parsed in 64 KiB blocks: each device batch stops at the
    blocks it may hold, so many short batches run; status and summary equal
    the oracle."""
    d, _, filters = data
    rng = np.random.default_rng(21)
    genome = open(d / "g1.fa").read().split(">")[1]
    g = "".join(genome.split("\n")[1:])
    reads = []
    for j in range(300):
        if j % 2:
            s = "".join(rng.choice(list("ACGT"), 20_000))
        else:
            a = int(rng.integers(0, len(g) - 20_000))
            s = g[a:a + 20_000]
        reads.append((f"L{j}", s))
    write_fastq(tmp_path / "long.fq", reads)
    bfs = " ".join(str(d / f"g{i}.bf") for i in range(3))
    print("[gpu] biobloomcategorizer long reads, 64 KiB blocks", flush=True)
    subprocess.run([CAT, "-p", str(tmp_path / "long"), "-f", bfs, "--status",
                    "--ingest_block=65536", str(tmp_path / "long.fq")], check=True, timeout=120)
    buf, off = O.pack_reads([s for _, s in reads])
    hits, _, cls, _ = O.classify(filters, buf, off)
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    status = open(tmp_path / "long_readStatus.tsv").read().splitlines()
    assert status == [f"{n}\t{names[c]}" for (n, _), c in zip(reads, cls)]
    assert open(tmp_path / "long_summary.tsv").read() == O.summary(["g0", "g1", "g2"], hits)


def test_gz_partitions_from_threads(data, tmp_path):
    """12k reads in batches that ramp up to 8192: batches of 4096 and more are
    formatted and gzip-compressed on threads, so each partition file holds
    several gzip members; the decompressed files, --status and the -d stdout
    equal the oracle's classes in input order.  Partitions that get no
    reads are valid, empty gzip files."""
    import gzip
    d, reads, filters = data
    rng = np.random.default_rng(5)
    many = [(f"m{i}", reads[int(rng.integers(0, len(reads)))][1]) for i in range(12000)]
    path = tmp_path / "many.fq"
    write_fastq(path, many)
    buf, off = O.pack_reads([s for _, s in many])
    _, _, cls, _ = O.classify(filters, buf, off)
    bfs = " ".join(str(d / f"g{i}.bf") for i in range(3))
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    print("[gpu] biobloomcategorizer --fq -g --status -d, 12k reads", flush=True)
    out = subprocess.run([CAT, "-p", str(tmp_path / "z"), "-f", bfs, "--fq", "-g", "--status",
                          "-d", "--batch_reads=8192", str(path)], check=True, timeout=120,
                         capture_output=True, text=True).stdout
    eof = b"\x1f\x8b\x08\x04\0\0\0\0\0\xff\x06\0BC\x02\0\x1b\0\x03\0\0\0\0\0\0\0\0\0"
    for ci, n in enumerate(names):
        want = "".join(f"@{rn} c\n{s}\n+\n{'I' * len(s)}\n"
                       for (rn, s), c in zip(many, cls) if c == ci)
        assert gzip.open(tmp_path / f"z_{n}.fq.gz", "rt").read() == want
        raw = (tmp_path / f"z_{n}.fq.gz").read_bytes()
        # BGZF: members with the 'BC' size subfield, then the empty end member
        assert raw[:4] == b"\x1f\x8b\x08\x04" and raw[12:14] == b"BC" and raw.endswith(eof)
    # zlib instead of libdeflate for the members: the same text
    env = dict(os.environ, BBT_NO_LIBDEFLATE="1")
    subprocess.run([CAT, "-p", str(tmp_path / "zz"), "-f", bfs, "--fq", "-g",
                    "--batch_reads=8192", str(path)], check=True, timeout=120, env=env)
    for n in names:
        assert gzip.open(tmp_path / f"zz_{n}.fq.gz").read() == \
            gzip.open(tmp_path / f"z_{n}.fq.gz").read()
    # a partition read back (BGZF input, members inflated on threads): every
    # read lands in its own class again
    g0 = [(rn, s) for (rn, s), c in zip(many, cls) if c == 0]
    subprocess.run([CAT, "-p", str(tmp_path / "back"), "-f", bfs, "--status", "-t", "8",
                    str(tmp_path / "z_g0.fq.gz")], check=True, timeout=120)
    back = open(tmp_path / "back_readStatus.tsv").read().splitlines()
    assert back == [f"{rn}\tg0" for rn, _ in g0]
    status = open(tmp_path / "z_readStatus.tsv").read().splitlines()
    assert status == [f"{rn}\t{names[c]}" for (rn, _), c in zip(many, cls)]
    assert out == "".join(f"@{rn} c\n{s}\n+\n{'I' * len(s)}\n"
                          for (rn, s), c in zip(many, cls) if c == 0)
    # random reads only: every partition but noMatch stays empty
    junk = [(f"j{i}", "".join(rng.choice(list("ACGT"), 150))) for i in range(50)]
    write_fastq(tmp_path / "junk.fq", junk)
    subprocess.run([CAT, "-p", str(tmp_path / "e"), "-f", bfs, "--fq", "-g",
                    str(tmp_path / "junk.fq")], check=True, timeout=120)
    for n in names:
        text = gzip.open(tmp_path / f"e_{n}.fq.gz", "rt").read()
        assert (text != "") == (n == "noMatch")
        if n != "noMatch":
            assert (tmp_path / f"e_{n}.fq.gz").read_bytes() == eof


def test_score_outputs_stdout_and_gz(data, tmp_path):
    """The printing paths (BioBloomClassifier.h:115-209): -w --fa puts every
    filter's score on multi-match headers, -b --fq the best score on every
    header, -d / -n print first-filter / other reads to stdout, -g gzips the
    partition files; scores and classes come from the oracle."""
    import gzip
    d, reads, filters = data
    sub = reads[:600]
    path = tmp_path / "in.fq"
    write_fastq(path, sub)
    buf, off = O.pack_reads([s for _, s in sub])
    bfs = " ".join(str(d / f"g{i}.bf") for i in range(3))
    names = ["g0", "g1", "g2", "noMatch", "multiMatch"]
    g = lambda v: "%g" % v  # iostream's default double format

    def run(args, prefix):
        print(f"[gpu] biobloomcategorizer {' '.join(args)}", flush=True)
        return subprocess.run([CAT, "-p", str(tmp_path / prefix), "-f", bfs] + args +
                              [str(path)], check=True, timeout=120, capture_output=True,
                              text=True).stdout

    # -w with --fa: SCORES mode
    run(["-w", "--fa"], "w")
    _, sc, cls, _ = O.classify(filters, buf, off, mode=O.SCORES, want_scores=True)
    for ci, n in enumerate(names):
        want = "".join(
            f">{rn} c" + ("".join(" " + g(x) for x in sc[r]) if c == 4 else "") + f"\n{s}\n"
            for r, ((rn, s), c) in enumerate(zip(sub, cls)) if c == ci)
        assert open(tmp_path / f"w_{n}.fa").read() == want
    # -b with --fq and -g: BESTHIT, gzip partitions
    run(["-b", "--fq", "-g"], "b")
    _, sc, cls, _ = O.classify(filters, buf, off, mode=O.BESTHIT, want_scores=True)
    for ci, n in enumerate(names):
        want = "".join(
            f"@{rn} c {g(max(0.0, float(sc[r].max())))}\n{s}\n+\n{'I' * len(s)}\n"
            for r, ((rn, s), c) in enumerate(zip(sub, cls)) if c == ci)
        assert gzip.open(tmp_path / f"b_{n}.fq.gz", "rt").read() == want
    # -d and -n: stdout
    _, _, cls, _ = O.classify(filters, buf, off)
    out = run(["-d"], "d")
    assert out == "".join(f"@{rn} c\n{s}\n+\n{'I' * len(s)}\n"
                          for (rn, s), c in zip(sub, cls) if c == 0)
    out = run(["-d", "-n"], "n")
    assert out == "".join(f"@{rn} c\n{s}\n+\n{'I' * len(s)}\n"
                          for (rn, s), c in zip(sub, cls) if c != 0)


def test_autotune_and_tuning_file(data, tmp_path):
    """--autotune times the knobs on the first batch and saves them to
    --tune_file; a later run takes them from the file.  Summaries equal the
    oracle's both times."""
    d, reads, filters = data
    tune = tmp_path / "tune.tsv"
    buf, off = O.pack_reads([s for _, s in reads])
    hits, _, _, _ = O.classify(filters, buf, off)
    want = O.summary(["g0", "g1", "g2"], hits)
    got = run_cat(d, ["--autotune", f"--tune_file={tune}", str(d / "all.fq")], "tuned")
    assert got == want
    lines = tune.read_text().splitlines()
    assert len(lines) >= 1 and all("\t" in ln for ln in lines)
    got = run_cat(d, [f"--tune_file={tune}", str(d / "all.fq")], "tuned2")
    assert got == want
