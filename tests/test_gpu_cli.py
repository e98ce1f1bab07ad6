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


@pytest.mark.parametrize("plain", [False, True])
def test_single_end_summary_and_status(data, plain):
    d, reads, filters = data
    extra = ["--plain"] if plain else []
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
