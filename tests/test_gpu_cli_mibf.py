"""End to end through the restated MIBF CLIs: biobloommimaker builds
<prefix>.bf / <prefix>_ids.txt from FASTA (spaced seeds, k-mer mode, IDs by
file), biobloommicategorizer classifies FASTQ (single and paired end, TSV,
FASTA and FASTQ output, -i), and the filter contents, every output line and
the summary are compared with the oracle on the same records and reads."""
import os
import random
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from test_oracle_mibf import make_refs, seeds_for

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MIMAKER = os.path.join(ROOT, "biobloom_amd", "biobloommimaker")
MICAT = os.path.join(ROOT, "biobloom_amd", "biobloommicategorizer")


def write_fasta(path, recs):
    with open(path, "w") as f:
        for name, seq in recs:
            f.write(f">{name} a comment\n")
            for i in range(0, len(seq), 70):
                f.write(seq[i:i + 70] + "\n")


def write_fastq(path, recs):
    with open(path, "w") as f:
        for name, seq in recs:
            f.write(f"@{name} c{name}\n{seq}\n+\n{'I' * len(seq)}\n")


def load_device(B, path):
    mf = B.Mibf.load(path)
    info = mf.info()
    bv = np.zeros(info.m // 64, np.uint64)
    data = np.zeros(max(1, info.pop), np.uint32)
    mf.download(bv, data)
    return mf, info, bv, data


def fmt_hits(names, row):
    return "\t" + ";".join(f"{names[h[0]]},{h[6]},{h[5]},{h[1]},{h[2]},{h[4]},{h[3]}"
                           for h in row)


def expected_lines(names, reads, rows, sats, out="tsv", hit_only=False, paired=False):
    """formatOutStr (MIBFClassifier.hpp:675-736) for every read or pair."""
    lines = []
    step = 2 if paired else 1
    for r, (row, sat) in enumerate(zip(rows, sats)):
        if hit_only and not row:
            continue
        for name, seq in reads[r * step:(r + 1) * step]:
            head = f"{name} c{name}"
            if out == "tsv":
                if not row:
                    lines.append(f"*\t{head}\t{sat}\t*")
                else:
                    lines.append(f"{names[row[0][0]]}\t{head}\t{sat}" + fmt_hits(names, row))
            else:
                h = (">" if out == "fa" else "@") + head
                if row:
                    h += f"\t{names[row[0][0]]}" + fmt_hits(names, row)
                lines.append(h)
                lines.append(seq)
                if out == "fq":
                    lines.append("+")
                    lines.append("I" * len(seq))
    return lines


def expected_summary(names, rows, nrec, hit_only_paired=False):
    above = [0] * len(names)
    unique = [0] * len(names)
    multi = none = 0
    for row in rows:
        if hit_only_paired and not row:
            continue
        for h in row:
            above[h[0]] += 1
        if not row:
            none += 1
        elif len(row) == 1:
            unique[row[0][0]] += 1
        else:
            multi += 1
    g = lambda x: format(x, ".6g")
    out = ["filter_id\thits\tmisses\tshared\trate_hit\trate_miss\trate_shared"]
    for i, n in enumerate(names):
        out.append(f"{n}\t{above[i]}\t{nrec - above[i]}\t{above[i] - unique[i]}\t"
                   f"{g(above[i] / nrec)}\t{g((nrec - above[i]) / nrec)}\t"
                   f"{g((above[i] - unique[i]) / nrec)}")
    out.append(f"multiMatch\t{multi}\t{nrec - multi}\t0\t{g(multi / nrec)}\t"
               f"{g((nrec - multi) / nrec)}\t0")
    out.append(f"noMatch\t{none}\t{nrec - none}\t0\t{g(none / nrec)}\t"
               f"{g((nrec - none) / nrec)}\t0")
    return "\n".join(out) + "\n"


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from biobloom_amd import lib as B
    B.load()
    d = tmp_path_factory.mktemp("micli")
    rng = random.Random(77)
    seeds = seeds_for(rng, span=25, weight=20)
    recs = make_refs(rng, n=5, L=4000)
    names = [f"ref{i}" for i, _ in recs]
    write_fasta(d / "refs.fa", [(n, s) for n, (_, s) in zip(names, recs)])
    print("[gpu] biobloommimaker -S", flush=True)
    subprocess.run([MIMAKER, "-p", "mi", "-S", " ".join(seeds), str(d / "refs.fa")],
                   check=True, timeout=120, cwd=d)
    reads = []
    for j in range(600):
        if j % 4 == 0:
            s = "".join(rng.choice("ACGT") for _ in range(rng.choice([20, 150])))
        else:
            _, ref = rng.choice(recs)
            a = rng.randrange(0, len(ref) - 200)
            s = list(ref[a:a + rng.choice([60, 150, 200])])
            for i in range(len(s)):
                if rng.random() < 0.01:
                    s[i] = rng.choice("ACGTN")
            s = "".join(s)
        reads.append((f"q{j}", s))
    write_fastq(d / "all.fq", reads)
    write_fastq(d / "r1.fq", reads[0::2])
    write_fastq(d / "r2.fq", reads[1::2])
    return B, d, seeds, recs, names, reads


def run_cat(d, args, prefix):
    print(f"[gpu] biobloommicategorizer {' '.join(args)}", flush=True)
    r = subprocess.run([MICAT, "-p", str(d / prefix), "-f", str(d / "mi.bf")] + args,
                       check=True, timeout=120, capture_output=True, text=True)
    return r.stdout.splitlines(), open(d / f"{prefix}_summary.tsv").read()


def test_maker_matches_oracle(data):
    B, d, seeds, recs, names, _ = data
    assert open(d / "mi_ids.txt").read().splitlines() == \
        [f"{i}\t{n}" for i, n in enumerate(names, 1)]
    mf, info, bv, dat = load_device(B, str(d / "mi.bf"))
    om = O.Mibf(recs, seeds)
    assert info.m == om.m and info.pop == om.pop
    assert np.array_equal(bv, om.bv)
    assert np.array_equal(dat[: om.pop], om.data[: om.pop])
    assert mf.seeds() == seeds


@pytest.mark.parametrize("out,hit_only", [("tsv", False), ("tsv", True), ("fa", False),
                                          ("fq", True)])
def test_categorizer_single_end(data, out, hit_only):
    B, d, seeds, recs, names, reads = data
    args = ([] if out == "tsv" else [f"--{out}"]) + (["-i"] if hit_only else [])
    args.append(str(d / "all.fq"))
    lines, summary = run_cat(d, args, f"se_{out}_{int(hit_only)}")
    om = O.Mibf(recs, seeds)
    rows, sats, _ = om.query([s for _, s in reads], score=1e-10, allowed_miss=len(seeds) - 1,
                             multi=1.0, min_nonsat=0, max_hits=64)
    full = ["repeat"] + names
    assert lines == expected_lines(full, reads, rows, sats, out, hit_only)
    assert summary == expected_summary(full, rows, len(reads))


def test_categorizer_options(data):
    """-s, -a, -m, -c and -b reach the query (reads sent as text here: the
    other tests send them 2-bit packed)."""
    B, d, seeds, recs, names, reads = data
    lines, summary = run_cat(d, ["--ascii_reads", "-s", "30", "-a", "2", "-m", "0.5", "-c", "2", "-b",
                                 str(d / "all.fq")], "opts")
    om = O.Mibf(recs, seeds)
    rows, sats, _ = om.query([s for _, s in reads], score=10 ** -3.0, allowed_miss=1,
                             multi=0.5, min_nonsat=2, max_hits=64, best_agree=True)
    full = ["repeat"] + names
    assert lines == expected_lines(full, reads, rows, sats)
    assert summary == expected_summary(full, rows, len(reads))


@pytest.mark.parametrize("hit_only", [False, True])
def test_categorizer_paired(data, hit_only):
    B, d, seeds, recs, names, reads = data
    args = ["-e"] + (["-i"] if hit_only else []) + [str(d / "r1.fq"), str(d / "r2.fq")]
    lines, summary = run_cat(d, args, f"pe_{int(hit_only)}")
    om = O.Mibf(recs, seeds)
    rows, sats, _ = om.query([s for _, s in reads], score=1e-10, allowed_miss=len(seeds) - 1,
                             multi=1.0, min_nonsat=0, max_hits=64, paired=True)
    full = ["repeat"] + names
    assert lines == expected_lines(full, reads, rows, sats, hit_only=hit_only, paired=True)
    assert summary == expected_summary(full, rows, len(reads) // 2, hit_only_paired=hit_only)


@pytest.mark.parametrize("paired", [False, True])
def test_categorizer_many_candidates(data, tmp_path, paired):
    """100 references share one segment: reads from it have dozens to more
    than 64 candidate IDs.  Every candidate is printed, in rank order (the
    records with more than the first query keeps are queried again), and
    lines and summary equal the oracle's."""
    B, d, seeds, *_ = data
    rng = random.Random(31)
    base = "".join(rng.choice("ACGT") for _ in range(1500))
    recs = [(i + 1, base) for i in range(100)]
    names = [f"s{i}" for i, _ in recs]
    write_fasta(tmp_path / "shared.fa", [(n, s) for n, (_, s) in zip(names, recs)])
    print("[gpu] biobloommimaker, 100 shared references", flush=True)
    subprocess.run([MIMAKER, "-p", "sh", "-S", " ".join(seeds), str(tmp_path / "shared.fa")],
                   check=True, timeout=120, cwd=tmp_path)
    reads = []
    for j, a in enumerate(range(0, 800, 53)):
        L = [150, 700, 40, 300][j % 4]
        reads.append((f"m{j}", base[a:a + L]))
    reads.append(("junk", "".join(rng.choice("ACGT") for _ in range(150))))
    if len(reads) % 2:
        reads.append(("junk2", "".join(rng.choice("ACGT") for _ in range(150))))
    om = O.Mibf(recs, seeds)
    rows, sats, _ = om.query([s for _, s in reads], score=1e-10, allowed_miss=len(seeds) - 1,
                             multi=0.0, max_hits=1024, paired=paired)
    assert max(len(r) for r in rows) > 64
    if paired:
        write_fastq(tmp_path / "a1.fq", reads[0::2])
        write_fastq(tmp_path / "a2.fq", reads[1::2])
        args = ["-e", "-m", "0", "--batch_reads=3", str(tmp_path / "a1.fq"),
                str(tmp_path / "a2.fq")]
    else:
        write_fastq(tmp_path / "a.fq", reads)
        args = ["-m", "0", "--batch_reads=5", str(tmp_path / "a.fq")]
    print(f"[gpu] biobloommicategorizer many candidates paired={paired}", flush=True)
    r = subprocess.run([MICAT, "-p", str(tmp_path / "mc"), "-f", str(tmp_path / "sh.bf")] + args,
                       check=True, timeout=120, capture_output=True, text=True)
    full = ["repeat"] + names
    assert r.stdout.splitlines() == expected_lines(full, reads, rows, sats, paired=paired)
    nrec = len(reads) // 2 if paired else len(reads)
    assert open(tmp_path / "mc_summary.tsv").read() == expected_summary(full, rows, nrec)


@pytest.mark.parametrize("paired", [False, True])
def test_categorizer_small_batches(data, paired):
    """Many small device batches (--batch_reads=37): batches are packed,
    queried and formatted at the same time, and the lines still come out in
    input order, equal to the oracle's."""
    B, d, seeds, recs, names, reads = data
    args = ["--batch_reads=37", "--fq"] + (["-e", str(d / "r1.fq"), str(d / "r2.fq")]
                                          if paired else [str(d / "all.fq")])
    lines, summary = run_cat(d, args, f"small_{int(paired)}")
    om = O.Mibf(recs, seeds)
    rows, sats, _ = om.query([s for _, s in reads], score=1e-10, allowed_miss=len(seeds) - 1,
                             multi=1.0, min_nonsat=0, max_hits=64, paired=paired)
    full = ["repeat"] + names
    assert lines == expected_lines(full, reads, rows, sats, "fq", paired=paired)
    nrec = len(reads) // 2 if paired else len(reads)
    assert summary == expected_summary(full, rows, nrec)


def test_maker_kmer_mode_and_ids_by_file(data, tmp_path):
    B, d, seeds, recs, names, _ = data
    write_fasta(tmp_path / "a.fa", [(names[0], recs[0][1]), (names[1], recs[1][1])])
    write_fasta(tmp_path / "b.fa", [(names[2], recs[2][1])])
    print("[gpu] biobloommimaker -k 25 -g 3 -F", flush=True)
    subprocess.run([MIMAKER, "-p", "km", "-k", "25", "-g", "3", "-F",
                    str(tmp_path / "a.fa"), str(tmp_path / "b.fa")],
                   check=True, timeout=120, cwd=tmp_path)
    assert open(tmp_path / "km_ids.txt").read().splitlines() == ["1\ta.fa", "2\tb.fa"]
    mf, info, bv, dat = load_device(B, str(tmp_path / "km.bf"))
    assert info.kmer_mode == 1 and info.nseeds == 3
    om = O.Mibf([(1, recs[0][1]), (1, recs[1][1]), (2, recs[2][1])], O.Kmer(25, 3))
    assert info.m == om.m
    assert np.array_equal(bv, om.bv)
    assert np.array_equal(dat[: om.pop], om.data[: om.pop])


def test_cli_errors(data, tmp_path):
    B, d, *_ = data
    r = subprocess.run([MIMAKER, "-p", "x", str(d / "refs.fa")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Please pick number of hash values" in r.stderr
    r = subprocess.run([MICAT, "-f", str(tmp_path / "none.bf"), str(d / "all.fq")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "File cannot be opened" in r.stderr
    r = subprocess.run([MICAT, "-e", "-f", str(d / "mi.bf"), str(d / "all.fq")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "paired end mode" in r.stderr
