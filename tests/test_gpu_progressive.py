"""Progressive filters on the GPU: bbt_eval_subtract (evalRead with a
subtract filter) and bbt_filter_insert_ordered against the oracle, and
biobloommaker -r end to end (two files, -i, -b bait, -s/-d, LENGTH scoring,
-e iterations, -n threshold, file lists, -P) against the oracle's one-thread
progressive loop started from the same seed filter."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_cli import MAKER, load_bf, write_fasta, write_fastq

pytestmark = pytest.mark.gpu
K, H = 25, 3


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    B.load()
    return torch, D, B


def rand_seq(rng, n):
    return "".join(rng.choice(list("ACGT"), n))


@pytest.mark.parametrize("scorer,thr", [(O.SIMPLE, 0.15), (O.SIMPLE, 0.5), (O.LENGTH, 40)])
def test_eval_subtract_matches_oracle(env, scorer, thr):
    torch, D, B = env
    rng = np.random.default_rng(3)
    g = rand_seq(rng, 30_000)
    fs = B.FilterSet.create([1 << 18, 1 << 16], [H, H], [K, K])
    f, s = O.Filter(1 << 18, H, K), O.Filter(1 << 16, H, K)
    f.insert_seq(g[:20_000])
    s.insert_seq(g[8_000:11_000])
    fs.upload(0, f.bits)
    fs.upload(1, s.bits)
    sub = B.FilterSet.create([1 << 16], [H], [K])
    sub.upload(0, s.bits)
    reads = []
    for i in range(3000):
        L = int(rng.integers(10, 400))
        a = int(rng.integers(0, len(g) - L))
        r = g[a:a + L] if i % 3 else rand_seq(rng, L)
        if i % 7 == 0:
            r = r[:L // 2] + "N" + r[L // 2 + 1:]
        reads.append(r)
    buf, off = O.pack_reads(reads)
    dr = D.DeviceReads(buf, off)
    out = torch.zeros(len(reads), dtype=torch.uint8, device="cuda")
    p = O.params(scorer=scorer, threshold=thr)
    for sfs, sref in ((None, None), (sub, s)):
        print(f"[gpu] eval_subtract scorer={scorer} thr={thr} sub={sref is not None}", flush=True)
        fs.eval_subtract(0, sfs, 0, dr.struct(), scorer, thr, 3, out.data_ptr())
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        want = np.array([O.eval_read_sub(f, sref, r, p) for r in reads], np.uint8)
        assert np.array_equal(got, want)
        assert 0 < got.sum() < len(reads)


def test_insert_ordered_counts_are_sequential(env):
    torch, D, B = env
    rng = np.random.default_rng(4)
    unit = rand_seq(rng, 300)
    read = (unit * 12)[:3000]  # repeated k-mers far apart
    fs = B.FilterSet.create([1 << 14], [H], [K])
    dseq = torch.from_numpy(np.frombuffer(read.encode(), np.uint8).copy()).cuda()
    c = (B.C.c_uint64 * 2)()
    B.check(B.load().bbt_filter_insert_ordered(fs._h, 0, B.C.c_void_p(dseq.data_ptr()),
                                               len(read), c, None))
    ref = O.Filter(1 << 14, H, K)
    assert (int(c[0]), int(c[1])) == ref.insert_seq(read)
    bits = np.zeros((1 << 14) // 8, np.uint8)
    fs.download(0, bits)
    assert np.array_equal(bits, ref.bits)


def make_inputs(tmp_path, rng):
    """A genome whose first 3 kbp seed the filter; read pairs walk along it
    (each overlapping the last), mixed with random pairs."""
    g = rand_seq(rng, 60_000)
    write_fasta(tmp_path / "seed.fa", [("seed", g[:3_000])])
    pairs = []
    for i in range(400):
        a = 2_800 + 70 * i
        pairs.append((g[a:a + 150], g[a + 30:a + 180]))
        if i % 3 == 0:
            pairs.append((rand_seq(rng, 150), rand_seq(rng, 150)))
    write_fastq(tmp_path / "r1.fq", [(f"p{i}/1", a) for i, (a, _) in enumerate(pairs)])
    write_fastq(tmp_path / "r2.fq", [(f"p{i}/2", b) for i, (_, b) in enumerate(pairs)])
    return g, pairs


def run_maker(args, tmp_path, prefix="prog"):
    print(f"[gpu] biobloommaker {' '.join(args)}", flush=True)
    r = subprocess.run([MAKER, "-p", prefix, "-o", str(tmp_path), "-k", str(K), "-f", "0.075"]
                       + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    seed = [int(x.split()[-1]) for x in r.stderr.splitlines()
            if x.startswith("Approximated (due to false positives)")][0]
    return r, load_bf(str(tmp_path / f"{prefix}.bf")), seed


def oracle_after(bf, g, seed_total, runs):
    """A filter with the seed inserted, grown by the oracle over runs of
    (pairs_or_reads, kwargs); returns (filter, total, total_reads, tagged)."""
    f = O.Filter(bf.m, bf.h, bf.k)
    f.insert_seq(g[:3_000])
    te, tr, tg = seed_total, 0, 0
    for items, kw in runs:
        te, r, t = O.progressive(f, items, total_entries=te, **kw)
        tr += r
        tg += t
    return f, te, tr, tg


@pytest.mark.parametrize("case", ["std", "inc", "length", "iter", "stop", "sub", "bait"])
def test_maker_progressive(env, tmp_path, case):
    rng = np.random.default_rng(20)
    g, pairs = make_inputs(tmp_path, rng)
    args = ["-r", "0.15", "-n", "40000"]
    kw = dict(score=0.15, expected=40_000)
    if case == "inc":
        args.append("-i")
        kw["inclusive"] = True
    if case == "length":
        args = ["-r", "60", "-n", "40000"]
        kw.update(scorer=O.LENGTH, score=60)
    if case == "iter":
        args += ["-e", "2"]
        kw["iterations"] = 2
    if case == "stop":
        args = ["-r", "0.15", "-n", "6000"]
        kw["expected"] = 6000
    if case == "sub":
        s = O.Filter(1 << 16, H, K)
        s.insert_seq(g[3_300:3_700])
        write_fasta(tmp_path / "rep.fa", [("rep", g[3_300:3_700])])
        subprocess.run([MAKER, "-p", "rep", "-o", str(tmp_path), "-k", str(K), "-f", "0.075",
                        str(tmp_path / "rep.fa")], check=True, timeout=120, capture_output=True)
        sref = load_bf(str(tmp_path / "rep.bf"))
        args += ["-s", str(tmp_path / "rep.bf")]
        kw["sub"] = sref
    if case == "bait":
        args += ["-b", "0.5"]
    r, bf, seed_total = run_maker(args + [str(tmp_path / "seed.fa"), str(tmp_path / "r1.fq"),
                                          str(tmp_path / "r2.fq")], tmp_path)
    if case == "bait":
        b = O.Filter(O.lib().orc_calc_optimal_size(3_000 - K + 1, 0.075, H), H, K)
        b.insert_seq(g[:3_000])
        kw.update(bait=b, bait_score=0.5)
    f, te, tr, tg = oracle_after(bf, g, seed_total, [(pairs, kw)])
    assert np.array_equal(bf.bits, f.bits)
    assert tg >= (1 if case in ("stop", "sub") else 50)
    if te >= kw["expected"]:
        assert f"K-mer threshold reached at read {tr}" in r.stderr
    else:
        assert f"K-mer threshold not reached, number of k-mers: {te}" in r.stderr
    txt = open(tmp_path / "prog.txt").read()
    assert f"num_entries={te}" in txt.replace(" ", "")


def test_maker_progressive_seed_already_enough(env, tmp_path):
    """Without -n the expected k-mers are the seed's own, which the seed
    load (less false positives) nearly reaches: growth stops at once."""
    rng = np.random.default_rng(21)
    g, pairs = make_inputs(tmp_path, rng)
    r, bf, seed_total = run_maker(["-r", "0.15", str(tmp_path / "seed.fa"),
                                   str(tmp_path / "r1.fq"), str(tmp_path / "r2.fq")], tmp_path)
    kw = dict(score=0.15, expected=3_000 - K + 1)
    f, te, tr, tg = oracle_after(bf, g, seed_total, [(pairs, kw)])
    assert np.array_equal(bf.bits, f.bits)
    assert f"K-mer threshold reached at read {tr}" in r.stderr and tr < 20


@pytest.mark.parametrize("paired", [True, False])
def test_maker_progressive_file_lists(env, tmp_path, paired):
    """-l: a list of file pairs (or single files), each read in turn."""
    rng = np.random.default_rng(22)
    g, pairs = make_inputs(tmp_path, rng)
    half = len(pairs) // 2
    groups = [pairs[:half], pairs[half:]]
    lines = []
    for j, grp in enumerate(groups):
        if paired:
            write_fastq(tmp_path / f"a{j}_1.fq", [(f"q{i}", a) for i, (a, _) in enumerate(grp)])
            write_fastq(tmp_path / f"a{j}_2.fq", [(f"q{i}", b) for i, (_, b) in enumerate(grp)])
            lines.append(f"{tmp_path}/a{j}_1.fq {tmp_path}/a{j}_2.fq")
        else:
            write_fastq(tmp_path / f"s{j}.fq", [(f"q{i}", a) for i, (a, _) in enumerate(grp)])
            lines.append(f"{tmp_path}/s{j}.fq")
    (tmp_path / "list.txt").write_text("\n".join(lines) + "\n")
    r, bf, seed_total = run_maker(["-r", "0.15", "-n", "30000", "-e", "2",
                                   "-l", str(tmp_path / "list.txt"),
                                   str(tmp_path / "seed.fa")], tmp_path)
    kw = dict(score=0.15, expected=30_000)
    if paired:
        runs = [(grp, kw) for _ in range(2) for grp in groups]
    else:
        runs = [([a for a, _ in grp], dict(kw, single=True)) for _ in range(2) for grp in groups]
    f, te, tr, tg = oracle_after(bf, g, seed_total, runs)
    assert np.array_equal(bf.bits, f.bits)
    assert tg >= 50
    if te >= kw["expected"]:
        assert f"K-mer threshold reached at read {tr}" in r.stderr
    else:
        assert f"K-mer threshold not reached, number of k-mers: {te}" in r.stderr


def test_maker_progressive_print_reads(env, tmp_path):
    """-P prints each tagged read with its new k-mers, the running total, its
    k-mers in the subtract filter, the tag index and the read count."""
    rng = np.random.default_rng(23)
    g, pairs = make_inputs(tmp_path, rng)
    r, bf, seed_total = run_maker(["-r", "0.15", "-n", "40000", "-P",
                                   str(tmp_path / "seed.fa"), str(tmp_path / "r1.fq"),
                                   str(tmp_path / "r2.fq")], tmp_path)
    heads = [x for x in r.stdout.splitlines() if x.startswith("@p")]
    f = O.Filter(bf.m, bf.h, bf.k)
    f.insert_seq(g[:3_000])
    te, reads_seen, tag = seed_total, 0, 0
    want = []
    p = O.params(threshold=0.15)
    for i, (a, b) in enumerate(pairs):
        reads_seen += 1
        if te >= 40_000:
            break
        if O.eval_read_sub(f, None, a, p) and O.eval_read_sub(f, None, b, p):
            tag += 1
            na = f.insert_seq(a)[0]
            nb = f.insert_seq(b)[0]
            te += na + nb
            want.append(f"@p{i}/1 {na} {te} 0 {tag} {reads_seen}")
            want.append(f"@p{i}/2 {nb} {te} 0 {tag} {reads_seen}")
    assert heads == want and len(want) >= 100
