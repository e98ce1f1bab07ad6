"""Full-size and large-filter parity on the device.

* The single-filter workload at its full size (1 x 5 Mbp genome, k=25,
  FPR 0.075; 100k single-end 150 bp reads, half from the genome with 1%
  substitutions, half random, each base N with probability 0.001): the filter
  built by biobloommaker and on the device equals the oracle's, both kernel
  paths and biobloomcategorizer --status equal the oracle's hits, classes,
  counters, summary.tsv and read status (BioBloomClassifier.cpp:33-100,
  SeqEval.h:37-114).
* A Bloom filter of 18 Gbit (m > 2^34: 64-bit positions and byte offsets
  above 2 GiB in the exact hash % m and the bit address) beside a small one,
  built on the device and probed by both paths, against the oracle.
* A multi-index Bloom filter of more than 2^32 bits, built and queried on the
  device, against the oracle.
* -D on the full single-filter workload with planted low-complexity runs.
"""
import os
import random
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAKER = os.path.join(ROOT, "biobloom_amd", "biobloommaker")
CAT = os.path.join(ROOT, "biobloom_amd", "biobloomcategorizer")


@pytest.fixture(scope="module")
def dev():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    B.load()
    return D, B


def oracle_copy(fs, i):
    inf = fs.info(i)
    f = O.Filter(inf.m, inf.h, inf.k)
    fs.download(i, f.bits)
    return f


def check_paths(D, B, fs, ofs, buf, off, paired=False):
    dr = D.DeviceReads(buf, off)
    oh, _, ocls, octr = O.classify(ofs, buf, off, paired=paired)
    for path in (B.PATH_PLAIN, B.PATH_BATCHED):
        print(f"[gpu] path={path} reads={len(off) - 1} filters={len(ofs)}", flush=True)
        res = D.classify(fs, dr, B.params(paired=paired, path=path))
        assert np.array_equal(res["hits"], oh), "hits differ"
        assert np.array_equal(res["cls"], ocls), "classes differ"
        assert int(res["counters"][0]) == int(octr[0]), "contains() calls differ"
        assert int(res["counters"][1]) == int(octr[1]), "bit lookups differ"
        assert int(res["counters"][3]) == 0
        ids = [fs.filter_id(i) for i in range(len(fs))]
        assert fs.summary_text(res["summary"], len(ocls)) == O.summary(ids, oh, paired)
    return oh, ocls


def test_single_filter_workload_full_size(dev, tmp_path):
    D, B = dev
    from biobloom_amd import synth
    from test_gpu_cli import load_bf, write_fasta, write_fastq
    w = synth.workload_single(1.0)
    g = w["genomes"][0]
    reads = w["reads"]
    assert len(g) == 5_000_000 and reads.shape == (100_000, 150)
    assert np.count_nonzero(reads == ord("N")) > 0
    # biobloommaker from FASTA, the device build, and the oracle insert agree
    write_fasta(tmp_path / "g.fa", [("chr", g.tobytes().decode())])
    print("[gpu] biobloommaker 5 Mbp", flush=True)
    subprocess.run([MAKER, "-p", "g", "-o", str(tmp_path), "-k", "25", "-f", "0.075",
                    str(tmp_path / "g.fa")], check=True, timeout=300)
    made = load_bf(str(tmp_path / "g.bf"))
    assert (made.m, made.h, made.k) == (27_387_648, 3, 25)
    of = O.Filter(made.m, made.h, made.k)
    of.insert_seq(bytes(g))
    assert np.array_equal(made.bits, of.bits)
    fs, _ = D.build_filters([g], 25, 0.075)
    dev_bits = oracle_copy(fs, 0)
    assert np.array_equal(dev_bits.bits, of.bits)
    # both kernel paths on the resident batch
    buf = np.ascontiguousarray(reads).reshape(-1)
    off = np.arange(reads.shape[0] + 1, dtype=np.uint64) * np.uint64(150)
    oh, ocls = check_paths(D, B, fs, [of], buf, off)
    assert 0.3 < oh.mean() < 0.7
    # the CLI with --status on FASTQ, batched and plain kernels
    names = [f"r{j}" for j in range(reads.shape[0])]
    write_fastq(tmp_path / "reads.fq", [(n, r.tobytes().decode()) for n, r in zip(names, reads)])
    want_summary = O.summary(["g"], oh)
    labels = ["g", "noMatch", "multiMatch"]
    want_status = [f"{n}\t{labels[c]}" for n, c in zip(names, ocls)]
    for extra in ([], ["--plain"]):
        print(f"[gpu] biobloomcategorizer --status {extra}", flush=True)
        prefix = tmp_path / ("out_plain" if extra else "out")
        subprocess.run([CAT, "-p", str(prefix), "-f", str(tmp_path / "g.bf"), "--status"]
                       + extra + [str(tmp_path / "reads.fq")], check=True, timeout=300)
        assert open(f"{prefix}_summary.tsv").read() == want_summary
        assert open(f"{prefix}_readStatus.tsv").read().splitlines() == want_status


def test_dust_single_workload_full_size(dev):
    """-D on the single-filter workload at full size (100k reads), every
    20th read with a planted low-complexity run: the LDS sdust kernel takes
    the reads whose interval lists stay short and hands the planted runs to
    the device-memory kernel; hits, classes and counters equal the oracle's
    for the default window, a short one and one past the LDS kernel."""
    D, B = dev
    from biobloom_amd import synth
    w = synth.workload_single(1.0)
    fs, _ = D.build_filters(w["genomes"], 25, 0.075)
    of = oracle_copy(fs, 0)
    rng = np.random.default_rng(31)
    units = [b"A", b"AC", b"CAG", b"ACGT", b"TTAGGG"]
    reads = []
    for i, r in enumerate(w["reads"]):
        r = r.tobytes()
        if i % 20 == 0:
            a = int(rng.integers(0, len(r)))
            r = r[:a] + units[i // 20 % len(units)] * int(rng.integers(4, 30)) + r[a:]
        reads.append(r)
    buf, off = O.pack_reads(reads)
    dr = D.DeviceReads(buf, off)
    for dt, dw in ((20, 64), (10, 32), (20, 80)):
        oh, _, ocls, octr = O.classify([of], buf, off, p=O.params(dust=True, dust_t=dt,
                                                                    dust_w=dw))
        for path in (B.PATH_PLAIN, B.PATH_BATCHED):
            print(f"[gpu] dust T={dt} W={dw} path={path} reads={len(reads)}", flush=True)
            res = D.classify(fs, dr, B.params(path=path, dust=True, dust_t=dt, dust_w=dw))
            assert np.array_equal(res["hits"], oh), "hits differ"
            assert np.array_equal(res["cls"], ocls), "classes differ"
            assert int(res["counters"][0]) == int(octr[0]), "contains() calls differ"
            assert int(res["counters"][1]) == int(octr[1]), "bit lookups differ"
            assert int(res["counters"][3]) == 0


def test_bloom_filter_above_2_34_bits(dev):
    """m = 18,000,000,000 bits (2.25 GB): hash % m above 2^32 and byte
    addresses above 2^31, next to a small filter in the same set."""
    D, B = dev
    import torch
    from biobloom_amd import synth
    rng = synth.rng_for(11)
    big_g = synth.codes_to_ascii(synth.random_codes(rng, 2_000_000))
    small_g = synth.codes_to_ascii(synth.random_codes(rng, 300_000))
    m_big = 18_000_000_000
    m_small = O.lib().orc_calc_optimal_size(300_000 - 24, 0.075, 3)
    fs = B.FilterSet.create([m_big, m_small], [3, 3], [25, 25])
    st = int(torch.cuda.current_stream().cuda_stream)
    for i, g in enumerate((big_g, small_g)):
        dg = torch.from_numpy(np.ascontiguousarray(g)).cuda()
        fs.insert(i, dg.data_ptr(), dg.numel(), st)
        fs.refresh(i, st)
        fs.set_id(i, f"g{i}")
    torch.cuda.synchronize()
    print("[gpu] 18 Gbit filter built", flush=True)
    ofs = []
    for i, g in enumerate((big_g, small_g)):
        of = O.Filter(fs.info(i).m, 3, 25)
        of.insert_seq(bytes(g))
        got = oracle_copy(fs, i)
        assert np.array_equal(got.bits, of.bits), f"filter {i} bits differ"
        # set bits spread over the whole array, past 2 GiB included
        if i == 0:
            nz = np.flatnonzero(of.bits)
            assert nz[-1] > (1 << 31) and nz[0] < (1 << 28)
        ofs.append(of)
    reads, _, _ = synth.fixed_len_reads(rng, [big_g, small_g], 20_000, 150, 0.01, 0.7,
                                        n_prob=0.001)
    buf = np.ascontiguousarray(reads).reshape(-1)
    off = np.arange(reads.shape[0] + 1, dtype=np.uint64) * np.uint64(150)
    oh, _ = check_paths(D, B, fs, ofs, buf, off)
    assert oh[:, 0].sum() > 1000 and oh[:, 1].sum() > 1000
    check_paths(D, B, fs, ofs, buf, off, paired=True)


def test_mibf_above_2_32_bits(dev):
    D, B = dev
    from test_gpu_mibf import classify_both, download, reads_from
    from test_oracle_mibf import make_refs, seeds_for
    rng = random.Random(2024)
    seeds = seeds_for(rng, span=25, weight=20)
    recs = make_refs(rng, n=8, L=20_000)
    m = 5_000_000_000 - 5_000_000_000 % 64
    print("[gpu] mibf 5 Gbit build", flush=True)
    om = O.Mibf(recs, seeds, m=m)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [f"ref{i}" for i, _ in recs], seeds, m=m)
    info, bv, data = download(mf)
    assert info.m == m and info.pop == om.pop
    assert np.array_equal(bv, om.bv)
    assert np.array_equal(data[: om.pop], om.data[: om.pop])
    nz = np.flatnonzero(bv)
    assert nz[-1] * 64 > (1 << 32)
    reads = reads_from(rng, recs, 20_000, lens=(100, 150, 250))
    buf, off = O.pack_reads(reads)
    dr = D.DeviceReads(buf, off)
    print("[gpu] mibf 5 Gbit query", flush=True)
    want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=2, multi=1.0, min_nonsat=0,
                                max_hits=8)
    got, gsat, gctr, _ = classify_both(D, B, mf, dr, B.mibf_params(1e-10, 1.0, 2, 0, 8))
    assert mf.info().direct == 1
    assert gsat == wsat
    assert got == want
    assert gctr[0] == wctr[0] and gctr[1] == wctr[1]
    assert sum(1 for r in want if r) > 5000
