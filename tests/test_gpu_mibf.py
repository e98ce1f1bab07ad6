"""Multi-index Bloom filter on the device against the oracle
(oracle/bbt_mibf.cpp): the built bit vector and ID slots bit for bit, the
per-read query results (candidates in rank order with all five statistics,
satCount, lookup counters), the stored file, and the argument checks."""
import random

import numpy as np
import pytest

import oracle_lib as O
from test_oracle_mibf import make_refs, seeds_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    B.load()
    return D, B


def reads_from(rng, recs, n, lens=(40, 100, 150), err=0.02, junk=0.3):
    reads = []
    for _ in range(n):
        if rng.random() > junk:
            _, ref = rng.choice(recs)
            L = rng.choice(lens)
            a = rng.randrange(0, len(ref) - L)
            r = list(ref[a:a + L])
            for i in range(len(r)):
                if rng.random() < err:
                    r[i] = rng.choice("ACGTN")
            reads.append("".join(r))
        else:
            reads.append("".join(rng.choice("ACGT") for _ in range(rng.choice([5, 10, 120]))))
    return reads


def classify_both(D, B, mf, dr, params):
    """D.mibf_classify through the direct-mapped table and again through the
    lines and the ID array: both layouts must give the same per-record
    results and lookup counters."""
    got = D.mibf_classify(mf, dr, params)
    if mf.info().direct:
        lp = B.MibfParams(params.score, params.multi_thresh, params.allowed_miss,
                          params.min_nonsat, params.max_hits, params.flags | B.MIBF_LINES)
        other = D.mibf_classify(mf, dr, lp)
        assert other[0] == got[0] and other[1] == got[1], "layouts disagree"
        assert np.array_equal(other[2], got[2]) and np.array_equal(other[3], got[3])
    return got


def device_reads(D, reads):
    buf, off = O.pack_reads(reads)
    return D.DeviceReads(buf, off)


def download(mf):
    info = mf.info()
    bv = np.zeros(info.m // 64, np.uint64)
    data = np.zeros(max(1, info.pop), np.uint32)
    mf.download(bv, data)
    return info, bv, data


@pytest.mark.parametrize("case", [(15, 11, 0.5, 1), (25, 20, 0.3, 2), (21, 15, 0.7, 3)])
def test_build_matches_oracle(dev, case):
    D, B = dev
    span, weight, occ, seed = case
    rng = random.Random(100 + seed)
    seeds = seeds_for(rng, span=span, weight=weight)
    recs = make_refs(rng, n=6, L=3000)
    print(f"[gpu] mibf build span={span} occ={occ}", flush=True)
    om = O.Mibf(recs, seeds, occupancy=occ)
    names = [f"ref{i}" for i, _ in recs]
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs], names, seeds, occupancy=occ)
    info, bv, data = download(mf)
    assert info.m == om.m, "occupancy sizing differs"
    assert info.pop == om.pop
    assert np.array_equal(bv, om.bv)
    assert np.array_equal(data[: om.pop], om.data[: om.pop])
    assert info.pop_saturated == int(np.count_nonzero(om.data[: om.pop] & 0x80000000))
    assert info.pop_saturated > 0
    assert mf.name(0) == "repeat" and mf.name(3) == "ref3"
    assert mf.seeds() == seeds
    for a in range(len(seeds)):
        assert np.allclose(mf.frame_probs(a), om.frame_probs(a), rtol=1e-12, atol=0)


def test_build_ids_by_file(dev):
    """Several records per ID (MIBFGen's idByFile mode)."""
    D, B = dev
    rng = random.Random(7)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=6, L=1200)
    ids = [1, 2, 1, 3, 2, 3]
    om = O.Mibf([(i, s) for i, (_, s) in zip(ids, recs)], seeds)
    mf = D.build_mibf([s for _, s in recs], ids, ["a.fa", "b.fa", "c.fa"], seeds, m=om.m)
    info, bv, data = download(mf)
    assert info.nids == 3
    assert np.array_equal(bv, om.bv)
    assert np.array_equal(data[: om.pop], om.data[: om.pop])


@pytest.mark.parametrize("seed", range(3))
def test_query_matches_oracle(dev, seed):
    D, B = dev
    rng = random.Random(200 + seed)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=7, L=2500)
    occ = rng.choice([0.3, 0.5])
    om = O.Mibf(recs, seeds, occupancy=occ)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = reads_from(rng, recs, 300)
    dr = device_reads(D, reads)
    for score, a, multi, mns in [(1e-10, 2, 1.0, 0), (1e-3, 1, 0.5, 0), (1e-10, 0, 1.0, 0),
                                 (0.1, 2, 0.0, 2)]:
        print(f"[gpu] mibf query seed={seed} score={score} miss={a} multi={multi}", flush=True)
        want, wsat, wctr = om.query(reads, score=score, allowed_miss=a, multi=multi,
                                    min_nonsat=mns, max_hits=8)
        got, gsat, gctr, _ = classify_both(D, B, mf, dr, B.mibf_params(score, multi, a, mns, 8))
        assert gsat == wsat
        for r in range(len(reads)):
            assert got[r] == want[r], (r, reads[r])
        assert gctr[0] == wctr[0] and gctr[1] == wctr[1] and gctr[2] == wctr[2]
        assert gctr[3] == 0
    assert sum(len(r) for r in want) > 0


@pytest.mark.parametrize("L", [700, 1300, 2600])
def test_query_long_reads(dev, L):
    """Longer reads take the larger per-read record stores (frames x seeds up
    to 8192)."""
    D, B = dev
    rng = random.Random(L)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=4, L=6000)
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = reads_from(rng, recs, 40, lens=(L, L // 2, 64), junk=0.1)
    dr = device_reads(D, reads)
    print(f"[gpu] mibf long reads L={L}", flush=True)
    want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=1, multi=0.5, max_hits=8)
    got, gsat, gctr, _ = classify_both(D, B, mf, dr, B.mibf_params(1e-10, 0.5, 1, 0, 8))
    assert gsat == wsat
    assert got == want
    assert gctr[0] == wctr[0] and gctr[1] == wctr[1]


def test_max_hits_overflow_counted(dev):
    """Candidates beyond max_hits are counted, the first max_hits kept."""
    D, B = dev
    rng = random.Random(9)
    seeds = seeds_for(rng)
    base = "".join(rng.choice("ACGT") for _ in range(800))
    recs = [(i + 1, base) for i in range(5)]  # five identical references
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = [base[a:a + 150] for a in range(0, 600, 50)]
    dr = device_reads(D, reads)
    want, wsat, wctr = om.query(reads, score=0.5, allowed_miss=2, multi=0.0, max_hits=2)
    got, gsat, gctr, nh = classify_both(D, B, mf, dr, B.mibf_params(0.5, 0.0, 2, 0, 2))
    assert got == want and gsat == wsat
    assert wctr[2] > 0 and gctr[2] == wctr[2]


@pytest.mark.parametrize("nids,L", [(100, 150), (100, 700), (1100, 300), (1100, 700)])
def test_more_than_64_candidates(dev, nids, L):
    """Reads from a segment that many references share have more than 64
    candidate IDs: such records go through the wide pass (counted in
    counters[3]) and equal the oracle, which keeps every candidate, in rank
    order, with the candidates beyond max_hits counted.  100 IDs take the
    per-ID LDS totals, 1100 the sorted frame records."""
    D, B = dev
    rng = random.Random(31)
    seeds = seeds_for(rng)
    base = "".join(rng.choice("ACGT") for _ in range(1500))
    recs = [(i + 1, base) for i in range(nids)]
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = [base[a:a + L] for a in range(0, 1500 - L, 97)] + reads_from(rng, recs, 20)
    dr = device_reads(D, reads)
    wide_seen = 0
    for score, multi, mh, agree in [(1e-10, 1.0, 64, False), (1e-10, 0.5, 8, False),
                                    (0.5, 0.0, 8, False), (0.5, 0.0, 64, True)]:
        print(f"[gpu] mibf wide nids={nids} L={L} score={score} multi={multi} max_hits={mh} "
              f"best_agree={agree}", flush=True)
        want, wsat, wctr = om.query(reads, score=score, allowed_miss=2, multi=multi,
                                    max_hits=mh, best_agree=agree)
        got, gsat, gctr, nh = classify_both(
            D, B, mf, dr, B.mibf_params(score, multi, 2, 0, mh, best_agree=agree))
        assert gsat == wsat
        for r in range(len(reads)):
            assert got[r] == want[r], (r, len(reads[r]))
        assert gctr[0] == wctr[0] and gctr[1] == wctr[1] and gctr[2] == wctr[2]
        assert not (nh == 0xFFFFFFFF).any()
        wide_seen += int(gctr[3])
    assert wide_seen > 0


def test_store_load_roundtrip(dev, tmp_path):
    D, B = dev
    rng = random.Random(11)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=4, L=1500)
    names = ["chrA", "chrB", "chrC", "chrD"]
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs], names, seeds)
    prefix = str(tmp_path / "multi")
    mf.store(prefix)
    ids_txt = open(prefix + "_ids.txt").read().splitlines()
    assert ids_txt == [f"{i}\t{n}" for i, n in enumerate(names, 1)]
    mf2 = B.Mibf.load(prefix + ".bf")
    i1, bv1, d1 = download(mf)
    i2, bv2, d2 = download(mf2)
    assert (i1.m, i1.pop, i1.pop_saturated, i1.nids) == (i2.m, i2.pop, i2.pop_saturated, i2.nids)
    assert np.array_equal(bv1, bv2) and np.array_equal(d1, d2)
    assert [mf2.name(i) for i in range(5)] == ["repeat"] + names
    reads = reads_from(rng, recs, 50)
    dr = device_reads(D, reads)
    p = B.mibf_params(1e-10, 1.0, 2, 0, 4)
    assert D.mibf_classify(mf, dr, p)[:2] == D.mibf_classify(mf2, dr, p)[:2]


def test_load_errors(dev, tmp_path):
    D, B = dev
    bad = tmp_path / "x.bf"
    bad.write_bytes(b"not a filter at all, no header" * 4)
    with pytest.raises(B.BbtError, match="status 3"):
        B.Mibf.load(str(bad))
    with pytest.raises(B.BbtError, match="status 2"):
        B.Mibf.load(str(tmp_path / "missing.bf"))


def test_records_without_separator_rejected(dev):
    torch = pytest.importorskip("torch")
    D, B = dev
    seq = np.frombuffer(b"ACGTACGTACGTACGTACGTACGTACGTACGTN", np.uint8)
    off = [0, 10, len(seq)]  # record 0 ends in 'G'
    dseq = torch.from_numpy(seq.copy()).cuda()
    doff = torch.tensor(off, dtype=torch.int64).cuda()
    did = torch.tensor([1, 2], dtype=torch.int32).cuda()
    with pytest.raises(B.BbtError, match="separator"):
        B.Mibf.build(dseq.data_ptr(), doff.data_ptr(), did.data_ptr(), off, ["a", "b"],
                     ["1101"], occupancy=0.5)
    with pytest.raises(B.BbtError, match="status 1"):
        B.Mibf.build(dseq.data_ptr(), doff.data_ptr(), did.data_ptr(), off, ["a", "b"],
                     ["1101", "111"], occupancy=0.5)


def test_kmer_mode_matches_oracle(dev, tmp_path):
    """biobloommimaker without -S: h ntHash values per k-mer."""
    D, B = dev
    rng = random.Random(31)
    recs = make_refs(rng, n=5, L=2000)
    km = O.Kmer(25, 4)
    om = O.Mibf(recs, km)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], None, kmer=(25, 4))
    info, bv, data = download(mf)
    assert info.kmer_mode == 1 and info.nseeds == 4 and info.span == 25
    assert mf.seeds() is None
    assert info.m == om.m
    assert np.array_equal(bv, om.bv)
    assert np.array_equal(data[: om.pop], om.data[: om.pop])
    reads = reads_from(rng, recs, 200)
    dr = device_reads(D, reads)
    print("[gpu] mibf k-mer mode query", flush=True)
    want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=0, multi=1.0, max_hits=8)
    got, gsat, gctr, _ = classify_both(D, B, mf, dr, B.mibf_params(1e-10, 1.0, 0, 0, 8))
    assert got == want and gsat == wsat
    assert gctr[0] == wctr[0] and gctr[1] == wctr[1]
    mf.store(str(tmp_path / "km"))
    mf2 = B.Mibf.load(str(tmp_path / "km.bf"))
    assert mf2.info().kmer_mode == 1
    assert classify_both(D, B, mf2, dr, B.mibf_params(1e-10, 1.0, 0, 0, 8))[:2] == (got, gsat)


@pytest.mark.parametrize("best_agree", [False, True])
def test_pairs_match_oracle(dev, best_agree):
    """Pairs pool both mates' frames with the reference's frameCount nesting
    (short first mates included); bestHitAgree on and off."""
    D, B = dev
    rng = random.Random(50 + best_agree)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=6, L=3000)
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = []
    for i in range(300):
        _, ref = rng.choice(recs)
        a = rng.randrange(0, len(ref) - 400)
        l1 = 5 if i % 7 == 0 else rng.choice([30, 100, 150])
        l2 = 8 if i % 11 == 0 else rng.choice([60, 150, 250])
        reads.append(ref[a:a + l1])
        reads.append(ref[a + 150:a + 150 + l2])
    dr = device_reads(D, reads)
    for score, a, multi in [(1e-10, 2, 1.0), (1e-3, 1, 0.0)]:
        print(f"[gpu] mibf pairs best_agree={best_agree} score={score} miss={a}", flush=True)
        want, wsat, wctr = om.query(reads, score=score, allowed_miss=a, multi=multi,
                                    max_hits=8, paired=True, best_agree=best_agree)
        got, gsat, gctr, _ = classify_both(
            D, B, mf, dr, B.mibf_params(score, multi, a, 0, 8, paired=True, best_agree=best_agree))
        assert len(got) == len(reads) // 2
        assert gsat == wsat
        assert got == want
        assert gctr[0] == wctr[0] and gctr[1] == wctr[1] and gctr[2] == wctr[2]
    assert sum(len(r) for r in want) > 0


def test_odd_paired_batch_rejected(dev):
    D, B = dev
    rng = random.Random(3)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=2, L=600)
    mf = D.build_mibf([s for _, s in recs], [1, 2], ["a", "b"], seeds)
    dr = device_reads(D, ["ACGT" * 40] * 3)
    out = D.MibfOutputs(2, 4)
    with pytest.raises(B.BbtError, match="odd"):
        mf.query_batch(dr.struct(), B.mibf_params(paired=True, max_hits=4), out.struct())


def test_many_ids_sort_path(dev):
    """More than 1023 IDs: per-ID totals come from sorting the frame records
    instead of the direct per-ID counters."""
    D, B = dev
    rng = random.Random(61)
    seeds = seeds_for(rng)
    base = ["".join(rng.choice("ACGT") for _ in range(400)) for _ in range(1100)]
    recs = [(i + 1, s) for i, s in enumerate(base)]
    # a few references share a segment so frames see several IDs
    for i in range(0, 1100, 50):
        j = (i + 7) % 1100
        recs[i] = (recs[i][0], recs[i][1][:100] + recs[j][1][100:250] + recs[i][1][250:])
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    assert mf.info().nids == 1100
    reads = []
    for _ in range(200):
        _, ref = rng.choice(recs)
        a = rng.randrange(0, 250)
        reads.append(ref[a:a + 150])
    reads += ["".join(rng.choice("ACGT") for _ in range(150)) for _ in range(20)]
    dr = device_reads(D, reads)
    print("[gpu] mibf 1100 ids", flush=True)
    want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=2, multi=0.5, max_hits=8)
    got, gsat, gctr, _ = classify_both(D, B, mf, dr, B.mibf_params(1e-10, 0.5, 2, 0, 8))
    assert gsat == wsat
    assert got == want
    assert gctr[0] == wctr[0] and gctr[1] == wctr[1]
    assert sum(len(r) for r in want) > 100


def test_query_reads_beyond_the_record_store(dev):
    """With at most 1023 IDs the per-ID totals live in LDS and records of any
    length are evaluated: reads of 5-20 kbp, single and paired."""
    D, B = dev
    rng = random.Random(71)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=4, L=30000)
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = []
    for L in [5000, 9000, 20000, 150, 20000, 7000]:
        _, ref = rng.choice(recs)
        a = rng.randrange(0, len(ref) - L)
        r = list(ref[a:a + L])
        for i in range(len(r)):
            if rng.random() < 0.03:
                r[i] = rng.choice("ACGTN")
        reads.append("".join(r))
    dr = device_reads(D, reads)
    print("[gpu] mibf 5-20 kbp reads", flush=True)
    for paired in [False, True]:
        want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=2, multi=0.5,
                                    max_hits=8, paired=paired)
        got, gsat, gctr, _ = classify_both(
            D, B, mf, dr, B.mibf_params(1e-10, 0.5, 2, 0, 8, paired=paired))
        assert gsat == wsat
        assert got == want
        assert gctr[0] == wctr[0] and gctr[1] == wctr[1] and gctr[3] == 0
    assert all(row for row in want)


def test_queries_on_two_streams(dev):
    """Two batches queried on two streams without a host wait between them:
    the filter's shared work queue must not mix them (the second call's
    kernels wait for the first's); both equal the oracle."""
    D, B = dev
    import torch
    rng = random.Random(77)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=6, L=3000)
    om = O.Mibf(recs, seeds, occupancy=0.5)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    ra = reads_from(rng, recs, 3000)
    rb = reads_from(rng, recs, 2000, lens=(150, 250))
    da, db = device_reads(D, ra), device_reads(D, rb)
    p = B.mibf_params(1e-10, 1.0, 2, 0, 8)
    oa, ob = D.MibfOutputs(len(ra), 8), D.MibfOutputs(len(rb), 8)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for rep in range(3):
        print(f"[gpu] mibf two streams rep={rep}", flush=True)
        oa.counters.zero_()
        ob.counters.zero_()
        oa.nhits.fill_(-1)
        ob.nhits.fill_(-1)
        torch.cuda.synchronize()
        mf.query_batch(da.struct(), p, oa.struct(), int(sa.cuda_stream))
        mf.query_batch(db.struct(), p, ob.struct(), int(sb.cuda_stream))
        torch.cuda.synchronize()
        for reads, out in ((ra, oa), (rb, ob)):
            want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=2, multi=1.0,
                                        min_nonsat=0, max_hits=8)
            got, gsat, gctr, _ = out.host()
            assert got == want and gsat == wsat
            assert gctr[0] == wctr[0] and gctr[1] == wctr[1]


def test_direct_table_and_id_zero_slots(dev, tmp_path, monkeypatch):
    """The direct-mapped table is built by default and not with
    BBT_MIBF_DIRECT=0; a loaded filter whose slots hold ID 0 (saturated or
    not) gives the same results through the table as through the lines and
    the ID array, and as the oracle on the same arrays."""
    import struct
    D, B = dev
    rng = random.Random(91)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=5, L=3000)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, occupancy=0.5)
    assert mf.info().direct == 1
    monkeypatch.setenv("BBT_MIBF_DIRECT", "0")
    mf0 = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                       [str(i) for i, _ in recs], seeds, occupancy=0.5)
    assert mf0.info().direct == 0
    monkeypatch.delenv("BBT_MIBF_DIRECT")
    prefix = str(tmp_path / "z")
    mf.store(prefix)
    raw = bytearray(open(prefix + ".bf", "rb").read())
    nseeds, span, nids, kmer, m, pop = struct.unpack_from("<IIIIQQ", raw, 8)
    data_at = 40 + nseeds * span + (m // 64) * 8
    data = np.frombuffer(raw, np.uint32, count=pop, offset=data_at).copy()
    pick = np.random.default_rng(5).choice(pop, size=pop // 5, replace=False)
    data[pick[: len(pick) // 2]] = 0                 # ID 0
    data[pick[len(pick) // 2:]] = 0x80000000        # ID 0, saturated
    raw[data_at:data_at + 4 * pop] = data.tobytes()
    open(prefix + ".bf", "wb").write(bytes(raw))
    mz = B.Mibf.load(prefix + ".bf")
    assert mz.info().direct == 1
    info, bv, dz = download(mz)
    assert np.array_equal(dz[:pop], data)
    om = O.Mibf.from_arrays(bv, dz, info.m, seeds, info.nids)
    reads = reads_from(rng, recs, 400)
    dr = device_reads(D, reads)
    for a in (0, 2):
        want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=a, multi=1.0,
                                    min_nonsat=0, max_hits=8)
        got, gsat, gctr, _ = classify_both(D, B, mz, dr, B.mibf_params(1e-10, 1.0, a, 0, 8))
        assert got == want and gsat == wsat
        assert gctr[0] == wctr[0] and gctr[1] == wctr[1]


def test_2bit_reads_query_like_ascii(dev):
    """MIBF queries of BBT_READS_2BIT batches equal the ASCII batch's and the
    oracle's, single and paired, with lower case and N in the reads."""
    D, B = dev
    rng = random.Random(55)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=6, L=3000)
    om = O.Mibf(recs, seeds, occupancy=0.5)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = [r.lower() if i % 3 == 0 else r for i, r in enumerate(reads_from(rng, recs, 800))]
    buf, off = O.pack_reads(reads)
    da, dp = D.DeviceReads(buf, off), D.DeviceReads.packed(buf, off)
    for paired in (False, True):
        print(f"[gpu] mibf 2-bit reads paired={paired}", flush=True)
        p = B.mibf_params(1e-10, 1.0, 2, 0, 8, paired=paired)
        want, wsat, _ = om.query(reads, score=1e-10, allowed_miss=2, multi=1.0, min_nonsat=0,
                                 max_hits=8, paired=paired)
        ga = D.mibf_classify(mf, da, p)
        gp = classify_both(D, B, mf, dp, p)
        assert gp[0] == ga[0] == want and gp[1] == ga[1] == wsat
        assert np.array_equal(gp[2], ga[2])


def test_long_records_with_many_ids(dev):
    """More than 1023 IDs (the sorted frame-record path) and records of more
    than 8192 frames x seeds: the records go to a device-memory store per
    workgroup; results equal the oracle's, single and paired."""
    D, B = dev
    rng = random.Random(73)
    seeds = seeds_for(rng)
    recs = make_refs(rng, n=4, L=20000)
    recs += [(5 + i, "".join(rng.choice("ACGT") for _ in range(120))) for i in range(1100)]
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    assert mf.info().nids > 1023
    reads = []
    for L in [4000, 9000, 150, 3500, 12000, 150, 6000, 300]:
        _, ref = recs[rng.randrange(0, 4)]
        a = rng.randrange(0, len(ref) - L)
        r = list(ref[a:a + L])
        for i in range(len(r)):
            if rng.random() < 0.03:
                r[i] = rng.choice("ACGTN")
        reads.append("".join(r))
    dr = device_reads(D, reads)
    print("[gpu] mibf long records, 1104 IDs", flush=True)
    for paired in [False, True]:
        want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=2, multi=0.5,
                                    max_hits=8, paired=paired)
        got, gsat, gctr, _ = classify_both(
            D, B, mf, dr, B.mibf_params(1e-10, 0.5, 2, 0, 8, paired=paired))
        assert gsat == wsat and got == want
        assert gctr[0] == wctr[0] and gctr[1] == wctr[1]
    assert all(row for row in want)


def test_long_records_many_candidates(dev):
    """Long reads (past the LDS record store) from a segment that 1100
    references share: more than 64 candidates, so the wide pass runs over the
    device-memory store too; equal to the oracle."""
    D, B = dev
    rng = random.Random(37)
    seeds = seeds_for(rng)
    base = "".join(rng.choice("ACGT") for _ in range(4200))
    recs = [(i + 1, base) for i in range(1100)]
    om = O.Mibf(recs, seeds)
    mf = D.build_mibf([s for _, s in recs], [i for i, _ in recs],
                      [str(i) for i, _ in recs], seeds, m=om.m)
    reads = [base[a:a + 3500] for a in range(0, 600, 150)] + reads_from(rng, recs, 10)
    dr = device_reads(D, reads)
    print("[gpu] mibf long records, 1100 shared IDs", flush=True)
    want, wsat, wctr = om.query(reads, score=1e-10, allowed_miss=2, multi=0.5, max_hits=16)
    got, gsat, gctr, nh = classify_both(D, B, mf, dr, B.mibf_params(1e-10, 0.5, 2, 0, 16))
    assert gsat == wsat and got == want
    assert gctr[0] == wctr[0] and gctr[1] == wctr[1] and gctr[2] == wctr[2]
    assert int(gctr[3]) > 0 and not (nh == 0xFFFFFFFF).any()
