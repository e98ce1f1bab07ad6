"""GPU parity: libbbt_hip through its C ABI against the CPU oracle, on the
same seeded inputs.  Bit-exact for hits, classes, summary counters, scores
and the contains()/lookup counters."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    B.load()
    return D, B


def oracle_filters_like(fs, B):
    out = []
    for i in range(len(fs)):
        inf = fs.info(i)
        f = O.Filter(inf.m, inf.h, inf.k)
        fs.download(i, f.bits)
        out.append(f)
    return out


def small_genomes(rng, n, glen, records=3):
    gs = []
    for _ in range(n):
        parts = [rng.choice(np.frombuffer(b"ACGT", np.uint8), size=glen // records)
                 for _ in range(records)]
        sep = np.frombuffer(b"\n", np.uint8)
        g = parts[0]
        for p in parts[1:]:
            g = np.concatenate([g, sep, p])
        gs.append(g)
    return gs


@pytest.mark.parametrize("k", [25, 11, 32])
def test_device_filter_build_matches_oracle(dev, k):
    D, B = dev
    print(f"[gpu] filter build k={k}", flush=True)
    rng = np.random.default_rng(k)
    genomes = small_genomes(rng, 3, 40_000)
    genomes[1][rng.integers(0, len(genomes[1]), 200)] = ord("N")
    fs, counts = D.build_filters(genomes, k, 0.075)
    for i, g in enumerate(genomes):
        inf = fs.info(i)
        of = O.Filter(inf.m, inf.h, inf.k)
        new, red = of.insert_seq(bytes(g))
        got = np.zeros_like(of.bits)
        fs.download(i, got)
        assert np.array_equal(got, of.bits)
        assert sum(counts[i]) == new + red
        assert inf.popcount == of.popcount()
        assert inf.fpr == of.fpr


PATHS = [0, 1]


def check_batch(D, B, fs, ofs, buf, off, paired=False, inclusive=False, mode=0,
                scorer=0, thr=0.15, streak=3, order=None, wave_min=0, dust=None):
    dr = D.DeviceReads(buf, off, order=order)
    dt, dw = dust if dust else (20, 64)
    p_or = O.params(scorer, thr, streak, bool(dust), dt, dw)
    want_scores = mode in (O.BESTHIT, O.SCORES)
    oh, osc, ocls, octr = O.classify(ofs, buf, off, paired, inclusive, mode, p_or,
                                     want_scores=want_scores)
    for path in PATHS:
        print(f"[gpu] mode={mode} scorer={scorer} thr={thr} streak={streak} "
              f"paired={paired} incl={inclusive} path={path} wave_min={wave_min} "
              f"n={len(off)-1}", flush=True)
        p = B.params(mode, scorer, thr, streak, paired, inclusive, path,
                     wave_min_len=wave_min, dust=bool(dust), dust_t=dt, dust_w=dw)
        res = D.classify(fs, dr, p)
        assert np.array_equal(res["hits"], oh), "hits differ"
        assert np.array_equal(res["cls"], ocls), "classes differ"
        if want_scores:
            assert np.array_equal(res["scores"].view(np.uint64), osc.view(np.uint64))
        assert int(res["counters"][0]) == int(octr[0]), "contains() calls differ"
        assert int(res["counters"][1]) == int(octr[1]), "bit lookups differ"
        assert int(res["counters"][2]) >= int(octr[1]), "issued lookups below reference"
        assert int(res["counters"][3]) == 0, "device error flag set"
        ids = [fs.filter_id(i) for i in range(len(fs))]
        nrec = len(ocls)
        txt = fs.summary_text(res["summary"], nrec)
        want = O.summary(ids, oh, paired, inclusive)
        assert txt == want
    return oh


@pytest.fixture(scope="module")
def multi(dev):
    D, B = dev
    from biobloom_amd import synth
    w = synth.workload_multi(scale=0.004, nfilters=4, genome_len=300_000)
    fs, _ = D.build_filters(w["genomes"], 25, 0.075)
    ofs = oracle_filters_like(fs, B)
    reads = w["reads"]
    # edge cases appended as an extra pair block: empty, short, N-only, lowercase
    extra = [b"", b"ACGT", b"N" * 200, bytes(reads[0]).lower(), b"A" * 24, b"C" * 25,
             bytes(reads[1][:60]) + b"NNNNN" + bytes(reads[1][60:]), b"ACGTN" * 40]
    buf, off = O.pack_reads([bytes(r) for r in reads] + extra)
    return fs, ofs, buf, off


def test_std_single_end_all_scorers(dev, multi):
    D, B = dev
    fs, ofs, buf, off = multi
    for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.2), (O.LENGTH, 40.0),
                        (O.BINOMIAL, 1e-10), (O.SIMPLE, 0.0), (O.SIMPLE, 0.9)]:
        check_batch(D, B, fs, ofs, buf, off, scorer=scorer, thr=thr)


def test_std_streak_variants(dev, multi):
    D, B = dev
    fs, ofs, buf, off = multi
    for streak in [0, 1, 5]:
        check_batch(D, B, fs, ofs, buf, off, streak=streak)


def test_paired_end(dev, multi):
    D, B = dev
    fs, ofs, buf, off = multi
    for inc in [False, True]:
        check_batch(D, B, fs, ofs, buf, off, paired=True, inclusive=inc)
        check_batch(D, B, fs, ofs, buf, off, paired=True, inclusive=inc, mode=O.ORDERED)


def test_modes(dev, multi):
    D, B = dev
    fs, ofs, buf, off = multi
    check_batch(D, B, fs, ofs, buf, off, mode=O.ORDERED)
    for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.15), (O.LENGTH, 30.0),
                        (O.BINOMIAL, 1e-10)]:
        check_batch(D, B, fs, ofs, buf, off, mode=O.BESTHIT, scorer=scorer, thr=thr)
        check_batch(D, B, fs, ofs, buf, off, mode=O.SCORES, scorer=scorer, thr=thr)


def test_whole_wave_walks_all_scorers(dev, multi):
    D, B = dev
    fs, ofs, buf, off = multi
    for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.2), (O.LENGTH, 40.0),
                        (O.BINOMIAL, 1e-10)]:
        check_batch(D, B, fs, ofs, buf, off, scorer=scorer, thr=thr, wave_min=26)
    check_batch(D, B, fs, ofs, buf, off, streak=0, wave_min=26)
    check_batch(D, B, fs, ofs, buf, off, paired=True, wave_min=26)
    for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.15), (O.LENGTH, 30.0),
                        (O.BINOMIAL, 1e-10)]:
        check_batch(D, B, fs, ofs, buf, off, mode=O.BESTHIT, scorer=scorer, thr=thr,
                    wave_min=26)
        check_batch(D, B, fs, ofs, buf, off, mode=O.SCORES, scorer=scorer, thr=thr,
                    wave_min=100)


def test_length_order_does_not_change_results(dev, multi):
    D, B = dev
    fs, ofs, buf, off = multi
    order = D.length_order(np.diff(off))
    check_batch(D, B, fs, ofs, buf, off, order=order)


def test_uneven_lengths(dev):
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(3)
    genomes = [synth.codes_to_ascii(synth.random_codes(rng, 200_000)) for _ in range(2)]
    longs, _ = synth.long_reads(rng, genomes, 60, median=2000, hi=20_000)
    shorts, _, _ = synth.fixed_len_reads(rng, genomes, 400, 150, 0.01, 0.7)
    items = longs + [s for s in shorts]
    buf, off = O.pack_reads([bytes(x) for x in items])
    fs, _ = D.build_filters(genomes, 25, 0.05)
    ofs = oracle_filters_like(fs, B)
    check_batch(D, B, fs, ofs, buf, off)
    check_batch(D, B, fs, ofs, buf, off, order=D.length_order(np.diff(off)))
    for scorer, thr in [(O.HARMONIC, 0.15), (O.LENGTH, 60.0), (O.BINOMIAL, 1e-10)]:
        check_batch(D, B, fs, ofs, buf, off, scorer=scorer, thr=thr)
    # BINOMIAL evalScore of reads past the device table (1024 k-mers): the
    # device counts hits, the host computes -10 log10 of the tail
    assert max(len(x) for x in longs) > 1100
    for mode in (O.BESTHIT, O.SCORES):
        check_batch(D, B, fs, ofs, buf, off, mode=mode, scorer=O.BINOMIAL, thr=1e-10)
        check_batch(D, B, fs, ofs, buf, off, mode=mode, scorer=O.BINOMIAL, thr=1e-10,
                    paired=True)


def test_runtime_k_filters(dev):
    """Filter sets whose k is not the compiled one: all at k = 31, and k = 21,
    31 and 25 in one set (each walk rolls its own filter's k).  Both paths,
    lane and whole-wave walks, every scorer, against the oracle."""
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(31)
    genomes = [synth.codes_to_ascii(synth.random_codes(rng, 120_000)) for _ in range(3)]
    shorts, _, _ = synth.fixed_len_reads(rng, genomes, 600, 150, 0.01, 0.6)
    longs, _ = synth.long_reads(rng, genomes, 20, median=900, hi=3000)
    reads = [bytes(x) for x in shorts] + [bytes(x) for x in longs] + [b"", b"ACGTN" * 12]
    buf, off = O.pack_reads(reads)
    L = B.load()
    for ks in ([31, 31, 31], [21, 31, 25]):
        print(f"[gpu] runtime k {ks}", flush=True)
        h = L.bbt_calc_optimal_hashnum(0.05)
        ms = [L.bbt_calc_optimal_size(len(g) - k + 1, 0.05, h) for g, k in zip(genomes, ks)]
        fs = B.FilterSet.create(ms, [h] * 3, ks)
        for i, g in enumerate(genomes):
            dg = torch.from_numpy(np.ascontiguousarray(g)).to("cuda")
            fs.insert(i, dg.data_ptr(), dg.numel())
            fs.refresh(i)
            fs.set_id(i, f"g{i}")
            del dg
        ofs = oracle_filters_like(fs, B)
        for scorer, thr in [(O.SIMPLE, 0.15), (O.HARMONIC, 0.2), (O.LENGTH, 40.0),
                            (O.BINOMIAL, 1e-10)]:
            check_batch(D, B, fs, ofs, buf, off, scorer=scorer, thr=thr)
            check_batch(D, B, fs, ofs, buf, off, scorer=scorer, thr=thr, wave_min=100)
        check_batch(D, B, fs, ofs, buf, off, paired=True)
        check_batch(D, B, fs, ofs, buf, off, mode=O.SCORES, scorer=O.HARMONIC, thr=0.15)
        check_batch(D, B, fs, ofs, buf, off, mode=O.ORDERED)
        if ks[0] == ks[1] == ks[2]:
            # one geometry: the interleaved read-major walks, lane and whole-wave
            fs.set_tuning(probe_ahead=5)
            check_batch(D, B, fs, ofs, buf, off, scorer=O.BINOMIAL, thr=1e-10)
            check_batch(D, B, fs, ofs, buf, off, wave_min=100, paired=True)
            fs.set_tuning()


def test_partial_interleave_groups(dev):
    """Sets with some filters of one geometry: two of one size and one of
    another; two groups (k 21 and k 25, two filters each).  The interleaved
    read-major walks (probe_ahead 4..7) read each group through its copy and
    the others through their own arrays; both paths equal the oracle."""
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(44)
    L = B.load()
    for lens, ks in (([90_000, 90_000, 140_000], [25, 25, 25]),
                     ([80_000, 80_000, 80_000, 80_000], [21, 25, 21, 25])):
        genomes = [synth.codes_to_ascii(synth.random_codes(rng, n)) for n in lens]
        shorts, _, _ = synth.fixed_len_reads(rng, genomes, 500, 150, 0.01, 0.6)
        longs, _ = synth.long_reads(rng, genomes, 10, median=800, hi=2500)
        reads = [bytes(x) for x in shorts] + [bytes(x) for x in longs] + [b"", b"N" * 40]
        buf, off = O.pack_reads(reads)
        h = L.bbt_calc_optimal_hashnum(0.05)
        ms = [L.bbt_calc_optimal_size(len(g) - k + 1, 0.05, h) for g, k in zip(genomes, ks)]
        fs = B.FilterSet.create(ms, [h] * len(ms), ks)
        for i, g in enumerate(genomes):
            dg = torch.from_numpy(np.ascontiguousarray(g)).to("cuda")
            fs.insert(i, dg.data_ptr(), dg.numel())
            fs.refresh(i)
            fs.set_id(i, f"p{i}")
            del dg
        ofs = oracle_filters_like(fs, B)
        for ahead in (4, 7):
            print(f"[gpu] partial groups lens={lens} k={ks} ahead={ahead}", flush=True)
            fs.set_tuning(probe_ahead=ahead)
            check_batch(D, B, fs, ofs, buf, off)
            check_batch(D, B, fs, ofs, buf, off, paired=True, wave_min=100)
            check_batch(D, B, fs, ofs, buf, off, mode=O.SCORES, scorer=O.LENGTH, thr=40.0)
        fs.set_tuning()
        check_batch(D, B, fs, ofs, buf, off)


@pytest.mark.parametrize("nf", [3, 4, 6, 12, 16])
def test_rows_walk_group_sizes(dev, nf):
    """The read-per-lane walk (probe_ahead 12) over sets of 3 to 16 filters of
    one geometry, i.e. interleaved groups of 4, 8 and 16 bytes per position:
    hits, classes and reference counters equal the oracle's with and without
    -D, single-end and paired.  With reads that reach the whole-wave length,
    sets of 3 to 8 walk each long read by a whole wave against all filters
    at once (larger sets leave the batch on the other walks); the same
    results."""
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(45, nf)
    genomes = [synth.codes_to_ascii(synth.random_codes(rng, 40_000)) for _ in range(nf)]
    shorts, _, _ = synth.fixed_len_reads(rng, genomes, 600, 150, 0.01, 0.6)
    odd = [b"", b"ACGT", b"N" * 100, b"A" * 24, b"C" * 25, b"ACGTN" * 40,
           bytes(shorts[0]).lower(), bytes(shorts[1][:70]) + b"NN" + bytes(shorts[1][70:])]
    reads = [bytes(x) for x in shorts] + odd
    buf, off = O.pack_reads(reads)
    fs, _ = D.build_filters(genomes, 25, 0.05)
    ofs = oracle_filters_like(fs, B)
    longs, _ = synth.long_reads(rng, genomes, 8, median=600, hi=1500)
    lbuf, loff = O.pack_reads(reads + [bytes(x) for x in longs])
    try:
        fs.set_tuning(probe_ahead=12)
        print(f"[gpu] rows walk nf={nf}", flush=True)
        check_batch(D, B, fs, ofs, buf, off)
        D.classify(fs, D.DeviceReads(buf, off), B.params())
        assert fs.last_walk_kernel() == "k_walk_rows"
        D.classify(fs, D.DeviceReads(lbuf, loff), B.params())
        assert fs.last_walk_kernel() == (
            "k_walk_group_long+k_walk_rows" if nf <= 8 else "k_walk_persistent")
        check_batch(D, B, fs, ofs, buf, off, thr=0.0, streak=0)
        check_batch(D, B, fs, ofs, buf, off, thr=0.9, streak=5)
        check_batch(D, B, fs, ofs, buf, off, dust=(20, 64))
        check_batch(D, B, fs, ofs, buf, off, paired=True, inclusive=True)
        check_batch(D, B, fs, ofs, lbuf, loff)
        if nf <= 8:
            print(f"[gpu] rows walk with long reads nf={nf}", flush=True)
            check_batch(D, B, fs, ofs, lbuf, loff, thr=0.0, streak=0)
            check_batch(D, B, fs, ofs, lbuf, loff, thr=0.9, streak=5)
            check_batch(D, B, fs, ofs, lbuf, loff, streak=1)
            check_batch(D, B, fs, ofs, lbuf, loff, dust=(20, 64))
            check_batch(D, B, fs, ofs, lbuf, loff, paired=True, wave_min=100)
            check_batch(D, B, fs, ofs, lbuf, loff, wave_min=64, thr=0.5)
    finally:
        fs.set_tuning()


@pytest.mark.parametrize("ng", [3, 5, 8])
def test_rows_walk_partial_group(dev, ng):
    """A set of one larger filter and ng filters of one geometry (the serving
    set's shape): the group by the read-per-lane walk (and whole-wave walks
    of long reads against all of it), the larger filter by the lane and wave
    walks over its own array in a second launch.  With probe_ahead 12, and
    by default (no knobs: the library takes 15 for SIMPLE in STD mode on a
    batch of 65536 reads or more, here a batch of 66k reads; smaller batches
    keep the lane and wave walks), hits, classes and reference counters equal
    the oracle's; LENGTH, which the read-per-lane walk does not take, too."""
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(46, ng)
    genomes = [synth.codes_to_ascii(synth.random_codes(rng, 150_000))]
    genomes += [synth.codes_to_ascii(synth.random_codes(rng, 40_000)) for _ in range(ng)]
    shorts, _, _ = synth.fixed_len_reads(rng, genomes, 800, 150, 0.01, 0.7)
    longs, _ = synth.long_reads(rng, genomes, 24, median=800, hi=4000)
    reads = [bytes(x) for x in shorts] + [bytes(x) for x in longs]
    reads += [b"", b"ACGT", b"N" * 300, bytes(longs[0]).lower(), b"ACGTN" * 80, b"C" * 25]
    buf, off = O.pack_reads(reads)  # an even count: pairs
    sbuf, soff = O.pack_reads([bytes(x) for x in shorts])
    fs, _ = D.build_filters(genomes, 25, 0.05)
    ofs = oracle_filters_like(fs, B)
    if ng == 8:
        print(f"[gpu] rows walk partial group ng={ng} default, 66k reads", flush=True)
        many, _, _ = synth.fixed_len_reads(rng, genomes, 66_000, 150, 0.01, 0.5)
        mbuf, moff = O.pack_reads([bytes(x) for x in many] + reads)
        check_batch(D, B, fs, ofs, mbuf, moff)
        check_batch(D, B, fs, ofs, mbuf, moff, paired=True)
        # the plain path ran last; a batched call names the group walks
        D.classify(fs, D.DeviceReads(mbuf, moff), B.params())
        assert fs.last_walk_kernel() == "k_walk_group_long+k_walk_rows+k_walk_persistent"
    try:
        for ahead in (None, 12):
            if ahead is not None:  # first the set as built, with no knobs of its own
                fs.set_tuning(probe_ahead=ahead)
            print(f"[gpu] rows walk partial group ng={ng} ahead={ahead}", flush=True)
            check_batch(D, B, fs, ofs, buf, off)
            D.classify(fs, D.DeviceReads(buf, off), B.params())
            assert fs.last_walk_kernel() == (
                "k_walk_persistent" if ahead is None  # a small batch: no group walks
                else "k_walk_group_long+k_walk_rows+k_walk_persistent")
            check_batch(D, B, fs, ofs, sbuf, soff)
            check_batch(D, B, fs, ofs, buf, off, thr=0.5, streak=1)
            check_batch(D, B, fs, ofs, buf, off, dust=(20, 64))
            check_batch(D, B, fs, ofs, buf, off, paired=True, wave_min=100)
            check_batch(D, B, fs, ofs, buf, off, scorer=O.LENGTH, thr=40.0)
    finally:
        fs.set_tuning()


@pytest.mark.parametrize("k,ng,extra", [(21, 5, False), (31, 4, True)])
def test_rows_walk_runtime_k(dev, k, ng, extra):
    """The group walks with a k other than 25 (the roll table and k read at
    run time): a whole set of 5 filters with k 21, and a group of 4 with k 31
    beside one larger filter; short and long reads, single-end, paired and
    graph-replayed batches equal the oracle."""
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(47, k)
    genomes = [synth.codes_to_ascii(synth.random_codes(rng, 120_000))] if extra else []
    genomes += [synth.codes_to_ascii(synth.random_codes(rng, 40_000)) for _ in range(ng)]
    shorts, _, _ = synth.fixed_len_reads(rng, genomes, 600, 150, 0.01, 0.7)
    longs, _ = synth.long_reads(rng, genomes, 16, median=700, hi=3000)
    reads = [bytes(x) for x in shorts] + [bytes(x) for x in longs] + [b"", b"N" * 64]
    buf, off = O.pack_reads(reads)
    fs, _ = D.build_filters(genomes, k, 0.05)
    ofs = oracle_filters_like(fs, B)
    try:
        fs.set_tuning(probe_ahead=15)
        print(f"[gpu] rows walk k={k} ng={ng} extra={extra}", flush=True)
        check_batch(D, B, fs, ofs, buf, off)
        check_batch(D, B, fs, ofs, buf, off, paired=True, thr=0.3)
        check_batch(D, B, fs, ofs, buf, off, dust=(20, 64), wave_min=100)
        dr = D.DeviceReads(buf, off)
        oh, _, ocls, octr = O.classify(ofs, buf, off, paired=True)
        out = D.Outputs(dr.nreads, len(fs), True, False)
        for rep in range(3):
            print(f"[gpu] rows walk k={k} graph replay {rep}", flush=True)
            out.hits.fill_(7)
            res = D.classify(fs, dr, B.params(paired=True, flags=B.FLAG_GRAPH), out=out)
            assert np.array_equal(res["hits"], oh) and np.array_equal(res["cls"], ocls)
            assert int(res["counters"][1]) == int(octr[1])
            eager = D.classify(fs, dr, B.params(paired=True))
            assert np.array_equal(eager["summary"], res["summary"])
    finally:
        fs.set_tuning()


def test_store_and_load_roundtrip(dev, tmp_path):
    D, B = dev
    rng = np.random.default_rng(5)
    genomes = small_genomes(rng, 2, 30_000)
    fs, counts = D.build_filters(genomes, 25, 0.075, ids=["alpha", "beta"])
    paths = []
    for i in range(2):
        prefix = str(tmp_path / f"f{i}")
        inf = fs.info(i)
        fs.store(i, prefix, 0.075, len(genomes[i]) - 24, counts[i][0], counts[i][1],
                 f"g{i}.fa")
        paths.append(prefix + ".bf")
    fs2 = B.FilterSet.load(paths)
    assert [fs2.filter_id(i) for i in range(2)] == ["alpha", "beta"]
    for i in range(2):
        a = np.zeros(fs.info(i).m // 8, np.uint8)
        b = np.zeros_like(a)
        fs.download(i, a)
        fs2.download(i, b)
        assert np.array_equal(a, b)
        assert fs2.info(i).fpr == fs.info(i).fpr


@pytest.mark.parametrize("ahead", [0, 7, 15])
def test_graph_replay_matches_eager(dev, multi, ahead):
    """Replays of a captured batch (ahead 7 and 15: with the interleaved copy
    made inside the graph; 15: read-per-lane walks) equal the eager batch and the oracle, replay after
    replay, with eager batches between them."""
    D, B = dev
    fs, ofs, buf, off = multi
    dr = D.DeviceReads(buf, off)
    oh, _, ocls, octr = O.classify(ofs, buf, off, paired=True)
    out = D.Outputs(dr.nreads, len(fs), True, False)
    fs.set_tuning(probe_ahead=ahead)
    try:
        for streak in [3, 1]:
            if streak == 1:
                oh, _, ocls, octr = O.classify(ofs, buf, off, paired=True,
                                               p=O.params(O.SIMPLE, 0.15, 1))
            for rep in range(4):
                print(f"[gpu] graph replay ahead={ahead} streak={streak} rep={rep}", flush=True)
                out.hits.fill_(7)
                res = D.classify(fs, dr, B.params(paired=True, streak=streak,
                                                  flags=B.FLAG_GRAPH), out=out)
                assert np.array_equal(res["hits"], oh)
                assert np.array_equal(res["cls"], ocls)
                assert int(res["counters"][1]) == int(octr[1])
                eager = D.classify(fs, dr, B.params(paired=True, streak=streak))
                assert np.array_equal(eager["summary"], res["summary"])
    finally:
        fs.set_tuning()


def test_autotune_keeps_results(dev):
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(3, 7)
    genomes = [synth.codes_to_ascii(synth.random_codes(rng, 150_000)) for _ in range(2)]
    longs, _ = synth.long_reads(rng, genomes, 40, median=3000, hi=12_000)
    shorts, _, _ = synth.fixed_len_reads(rng, genomes, 2000, 150, 0.01, 0.7)
    buf, off = O.pack_reads([bytes(x) for x in longs + list(shorts)])
    fs, _ = D.build_filters(genomes, 25, 0.05)
    ofs = oracle_filters_like(fs, B)
    oh, _, ocls, octr = O.classify(ofs, buf, off)
    dr = D.DeviceReads(buf, off)
    out = D.Outputs(dr.nreads, len(fs), False, False)
    print("[gpu] autotune", flush=True)
    best = fs.autotune(dr.struct(), B.params(), out.struct(),
                       int(torch.cuda.current_stream().cuda_stream))
    assert best["walk_ms"] > 0
    assert fs.get_tuning()["wave_min_len"] == best["wave_min_len"]
    res = D.classify(fs, dr, B.params())
    assert np.array_equal(res["hits"], oh) and np.array_equal(res["cls"], ocls)
    assert int(res["counters"][1]) == int(octr[1])


def test_dust(dev, multi):
    """-D: the sdust pre-pass on the device and the walks that skip its
    windows, against the oracle, on reads with planted low-complexity runs."""
    D, B = dev
    fs, ofs, buf, off = multi
    rng = np.random.default_rng(12)
    reads = [bytes(buf[int(off[i]):int(off[i + 1])]) for i in range(len(off) - 1)]
    units = [b"A", b"AC", b"CAG", b"ACGT", b"T"]
    planted = []
    for i, r in enumerate(reads[:400]):
        a = int(rng.integers(0, max(1, len(r))))
        u = units[i % len(units)]
        planted.append(r[:a] + u * int(rng.integers(5, 40)) + r[a:])
    planted += [b"A" * 300, b"AC" * 150 + b"N" + b"G" * 60, b"ACGTTGCA" + b"N" + b"A" * 40]
    if len(planted) % 2:
        planted.append(b"CA" * 40)
    pbuf, poff = O.pack_reads(planted)
    # windows above 64 too (the scratch grows as W^2; sdust's work per base
    # on repetitive sequence grows with it, on the host as on the device)
    for dust in [(20, 64), (10, 32), (20, 100)]:
        check_batch(D, B, fs, ofs, pbuf, poff, dust=dust)
        check_batch(D, B, fs, ofs, pbuf, poff, dust=dust, wave_min=26)
    check_batch(D, B, fs, ofs, pbuf, poff, dust=(20, 64), paired=True)
    check_batch(D, B, fs, ofs, pbuf, poff, dust=(20, 64), mode=O.ORDERED)
    for scorer, thr in [(O.HARMONIC, 0.15), (O.LENGTH, 30.0), (O.BINOMIAL, 1e-10)]:
        check_batch(D, B, fs, ofs, pbuf, poff, dust=(20, 64), scorer=scorer, thr=thr)
    check_batch(D, B, fs, ofs, pbuf, poff, dust=(20, 64), mode=O.BESTHIT)
    check_batch(D, B, fs, ofs, pbuf, poff, dust=(20, 64), mode=O.SCORES, scorer=O.LENGTH,
                thr=30.0)


def test_binomial_thresholds_follow_uploads(dev):
    """BINOMIAL thresholds depend on each filter's FPR ((popcount / m)^h):
    after new bits are uploaded into a filter, the next batch must use
    thresholds made for the new FPR (checked against the oracle on the new
    bits)."""
    D, B = dev
    rng = np.random.default_rng(31)
    genomes = small_genomes(rng, 2, 60_000)
    fs, _ = D.build_filters(genomes, 25, 0.075)
    ofs = oracle_filters_like(fs, B)
    from biobloom_amd import synth
    reads, _, _ = synth.fixed_len_reads(synth.rng_for(9), genomes, 600, 150, 0.03, 0.5)
    buf, off = O.pack_reads([bytes(r) for r in reads])
    check_batch(D, B, fs, ofs, buf, off, scorer=O.BINOMIAL, thr=1e-10)
    # a denser filter 0: its bits OR'ed with 10% random bits (a higher FPR)
    extra = np.packbits(rng.random(ofs[0].bits.size * 8) < 0.1, bitorder="little")
    denser = ofs[0].bits | extra.astype(ofs[0].bits.dtype)
    fs.upload(0, denser)
    assert fs.info(0).fpr > ofs[0].fpr
    ofs2 = oracle_filters_like(fs, B)
    assert ofs2[0].fpr == fs.info(0).fpr
    print("[gpu] binomial after upload", flush=True)
    check_batch(D, B, fs, ofs2, buf, off, scorer=O.BINOMIAL, thr=1e-10)
    check_batch(D, B, fs, ofs2, buf, off, scorer=O.BINOMIAL, thr=1e-10, mode=O.BESTHIT)


def test_2bit_reads_classify_like_ascii(dev, multi):
    """BBT_READS_2BIT batches (codes + N mask, unpacked on the device) give the
    ASCII batch's hits, scores, classes, summary and counters, against the
    oracle on the ASCII text: lower case, N and other bytes included, every
    mode and scorer, both paths, -D and paired end."""
    D, B = dev
    from test_ingest_host import pack_model
    fs, ofs, buf, off = multi
    rng = np.random.default_rng(21)
    reads = [bytes(buf[int(off[i]):int(off[i + 1])]) for i in range(len(off) - 1)]
    odd = []
    for i, r in enumerate(reads[:600]):
        a = bytearray(r)
        for j in rng.integers(0, max(1, len(a)), 3) if len(a) else []:
            a[j] = rng.choice(list(b"acgtnNRY-"))
        odd.append(bytes(a))
    reads = odd + reads[600:]
    if len(reads) % 2:
        reads.append(b"ACGT" * 40)
    pbuf, poff = O.pack_reads(reads)
    data, cb = D.pack_2bit(pbuf, poff)
    want_cb, want = pack_model(reads)
    assert cb == want_cb and np.array_equal(data, want)
    da, dp = D.DeviceReads(pbuf, poff), D.DeviceReads.packed(pbuf, poff)
    assert dp.seq.numel() * 8 < da.seq.numel() * 3 + 64
    cases = [dict(), dict(scorer=O.HARMONIC, thr=0.2), dict(scorer=O.BINOMIAL, thr=1e-10),
             dict(mode=O.BESTHIT, scorer=O.LENGTH, thr=30.0), dict(mode=O.ORDERED),
             dict(paired=True), dict(dust=(20, 64)), dict(mode=O.SCORES)]
    for c in cases:
        mode, scorer, thr = c.get("mode", 0), c.get("scorer", 0), c.get("thr", 0.15)
        paired, dust = c.get("paired", False), c.get("dust")
        dt, dw = dust if dust else (20, 64)
        want_scores = mode in (O.BESTHIT, O.SCORES)
        oh, osc, ocls, octr = O.classify(ofs, pbuf, poff, paired, False, mode,
                                         O.params(scorer, thr, 3, bool(dust), dt, dw),
                                         want_scores=want_scores)
        for path in PATHS:
            print(f"[gpu] 2-bit reads {c} path={path}", flush=True)
            p = B.params(mode, scorer, thr, 3, paired, False, path, dust=bool(dust),
                         dust_t=dt, dust_w=dw)
            ra, rp = D.classify(fs, da, p), D.classify(fs, dp, p)
            for key in ("hits", "cls", "summary", "counters"):
                assert np.array_equal(ra[key], rp[key]), (c, key)
            if want_scores:
                assert np.array_equal(ra["scores"].view(np.uint64), rp["scores"].view(np.uint64))
                assert np.array_equal(rp["scores"].view(np.uint64), osc.view(np.uint64))
            assert np.array_equal(rp["hits"], oh) and np.array_equal(rp["cls"], ocls)
            assert int(rp["counters"][1]) == int(octr[1])


@pytest.mark.parametrize("ahead", [1, 2, 3, 4, 5, 6, 7, 12, 15])
def test_probe_unroll_keeps_results(dev, multi, ahead):
    """probe_ahead (bit 0: a contains() call's lookups issued together; bit
    1: the next window's lookups issued before the pending call's bits are
    taken; bit 2: read-major walks over the filters interleaved byte by byte,
    which this set of one geometry takes; bit 3 with bit 2: one read per lane
    against all four filters, for SIMPLE evalRead in STD mode) changes no hit,
    score, class or reference counter; it only issues more lookups."""
    D, B = dev
    fs, ofs, buf, off = multi
    try:
        for refill in (1, 32):
            fs.set_tuning(refill_min=refill, probe_ahead=ahead)
            check_batch(D, B, fs, ofs, buf, off, dust=(20, 64))
            for scorer, thr in [(O.SIMPLE, 0.15), (O.LENGTH, 40.0), (O.BINOMIAL, 1e-10)]:
                check_batch(D, B, fs, ofs, buf, off, scorer=scorer, thr=thr)
            check_batch(D, B, fs, ofs, buf, off, paired=True)
            check_batch(D, B, fs, ofs, buf, off, mode=O.SCORES, scorer=O.HARMONIC, thr=0.15)
    finally:
        fs.set_tuning()


def test_tuning_table_carries_knobs(dev, tmp_path):
    """bbt_autotune records its choice under the batch's geometry; another
    filter set of the same geometry without knobs of its own takes it, also
    after a save and load; results stay the oracle's."""
    D, B = dev
    from biobloom_amd import synth
    rng = synth.rng_for(3, 9)
    genomes = [synth.codes_to_ascii(synth.random_codes(rng, 120_000)) for _ in range(2)]
    longs, _ = synth.long_reads(rng, genomes, 30, median=2500, hi=8000)
    shorts, _, _ = synth.fixed_len_reads(rng, genomes, 1500, 150, 0.01, 0.7)
    buf, off = O.pack_reads([bytes(x) for x in longs + list(shorts)])
    fs, _ = D.build_filters(genomes, 25, 0.05)
    dr = D.DeviceReads(buf, off)
    out = D.Outputs(dr.nreads, len(fs), False, False)
    p = B.params()
    print("[gpu] tuning table", flush=True)
    best = fs.autotune(dr.struct(), p, out.struct(), int(torch.cuda.current_stream().cuda_stream))
    got = fs.tuning_lookup(dr.struct(), p)
    assert got == {k: best[k] for k in got}
    path = tmp_path / "tune.tsv"
    B.tuning_save(path)
    B.tuning_load(path)
    fs2, _ = D.build_filters(genomes, 25, 0.05)
    assert fs2.tuning_lookup(dr.struct(), p) == got
    assert fs2.tuning_lookup(dr.struct(), B.params(paired=True)) is None
    ofs = oracle_filters_like(fs2, B)
    oh, _, ocls, octr = O.classify(ofs, buf, off)
    res = D.classify(fs2, dr, p)
    assert np.array_equal(res["hits"], oh) and np.array_equal(res["cls"], ocls)
    assert int(res["counters"][1]) == int(octr[1])
