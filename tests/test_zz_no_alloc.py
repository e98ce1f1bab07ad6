"""bbt_eval_batch_noalloc through the C ABI: the same results as
bbt_eval_batch (and the oracle) on two calls in a row, the second with a
smaller batch, on both kernel paths, single and paired end, every mode; the
size limit and the rejected options return status codes."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def env():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    from biobloom_amd import synth
    B.load()
    w = synth.workload_multi(scale=0.003, nfilters=3, genome_len=200_000)
    fs, _ = D.build_filters(w["genomes"], 25, 0.075)
    ofs = []
    for i in range(len(fs)):
        inf = fs.info(i)
        f = O.Filter(inf.m, inf.h, inf.k)
        fs.download(i, f.bits)
        ofs.append(f)
    reads = [bytes(r) for r in w["reads"]]
    rng = np.random.default_rng(3)
    longs, _ = synth.long_reads(synth.rng_for(3, 11), w["genomes"], 30, median=1500, hi=6000)
    reads += [bytes(x) for x in longs] + [b"", b"ACGT", b"N" * 90]
    if len(reads) % 2:
        reads.append(bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), 150)))
    return D, B, fs, ofs, reads


def noalloc(B, fs, dr, params, out):
    L = B.load()
    st = L.bbt_eval_batch_noalloc(fs.handle, C.byref(dr.struct()), C.byref(params),
                                  C.byref(out.struct()),
                                  C.c_void_p(int(torch.cuda.current_stream().cuda_stream)))
    return st, (L.bbt_last_error() or b"").decode()


def run_both(D, B, fs, reads, params):
    buf, off = O.pack_reads(reads)
    dr = D.DeviceReads(buf, off)
    want_scores = params.mode in (B.BESTHIT, B.SCORES)
    ref = D.classify(fs, dr, params, want_scores=want_scores)
    out = D.Outputs(dr.nreads, len(fs), bool(params.paired), want_scores)
    out.hits.fill_(9)
    out.cls.fill_(-1)
    out.zero_counts()
    st, msg = noalloc(B, fs, dr, params, out)
    assert st == 0, msg
    torch.cuda.synchronize()
    got = out.host()
    for key in ("hits", "cls", "summary"):
        assert np.array_equal(got[key], ref[key]), key
    # contains() calls, bit lookups and the error flag; the issued-lookup
    # counter follows the load scheduling, which without allocation has no
    # interleaved copies to read (only at least the reference lookups)
    for i in (0, 1, 3):
        assert int(got["counters"][i]) == int(ref["counters"][i]), f"counter {i}"
    assert int(got["counters"][2]) >= int(got["counters"][1])
    if want_scores:
        assert np.array_equal(got["scores"].view(np.uint64), ref["scores"].view(np.uint64))
    return got, buf, off


@pytest.mark.parametrize("path", [0, 1])
def test_noalloc_equals_eval_batch(env, path):
    D, B, fs, ofs, reads = env
    for paired in (False, True):
        for mode, scorer, thr in [(B.STD, B.SIMPLE, 0.15), (B.ORDERED, B.SIMPLE, 0.15),
                                  (B.BESTHIT, B.HARMONIC, 0.15), (B.SCORES, B.LENGTH, 30.0)]:
            p = B.params(mode, scorer, thr, 3, paired, False, path)
            # two calls in a row on the shared scratch, the second smaller
            for n in (len(reads), len(reads) // 3 * 2 // 2 * 2):
                print(f"[gpu] noalloc path={path} paired={paired} mode={mode} n={n}",
                      flush=True)
                got, buf, off = run_both(D, B, fs, reads[:n], p)
                if mode == B.STD:
                    oh, _, ocls, octr = O.classify(ofs, buf, off, paired=paired)
                    assert np.array_equal(got["hits"], oh)
                    assert np.array_equal(got["cls"], ocls)
                    assert int(got["counters"][1]) == int(octr[1])


def test_noalloc_binomial_after_tables_are_built(env):
    D, B, fs, ofs, reads = env
    p = B.params(B.STD, B.BINOMIAL, 1e-7, 3, False, False, 1)
    buf, off = O.pack_reads(reads)
    dr = D.DeviceReads(buf, off)
    out = D.Outputs(dr.nreads, len(fs), False, False)
    st, msg = noalloc(B, fs, dr, p, out)
    assert st == 6 and "not built" in msg
    print("[gpu] noalloc binomial", flush=True)
    run_both(D, B, fs, reads, p)      # bbt_eval_batch builds the tables, then the same
    run_both(D, B, fs, reads[:100], p)


def test_noalloc_limits(env):
    D, B, fs, ofs, reads = env
    buf, off = O.pack_reads(reads[:10])
    dr = D.DeviceReads(buf, off)
    out = D.Outputs(dr.nreads, len(fs), False, False)
    p = B.params()
    p.dust = 1
    p.dust_t, p.dust_w = 20, 64
    st, msg = noalloc(B, fs, dr, p, out)
    assert st == 6 and "dust" in msg
    p = B.params(flags=B.FLAG_GRAPH)
    st, _ = noalloc(B, fs, dr, p, out)
    assert st == 6
    big = dr.struct()
    big.nreads = (1 << 20) + 1   # rejected before any buffer is touched
    st = B.load().bbt_eval_batch_noalloc(fs.handle, C.byref(big), C.byref(B.params()),
                                         C.byref(out.struct()), None)
    assert st == 1 and "1048576" in B.load().bbt_last_error().decode()
