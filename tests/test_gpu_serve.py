"""Serving loop (bbt_server): every request's results are the same whether it
is classified alone or merged with others into shared launches, under a
device-memory budget that forces requests to wait, in plain and batched
modes, single- and paired-end; and the server's argument checks."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    from biobloom_amd import synth
    B.load()
    rng = np.random.default_rng(7)
    genomes = [synth.codes_to_ascii(rng.integers(0, 4, size=400_000, dtype=np.uint8))
               for _ in range(3)]
    fs, _ = D.build_filters(genomes, 25, 0.075, ids=["g0", "g1", "g2"])
    return torch, D, B, synth, fs, genomes


def make_requests(synth, genomes, n, seed, paired=False):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        nr = int(rng.integers(20, 600)) * (2 if paired else 1)
        items = []
        for j in range(nr):
            if rng.random() < 0.2 and not paired:
                L = int(rng.integers(200, 3000))
            else:
                L = 150
            if rng.random() < 0.6:
                g = genomes[int(rng.integers(0, len(genomes)))]
                a = int(rng.integers(0, len(g) - L))
                items.append(g[a:a + L].copy())
            else:
                items.append(synth.codes_to_ascii(rng.integers(0, 4, size=L, dtype=np.uint8)))
        buf, off = synth.pack(items)
        reqs.append((buf, off))
    return reqs


def alone(D, B, fs, buf, off, params, counters=False):
    dr = D.DeviceReads(buf, off)
    res = D.classify(fs, dr, params)
    if counters:
        return res["hits"], res["cls"], res["counters"]
    return res["hits"], res["cls"]


def run_server(srv, reqs):
    ids = [srv.submit(buf, off) for buf, off in reqs]
    srv.drain(120.0)
    out = []
    for rid, (buf, off) in zip(ids, reqs):
        h, c, st = srv.result(rid, len(off) - 1)
        srv.release(rid)
        out.append((h, c, st))
    return out


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_alone_equals_mixed(env, mode):
    torch, D, B, synth, fs, genomes = env
    params = B.params()
    reqs = make_requests(synth, genomes, 40, 1)
    m = B.SERVE_PLAIN if mode == "plain" else B.SERVE_BATCHED
    print(f"[gpu] server {mode}: 40 requests", flush=True)
    srv = B.Server(fs, params, slots=3, mode=m, device_bytes=64 << 20,
                   max_wave_reads=1 << 14, max_wave_bases=8 << 20)
    got = run_server(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["requests_done"] == len(reqs)
    probes = lookups = 0
    for (buf, off), (h, c, info) in zip(reqs, got):
        wh, wc, ctr = alone(D, B, fs, buf, off, params, counters=True)
        assert np.array_equal(h, wh) and np.array_equal(c, wc)
        assert info["done_us"] >= info["admit_us"] >= info["submit_us"]
        probes += int(ctr[0])
        lookups += int(ctr[1])
    # the reference walk's contains() calls and lookups do not depend on how
    # requests were merged into launches
    assert st["probes"] == probes and st["lookups"] == lookups > 0
    if mode == "plain":
        assert st["mean_wave_requests"] == 1.0
    else:
        # requests submitted back to back are merged into shared launches
        assert st["mean_wave_requests"] > 1.0
        assert st["max_waves_in_flight"] >= 1


def test_results_match_oracle(env):
    torch, D, B, synth, fs, genomes = env
    params = B.params()
    reqs = make_requests(synth, genomes, 6, 2)
    ofs = []
    for i in range(len(fs)):
        inf = fs.info(i)
        f = O.Filter(inf.m, inf.h, inf.k)
        fs.download(i, f.bits)
        ofs.append(f)
    srv = B.Server(fs, params, slots=2, device_bytes=32 << 20, max_wave_reads=1 << 13,
                   max_wave_bases=4 << 20)
    got = run_server(srv, reqs)
    srv.close()
    for (buf, off), (h, c, _) in zip(reqs, got):
        oh, _, ocls, _ = O.classify(ofs, buf, off)
        assert np.array_equal(h, oh) and np.array_equal(c, ocls)


def test_device_memory_admission(env):
    """An arena that holds about one launch: requests wait for memory held by
    launches in flight, and all finish with the same results."""
    torch, D, B, synth, fs, genomes = env
    params = B.params()
    reqs = make_requests(synth, genomes, 24, 3)
    wave_reads, wave_bases = 1 << 12, 2 << 20
    need = wave_bases + 8 * wave_reads + wave_reads * len(fs) + 4 * wave_reads + 4096
    srv = B.Server(fs, params, slots=4, device_bytes=need, max_wave_reads=wave_reads,
                   max_wave_bases=wave_bases)
    got = run_server(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["requests_done"] == len(reqs)
    assert st["arena_high_water"] <= st["arena_bytes"]
    for (buf, off), (h, c, _) in zip(reqs, got):
        wh, wc = alone(D, B, fs, buf, off, params)
        assert np.array_equal(h, wh) and np.array_equal(c, wc)


def test_paired_requests(env):
    torch, D, B, synth, fs, genomes = env
    params = B.params(paired=True)
    reqs = make_requests(synth, genomes, 12, 4, paired=True)
    srv = B.Server(fs, params, slots=2, device_bytes=64 << 20, max_wave_reads=1 << 14,
                   max_wave_bases=8 << 20)
    got = run_server(srv, reqs)
    srv.close()
    for (buf, off), (h, c, _) in zip(reqs, got):
        wh, wc = alone(D, B, fs, buf, off, params)
        assert np.array_equal(h, wh) and np.array_equal(c, wc)
        assert len(c) == (len(off) - 1) // 2


def test_server_argument_checks(env):
    torch, D, B, synth, fs, genomes = env
    with pytest.raises(B.BbtError, match="device_bytes"):
        B.Server(fs, B.params(), device_bytes=4096, max_wave_reads=1 << 14,
                 max_wave_bases=8 << 20)
    with pytest.raises(B.BbtError, match="configuration"):
        B.Server(fs, B.params(), slots=0)
    srv = B.Server(fs, B.params(paired=True), slots=1, device_bytes=16 << 20,
                   max_wave_reads=1024, max_wave_bases=1 << 20)
    buf, off = synth.pack([np.frombuffer(b"ACGT" * 40, np.uint8)] * 3)
    with pytest.raises(B.BbtError, match="odd"):
        srv.submit(buf, off)
    big, boff = synth.pack([np.frombuffer(b"ACGT" * 40, np.uint8)] * 2000)
    with pytest.raises(B.BbtError, match="larger than one wave"):
        srv.submit(big, boff)
    srv.close()


def test_host_memory_admission(env):
    """A pinned host arena that holds about one request: submit waits until
    launched requests release it; every result still equals the alone run."""
    torch, D, B, synth, fs, genomes = env
    params = B.params()
    reqs = make_requests(synth, genomes, 16, 5)
    biggest = max(int(off[-1]) for _, off in reqs)
    srv = B.Server(fs, params, slots=2, device_bytes=64 << 20, max_wave_reads=1 << 14,
                   max_wave_bases=biggest + 4096, host_bytes=biggest + 4096)
    got = run_server(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["requests_done"] == len(reqs)
    for (buf, off), (h, c, _) in zip(reqs, got):
        wh, wc = alone(D, B, fs, buf, off, params)
        assert np.array_equal(h, wh) and np.array_equal(c, wc)


def test_large_request_copy(env):
    """A request of 9 MB of bases goes into the pinned arena through the
    server's split copy (requests of 4 MiB and more): its results, and those
    of a small request behind it, equal the requests alone."""
    torch, D, B, synth, fs, genomes = env
    params = B.params()
    rng = np.random.default_rng(17)
    g = genomes[1]
    starts = rng.integers(0, len(g) - 150, size=60_000)
    items = [g[a:a + 150].copy() for a in starts[:30_000]]
    items += [synth.codes_to_ascii(rng.integers(0, 4, size=150, dtype=np.uint8))
              for _ in range(30_000)]
    big = synth.pack(items)
    small = make_requests(synth, genomes, 1, 18)[0]
    print("[gpu] server: one 9 MB request and one small one", flush=True)
    srv = B.Server(fs, params, slots=2, device_bytes=256 << 20, max_wave_reads=1 << 17,
                   max_wave_bases=16 << 20)
    got = run_server(srv, [big, small])
    srv.close()
    for (buf, off), (h, c, _) in zip([big, small], got):
        wh, wc = alone(D, B, fs, buf, off, params)
        assert np.array_equal(h, wh) and np.array_equal(c, wc)
    assert int(got[0][1].size) == 60_000 and (got[0][1][:30_000] == 1).mean() > 0.9


@pytest.mark.parametrize("mode,scorer,thr", [("besthit", 0, 0.15), ("scores", 2, 0.15),
                                             ("besthit", 3, 1e-10)])
def test_score_modes_served(env, mode, scorer, thr):
    """BESTHIT and SCORES requests through the server: hits, classes and the
    evalScore values of every request equal the request classified alone and
    the oracle's (BioBloomClassifier.cpp:1194-1237), merged into shared
    launches."""
    torch, D, B, synth, fs, genomes = env
    m = B.BESTHIT if mode == "besthit" else B.SCORES
    params = B.params(mode=m, scorer=scorer, threshold=thr)
    reqs = make_requests(synth, genomes, 10, 71 + scorer)
    ofs = []
    for i in range(len(fs)):
        inf = fs.info(i)
        f = O.Filter(inf.m, inf.h, inf.k)
        fs.download(i, f.bits)
        ofs.append(f)
    srv = B.Server(fs, params, slots=2, device_bytes=256 << 20, max_wave_reads=1 << 14,
                   max_wave_bases=16 << 20)
    ids = [srv.submit(buf, off) for buf, off in reqs]
    srv.drain(120.0)
    waves = set()
    for rid, (buf, off) in zip(ids, reqs):
        n = len(off) - 1
        h, c, st = srv.result(rid, n)
        sc = srv.scores(rid, n)
        srv.release(rid)
        waves.add(st["wave"])
        print(f"[gpu] served {mode} scorer={scorer} request of {n} reads", flush=True)
        dr = D.DeviceReads(buf, off)
        res = D.classify(fs, dr, params)
        assert np.array_equal(h, res["hits"]) and np.array_equal(c, res["cls"])
        assert np.array_equal(sc.view(np.uint64), res["scores"].view(np.uint64))
        oh, osc, ocls, _ = O.classify(ofs, buf, off, False, False, m,
                                      O.params(scorer, thr, 3), want_scores=True)
        assert np.array_equal(h, oh) and np.array_equal(c, ocls)
        assert np.array_equal(sc.view(np.uint64), osc.view(np.uint64))
    srv.close()
    assert len(waves) < len(reqs)  # requests shared launches
