"""Benchmark of the MI355X read-classification path.

One step = one bbt_eval_batch over the whole batch of reads already resident
in HBM: every (read, filter) SeqEval walk, the ResultsManager classes and the
summary counters.  Default workload ("multi"): 10 filters from 20 Mbp
genomes (k=25, FPR 0.075, h=3), 2M read pairs 2x150 bp, paired-end
intersection mode, default SIMPLE scorer (-s 0.15, -r 3).

Prints one JSON line: reads/s of the batched persistent path (value), the
plain per-batch path beside it, k-mer probes/s, the roofline of the walk
kernel and the CPU oracle baseline on a bounded sample.

--workload serve: the serving loop (one 3 Gbp filter + 8 contaminant
filters resident in HBM; requests of 1k-20k reads arrive Poisson at each
--qps); reads/s and p50/p99 request latency of continuous batching
(bbt_server) against the plain per-request launch path.

--workload e2e: the same reads and filters end to end through the
biobloomcategorizer CLI: FASTQ files on disk (plain, and gzip with --e2e-gz)
-> parallel parse -> device batches -> summary.tsv; reads/s over the CLI's
filtering time, with the summary checked against an in-process
bbt_eval_batch of the same reads.

--workload mibf-e2e: the same through the biobloommicategorizer CLI from a
FASTQ file, per-read TSV lines included.

--workload mibf: the multi-index Bloom filter instead (200 x 5 Mbp
references, 3 spaced seeds span 25 weight 20, 2M x 150 bp reads, the
biobloommicategorizer defaults -s 100 -a 1 -m 1 -c 0); one step is one
bbt_mibf_query_batch over all reads; probes are (frame, seed) lookups.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-mer probes/s + reads/s on 1 MI355X (batched vs plain GPU); p50/p99 latency"
HBM_PEAK_GBS = 8000.0
LINE_BYTES = 64  # one random 64 B line per bit lookup


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name, scale):
    from biobloom_amd import synth
    t = time.time()
    if name == "multi":
        w = synth.workload_multi(scale=scale)
        reads = w["reads"]
        buf = np.ascontiguousarray(reads).reshape(-1)
        off = np.arange(reads.shape[0] + 1, dtype=np.uint64) * np.uint64(reads.shape[1])
    elif name == "single":
        w = synth.workload_single(scale=scale)
        reads = w["reads"]
        buf = np.ascontiguousarray(reads).reshape(-1)
        off = np.arange(reads.shape[0] + 1, dtype=np.uint64) * np.uint64(reads.shape[1])
    elif name == "uneven":
        w = synth.workload_uneven(scale=scale)
        buf, off = synth.pack(w["reads_list"])
    else:
        raise SystemExit(f"unknown workload {name}")
    log(f"workload {name}: {len(off) - 1} reads, {int(off[-1])} bases, "
        f"{len(w['genomes'])} genomes ({time.time() - t:.1f}s)")
    return w, buf, off


def time_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def cpu_baseline(fs, buf, off, paired, sample_records, threads, B, gpu_hits):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ofs = []
    for i in range(len(fs)):
        inf = fs.info(i)
        f = O.Filter(inf.m, inf.h, inf.k)
        fs.download(i, f.bits)
        ofs.append(f)
    nreads = sample_records * (2 if paired else 1)
    sbuf = buf[: int(off[nreads])]
    soff = off[: nreads + 1]
    O.classify(ofs, sbuf, soff, paired=paired, nthreads=threads)  # warm
    t = time.time()
    hits, _, _, _ = O.classify(ofs, sbuf, soff, paired=paired, nthreads=threads)
    dt = time.time() - t
    parity = bool(np.array_equal(hits, gpu_hits[:nreads]))
    return nreads / dt, dt, parity


def cpu_baseline_mibf(mf, buf, off, nsample, params, gpu_rows, gpu_sat):
    """Oracle MIBF query (OpenMP over reads) on the first nsample reads,
    against host copies of the device-built bit vector and ID array."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    info = mf.info()
    bv = np.zeros(info.m // 64, np.uint64)
    data = np.zeros(max(1, info.pop), np.uint32)
    mf.download(bv, data)
    seeds = mf.seeds() or O.Kmer(info.span, info.nseeds)
    om = O.Mibf.from_arrays(bv, data, info.m, seeds, info.nids)
    sbuf = buf[: int(off[nsample])]
    soff = off[: nsample + 1]
    reads = [bytes(sbuf[int(soff[i]):int(soff[i + 1])]) for i in range(nsample)]
    kw = dict(score=params.score, allowed_miss=params.allowed_miss,
              multi=params.multi_thresh, min_nonsat=params.min_nonsat,
              max_hits=params.max_hits)
    om.query(reads[: min(1000, nsample)], **kw)  # warm
    t = time.time()
    rows, sat, _ = om.query(reads, **kw)
    dt = time.time() - t
    parity = rows == gpu_rows[:nsample] and sat == gpu_sat[:nsample]
    return nsample / dt, dt, bool(parity)


def cpu_baseline_mibf_e2e(mf, path, width, nsample):
    """biobloommicategorizer's path on the host cores over the first nsample
    reads of the FASTQ file (records of `width` bytes): read the bytes, parse
    them with the oracle's kseq_read restatement (one thread), query the
    oracle MIBF (OpenMP over reads) with the CLI's defaults, and count the
    summary (hits, unique hits, multi, none per ID)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    info = mf.info()
    bv = np.zeros(info.m // 64, np.uint64)
    data = np.zeros(max(1, info.pop), np.uint32)
    mf.download(bv, data)
    seeds = mf.seeds() or O.Kmer(info.span, info.nseeds)
    om = O.Mibf.from_arrays(bv, data, info.m, seeds, info.nids)
    kw = dict(score=1e-10, allowed_miss=info.nseeds - 1, multi=1.0, min_nonsat=0, max_hits=64)
    t0 = time.time()
    with open(path, "rb") as fh:
        text = fh.read(nsample * width)
    sbuf, soff, _ = O.kseq_seqs(text)
    t1 = time.time()
    n = len(soff) - 1
    reads = [bytes(sbuf[int(soff[i]):int(soff[i + 1])]) for i in range(n)]
    rows, _, _ = om.query(reads, **kw)
    t2 = time.time()
    above = np.zeros(info.nids + 1, np.int64)
    unique = np.zeros(info.nids + 1, np.int64)
    multi = none = 0
    for row in rows:
        for h in row:
            above[h[0]] += 1
        if not row:
            none += 1
        elif len(row) == 1:
            unique[row[0][0]] += 1
        else:
            multi += 1
    t3 = time.time()
    omp = os.environ.get("OMP_NUM_THREADS")
    return {"value": n / (t3 - t0), "unit": "reads/s",
            "cores": int(omp) if omp and omp.isdigit() else os.cpu_count(), "kind": "port",
            "sample": f"first {n} reads: read + kseq parse (1 thread) {t1 - t0:.2f}s, MIBF "
                      f"query (OpenMP) {t2 - t1:.2f}s, summary counts {t3 - t2:.2f}s"}


def bench_mibf(args, torch, D, B):
    from biobloom_amd import synth
    t = time.time()
    w = synth.workload_mibf(scale=args.scale)
    reads = w["reads"]
    buf = np.ascontiguousarray(reads).reshape(-1)
    off = np.arange(reads.shape[0] + 1, dtype=np.uint64) * np.uint64(reads.shape[1])
    nreads = reads.shape[0]
    log(f"workload mibf: {len(w['refs'])} references x {len(w['refs'][0])} bp, "
        f"{nreads} reads, seeds {w['seeds']} ({time.time() - t:.1f}s)")
    t = time.time()
    mf = D.build_mibf(w["refs"], list(range(1, len(w["refs"]) + 1)), w["names"], w["seeds"],
                      occupancy=0.5)
    info = mf.info()
    log(f"MIBF built on device ({time.time() - t:.2f}s): m={info.m} pop={info.pop} "
        f"saturated={info.pop_saturated} ids={info.nids}")
    del w
    dr = D.DeviceReads(buf, off)
    # biobloommicategorizer defaults: -s 100 (score 1e-10), -a 1 (allowed
    # misses = 3 - 1), -m 1, -c 0 (BioBloomMICategorizer.cpp:132, Options)
    params = B.mibf_params(score=1e-10, multi_thresh=1.0, allowed_miss=info.nseeds - 1,
                           min_nonsat=0, max_hits=8)
    out = D.MibfOutputs(nreads, params.max_hits)
    stream = int(torch.cuda.current_stream().cuda_stream)
    rs, os_ = dr.struct(), out.struct()

    def step():
        mf.query_batch(rs, params, os_, stream)

    out.counters.zero_()
    ms = time_steps(step, args.steps, args.warmup, torch)
    out.counters.zero_()
    step()
    torch.cuda.synchronize()
    rows, sat, ctr, nh = out.host()
    lookups, gathers = int(ctr[0]), int(ctr[1])
    assert int(ctr[3]) == 0, "records not evaluated"
    classified = sum(1 for r in rows if r)
    reads_per_s = nreads / (ms / 1e3)
    alg_bytes = (lookups + gathers) * LINE_BYTES
    achieved = alg_bytes / (ms / 1e3) / 1e9
    log(f"query: {ms:.3f} ms/step, {reads_per_s / 1e6:.2f} M reads/s, "
        f"{lookups / (ms / 1e3) / 1e9:.2f} G lookups/s, classified {classified}/{nreads}")
    gather = {}
    if not args.no_gather:
        lps, _ = B.measure_gather(8 << 30, 64, stream)
        gather = {"hbm_8GiB": {"bytes": 8 << 30, "lines_per_s": lps,
                               "gbs_at_64B": lps * LINE_BYTES / 1e9}}
        log(f"random gather over 8192 MiB: {lps / 1e9:.1f} G lines/s")
    cpu = None
    if not args.no_cpu:
        nsample = min(args.cpu_sample, nreads)
        v, dt, parity = cpu_baseline_mibf(mf, buf, off, nsample, params, rows, sat)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": v, "unit": "reads/s", "cores": threads, "kind": "port",
               "sample": f"first {nsample} reads of the workload, {dt:.2f}s",
               "parity_with_gpu": parity}
        log(f"cpu oracle: {v / 1e6:.4f} M reads/s on {threads} threads, parity={parity}")
    traffic = load_traffic("mibf")
    line = {
        "metric": METRIC, "value": reads_per_s, "unit": "reads/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "multi-index Bloom filter: 200 x 5 Mbp references (5% shared "
                               "10 kbp segments), 3 spaced seeds span 25 weight 20, "
                               "occupancy 0.5, 32-bit IDs, 2M x 150 bp reads",
                   "scale": args.scale, "reads": nreads, "ids": info.nids, "m": info.m,
                   "pop": info.pop},
        "probes_per_s": lookups / (ms / 1e3),
        "probes_per_step": lookups,
        "id_gathers_per_step": gathers,
        "classified_reads": classified,
        "roofline": {"bound": "hbm", "kernel": "k_mibf_query", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_gbs": (traffic["bytes_per_launch"] / (ms / 1e3) / 1e9
                                     if traffic else None),
                     "traffic_frac": (traffic["bytes_per_launch"] / (ms / 1e3) / 1e9
                                      / HBM_PEAK_GBS if traffic else None),
                     "algorithmic_bytes": alg_bytes, "random_gather": gather or None,
                     "frac_of_random_gather": (achieved / gather["hbm_8GiB"]["gbs_at_64B"]
                                               if gather else None)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def serve_run(B, fs, params, pool, mode, qps, seconds, max_requests, seed, want):
    """Poisson arrivals at qps submitted in real time; returns latency and
    throughput figures and whether every result equals the request alone."""
    plain = mode == "plain"
    srv = B.Server(fs, params, slots=1 if plain else 4,
                   mode=B.SERVE_PLAIN if plain else B.SERVE_BATCHED,
                   device_bytes=4 << 30, max_wave_reads=1 << 18, max_wave_bases=512 << 20)
    n = max(20, min(max_requests, int(qps * seconds)))
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, size=n))
    ids = []
    t0 = time.perf_counter()
    for i in range(n):
        dt = arrivals[i] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        pi = i % len(pool)
        ids.append((srv.submit(*pool[pi]), pi))
    srv.drain(900.0)
    lat, reads, first, last, equal = [], 0, None, 0.0, True
    for rid, pi in ids:
        nr = len(pool[pi][1]) - 1
        h, c, st = srv.result(rid, nr)
        srv.release(rid)
        lat.append((st["done_us"] - st["submit_us"]) / 1e3)
        reads += nr
        first = st["submit_us"] if first is None else min(first, st["submit_us"])
        last = max(last, st["done_us"])
        equal &= bool(np.array_equal(h, want[pi][0]) and np.array_equal(c, want[pi][1]))
    stats = srv.stats()
    srv.close()
    p50, p99 = np.percentile(lat, [50, 99])
    return {"qps": qps, "requests": n, "reads": reads, "p50_ms": float(p50),
            "p99_ms": float(p99), "mean_ms": float(np.mean(lat)),
            "reads_per_s": reads / ((last - first) / 1e6),
            "offered_reads_per_s": reads / arrivals[-1] if arrivals[-1] > 0 else None,
            "waves": stats["waves"], "mean_wave_requests": stats["mean_wave_requests"],
            "submit_ms": stats["submit_ms"], "pack_ms": stats["pack_ms"],
            "device_ms": stats["device_ms"], "span_ms": (last - first) / 1e3,
            "max_waves_in_flight": stats["max_waves_in_flight"],
            "probes": stats["probes"], "lookups": stats["lookups"],
            "lookups_per_s": stats["lookups"] / ((last - first) / 1e6),
            "equal_to_alone": equal}


def bench_serve(args, torch, D, B):
    from biobloom_amd import synth
    t = time.time()
    w = synth.workload_serve(scale=args.scale)
    pool = w["requests"]
    log(f"workload serve: genome {len(w['genomes'][0])} bp + {len(w['genomes']) - 1} "
        f"contaminants, request pool {len(pool)} (mean "
        f"{np.mean([len(o) - 1 for _, o in pool]):.0f} reads) ({time.time() - t:.1f}s)")
    t = time.time()
    fs, _ = D.build_filters(w["genomes"], w["k"], w["fpr"])
    torch.cuda.synchronize()
    big = fs.info(0)
    log(f"filters built on device ({time.time() - t:.1f}s): big m={big.m} "
        f"({big.m / 8 / 2**30:.2f} GiB), h={big.h}")
    del w
    params = B.params()
    # each pool request alone through bbt_eval_batch: the results every
    # served copy must reproduce
    want = []
    t = time.time()
    for buf, off in pool:
        r = D.classify(fs, D.DeviceReads(buf, off), params)
        want.append((r["hits"], r["cls"]))
    log(f"pool classified alone ({time.time() - t:.1f}s)")
    qps_list = [float(x) for x in args.qps.split(",")]
    runs = {"plain": [], "batched": []}
    for mode in ("plain", "batched"):
        for i, qps in enumerate(qps_list):
            r = serve_run(B, fs, params, pool, mode, qps, args.serve_seconds,
                          args.max_requests, 1000 + i, want)
            runs[mode].append(r)
            log(f"{mode} qps={qps:g}: {r['reads_per_s'] / 1e6:.2f} M reads/s, p50 "
                f"{r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, waves {r['waves']}, "
                f"{r['mean_wave_requests']:.2f} req/wave, equal={r['equal_to_alone']}; "
                f"span {r['span_ms']:.0f} ms: submit {r['submit_ms']:.0f}, pack "
                f"{r['pack_ms']:.0f}, device {r['device_ms']:.0f} ms")
    top = runs["batched"][-1]
    # the random-line rate over a buffer of the big filter's size: what the
    # saturated server's lookups (one random 64 B line each) are measured against
    gbytes = 1 << max(12, int(np.ceil(np.log2(max(big.m // 8, 4096)))))
    glps, _ = B.measure_gather(gbytes, 64, 0)
    log(f"random gather over {gbytes >> 20} MiB: {glps / 1e9:.1f} G lines/s; saturated "
        f"server: {top['lookups_per_s'] / 1e9:.1f} G lookups/s")
    line = {
        "metric": METRIC, "value": top["reads_per_s"], "unit": "reads/s", "n_gpus": 1,
        "steps": top["requests"], "warmup": 0,
        "ms_per_step": top["mean_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "serving: 3 Gbp genome (50% 300 bp repeats, 10% divergence) "
                               "+ 8 x 20 Mbp contaminant filters, k=25 FPR 0.075; Poisson "
                               "requests of 1k-20k reads, 80% 150 bp / 20% log-normal long",
                   "scale": args.scale, "qps": qps_list, "filters": len(fs),
                   "big_filter_bits": big.m},
        "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
        "saturated_lookups_per_s": top["lookups_per_s"],
        "random_gather": {"bytes": gbytes, "lines_per_s": glps},
        "saturated_frac_of_random_gather": top["lookups_per_s"] / glps,
        "serving": runs,
        "alone_equals_mixed": all(r["equal_to_alone"] for m in runs.values() for r in m),
    }
    print(json.dumps(line), flush=True)


def write_bgzf(f, data, level=1, chunk=65280):
    """data as BGZF (bgzip's blocked gzip): members of at most `chunk` bytes
    of text with the 'BC' size subfield, then the empty end-of-file member."""
    import struct
    import zlib
    for i in range(0, len(data), chunk):
        part = data[i:i + chunk]
        c = zlib.compressobj(level, zlib.DEFLATED, -15)
        payload = c.compress(part) + c.flush()
        f.write(b"\x1f\x8b\x08\x04\0\0\0\0\0\xff" + struct.pack("<H", 6) + b"BC" +
                struct.pack("<HH", 2, len(payload) + 25) + payload +
                struct.pack("<II", zlib.crc32(part), len(part)))
    f.write(b"\x1f\x8b\x08\x04\0\0\0\0\0\xff\x06\0BC\x02\0\x1b\0\x03\0\0\0\0\0\0\0\0\0")


def write_fastq(path, reads, mate, gz=False):
    """FASTQ with fixed-width names @r<9 digits>/<mate>, qualities 'I';
    gz: False, True (one gzip member) or "bgzf"."""
    import gzip
    n, L = reads.shape
    w = 2 * L + 18
    a = np.empty((n, w), np.uint8)
    idx = np.arange(n, dtype=np.int64)
    a[:, 0] = ord("@")
    a[:, 1] = ord("r")
    for d in range(9):
        a[:, 2 + d] = 48 + (idx // 10 ** (8 - d)) % 10
    a[:, 11] = ord("/")
    a[:, 12] = ord(str(mate))
    a[:, 13] = 10
    a[:, 14:14 + L] = reads
    a[:, 14 + L] = 10
    a[:, 15 + L] = ord("+")
    a[:, 16 + L] = 10
    a[:, 17 + L:17 + 2 * L] = ord("I")
    a[:, 17 + 2 * L] = 10
    if gz == "bgzf":
        with open(path, "wb") as f:
            write_bgzf(f, memoryview(a.reshape(-1)))
    elif gz:
        with gzip.open(path, "wb", compresslevel=1) as f:
            f.write(a.tobytes())
    else:
        a.tofile(path)
    return a.nbytes


def bench_e2e(args, torch, D, B):
    import subprocess
    import tempfile
    w, buf, off = make_workload("multi", args.scale)
    t = time.time()
    fs, counts = D.build_filters(w["genomes"], w["k"], w["fpr"])
    torch.cuda.synchronize()
    log(f"filters built on device ({time.time() - t:.2f}s)")
    reads = w["reads"]
    npairs = reads.shape[0] // 2
    # the in-process result the CLI's summary must equal
    dr = D.DeviceReads(buf, off)
    res = D.classify(fs, dr, B.params(paired=True))
    want = fs.summary_text(res["summary"], npairs)
    del dr
    cli = os.path.join(ROOT, "biobloom_amd", "biobloomcategorizer")
    if args.e2e_keep:
        tmp = os.path.abspath(args.e2e_keep)
        os.makedirs(tmp, exist_ok=True)
    else:
        tmp = tempfile.mkdtemp(prefix="bbt_e2e_", dir=os.environ.get("TMPDIR", "/tmp"))
    bfs = []
    for i in range(len(fs)):
        inf = fs.info(i)
        pre = os.path.join(tmp, f"g{i}")
        glen = len(w["genomes"][i])
        fs.store(i, pre, w["fpr"], glen - w["k"] + 1, counts[i][0], counts[i][1], f"g{i}.fa")
        bfs.append(pre + ".bf")
    runs = []
    cpu = None
    variants = [("fastq", False)] + ([("fastq.gz", True), ("fastq.bgz", "bgzf")]
                                     if args.e2e_gz else [])
    for ext, gz in variants:
        t = time.time()
        f1, f2 = os.path.join(tmp, f"r1.{ext}"), os.path.join(tmp, f"r2.{ext}")
        nbytes = write_fastq(f1, reads[0::2], 1, gz) + write_fastq(f2, reads[1::2], 2, gz)
        disk = os.path.getsize(f1) + os.path.getsize(f2)
        log(f"wrote {ext} pair files: {nbytes / 1e6:.0f} MB text, {disk / 1e6:.0f} MB on "
            f"disk ({time.time() - t:.1f}s)")
        for threads in args.e2e_threads:
            prefix = os.path.join(tmp, f"out_{ext}_{threads}")
            cmd = ([cli, "-e", "-p", prefix, "-f", " ".join(bfs), "-t", str(threads)]
                   + args.e2e_args.split() + [f1, f2])
            t = time.time()
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            wall = time.time() - t
            if r.returncode != 0:
                raise SystemExit(f"biobloomcategorizer failed: {r.stderr[-2000:]}")
            tl = [x for x in r.stderr.splitlines() if x.startswith("Filtering time:")][0]
            ft = float(tl.split()[2])
            log(tl)
            got = open(prefix + "_summary.tsv").read()
            item = {"input": ext, "threads": threads, "filtering_s": ft, "wall_s": wall,
                    "reads_per_s": 2 * npairs / ft, "pairs_per_s": npairs / ft,
                    "text_gb_per_s": nbytes / ft / 1e9, "summary_equal": got == want}
            runs.append(item)
            log(f"e2e {ext} -t {threads}: filtering {ft:.2f} s (wall {wall:.2f} s), "
                f"{item['reads_per_s'] / 1e6:.2f} M reads/s, {item['text_gb_per_s']:.2f} GB/s "
                f"of FASTQ, summary equal={item['summary_equal']}")
        if not gz and not args.no_cpu and cpu is None:
            cpu = cpu_baseline_e2e(fs, f1, f2, 2 * reads.shape[1] + 18, min(npairs, 200_000),
                                   min(16, os.cpu_count() or 1))
            log(f"cpu e2e baseline: {cpu['value'] / 1e6:.3f} M reads/s ({cpu['sample']})")
        if args.e2e_keep:
            log("kept: " + " ".join(cmd))
        else:
            os.remove(f1)
            os.remove(f2)
    best = max((r for r in runs if r["input"] == "fastq"), key=lambda r: r["reads_per_s"])
    line = {
        "metric": METRIC, "value": best["reads_per_s"], "unit": "reads/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": best["filtering_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "end to end: biobloomcategorizer -e on two FASTQ files "
                               "(2M pairs 2x150 bp) with 10 filters from 20 Mbp genomes, "
                               "k=25 FPR 0.075; parse + device batches + summary.tsv",
                   "scale": args.scale, "pairs": npairs},
        "e2e": runs,
        "summary_equal": all(r["summary_equal"] for r in runs),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if not args.e2e_keep:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline_e2e(fs, f1, f2, width, npairs, threads):
    """The reference's path end to end on the host cores, over the first
    npairs pairs of the two FASTQ files (records of `width` bytes): read the
    bytes, parse them with the oracle's kseq_read restatement (one thread, as
    the reference's parse is a critical section), classify every pair against
    every filter with the oracle on `threads` threads, and write the summary
    text."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ofs, ids = [], []
    for i in range(len(fs)):
        inf = fs.info(i)
        f = O.Filter(inf.m, inf.h, inf.k)
        fs.download(i, f.bits)
        ofs.append(f)
        ids.append(f"g{i}")
    t0 = time.time()
    mates = []
    for path in (f1, f2):
        with open(path, "rb") as fh:
            data = fh.read(npairs * width)
        mates.append(O.kseq_seqs(data))
    t1 = time.time()
    (b1, o1, _), (b2, o2, _) = mates
    n = min(len(o1), len(o2)) - 1
    L = int(o1[1] - o1[0])
    if not (np.all(np.diff(o1[: n + 1]) == L) and np.all(np.diff(o2[: n + 1]) == L)):
        raise SystemExit("cpu e2e baseline expects fixed-length reads")
    buf = np.stack([b1[: n * L].reshape(n, L), b2[: n * L].reshape(n, L)], 1).reshape(-1)
    buf = np.concatenate([buf, np.zeros(1, np.uint8)])
    off = np.arange(2 * n + 1, dtype=np.uint64) * np.uint64(L)
    hits, _, _, _ = O.classify(ofs, buf, off, paired=True, nthreads=threads)
    t2 = time.time()
    text = O.summary(ids, hits, paired=True)
    t3 = time.time()
    return {"value": 2 * n / (t3 - t0), "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"first {n} pairs of both files: read + kseq parse (1 thread) "
                      f"{t1 - t0:.2f}s, classify ({threads} threads) {t2 - t1:.2f}s, "
                      f"summary {t3 - t2:.3f}s",
            "summary_bytes": len(text)}


def bench_mibf_e2e(args, torch, D, B):
    """biobloommicategorizer end to end: the MIBF stored to files, the reads
    as FASTQ, one CLI run writing the per-read TSV lines and summary.tsv."""
    import subprocess
    import tempfile
    from biobloom_amd import synth
    w = synth.workload_mibf(scale=args.scale)
    reads = w["reads"]
    nreads = reads.shape[0]
    mf = D.build_mibf(w["refs"], list(range(1, len(w["refs"]) + 1)), w["names"], w["seeds"],
                      occupancy=0.5)
    tmp = tempfile.mkdtemp(prefix="bbt_mibf_e2e_", dir=os.environ.get("TMPDIR", "/tmp"))
    pre = os.path.join(tmp, "multi")
    mf.store(pre)
    del w
    fq = os.path.join(tmp, "reads.fq")
    nbytes = write_fastq(fq, reads, 1)
    cli = os.path.join(ROOT, "biobloom_amd", "biobloommicategorizer")
    runs = []
    for threads in args.e2e_threads:
        out = os.path.join(tmp, f"out_{threads}.tsv")
        t = time.time()
        with open(out, "w") as fo:
            r = subprocess.run([cli, "-f", pre + ".bf", "-p", os.path.join(tmp, f"o{threads}"),
                                "-t", str(threads), fq], stdout=fo, stderr=subprocess.PIPE,
                               text=True, timeout=900)
        wall = time.time() - t
        if r.returncode != 0:
            raise SystemExit(f"biobloommicategorizer failed: {r.stderr[-2000:]}")
        tl = [x for x in r.stderr.splitlines() if x.startswith("Filtering time:")][0]
        log(tl)
        ft = float(tl.split()[2])
        lines = sum(1 for _ in open(out))
        item = {"threads": threads, "filtering_s": ft, "wall_s": wall,
                "reads_per_s": nreads / ft, "text_gb_per_s": nbytes / ft / 1e9,
                "output_lines": lines}
        runs.append(item)
        log(f"mibf e2e -t {threads}: filtering {ft:.2f} s (wall {wall:.2f} s), "
            f"{item['reads_per_s'] / 1e6:.2f} M reads/s, {lines} lines")
        os.remove(out)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_mibf_e2e(mf, fq, 2 * reads.shape[1] + 18, min(nreads, 200_000))
        log(f"cpu e2e baseline: {cpu['value'] / 1e6:.3f} M reads/s ({cpu['sample']})")
    best = max(runs, key=lambda r: r["reads_per_s"])
    line = {
        "metric": METRIC, "value": best["reads_per_s"], "unit": "reads/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": best["filtering_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "end to end: biobloommicategorizer on a FASTQ file of 2M x 150 bp "
                               "reads against the 200 x 5 Mbp multi-index Bloom filter; parse + "
                               "device batches + per-read TSV lines + summary.tsv",
                   "scale": args.scale, "reads": nreads},
        "e2e": runs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    import shutil
    shutil.rmtree(tmp, ignore_errors=True)


def load_traffic(workload):
    p = os.path.join(ROOT, "profiles", f"traffic_walk_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="multi",
                    choices=["multi", "single", "uneven", "mibf", "serve", "e2e", "mibf-e2e"])
    ap.add_argument("--e2e-args", default="",
                    help="e2e: extra biobloomcategorizer options, e.g. --batch_reads=524288")
    ap.add_argument("--e2e-keep", default=None, metavar="DIR",
                    help="e2e: write the filters and FASTQ files into DIR and keep them "
                         "(to run the CLI again, e.g. under a profiler)")
    ap.add_argument("--e2e-threads", type=lambda v: [int(x) for x in v.split(",")],
                    default=[16, 4, 1], help="e2e: CLI parse-thread counts to run")
    ap.add_argument("--e2e-gz", action="store_true", help="e2e: also gzip inputs")
    ap.add_argument("--qps", default="10,50,200,400,800,1600",
                    help="serve: request arrival rates to sweep")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="serve: arrival window per rate")
    ap.add_argument("--max-requests", type=int, default=400)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-sample", type=int, default=100_000,
                    help="records (pairs for paired workloads) in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--path", default="both", choices=["both", "batched", "plain"])
    ap.add_argument("--wave-min", type=int, default=0,
                    help="batched path: reads this long are walked by a whole wave")
    ap.add_argument("--refill-min", type=int, default=0,
                    help="batched path: idle lanes that trigger a queue refill")
    ap.add_argument("--blocks-per-cu", type=int, default=0,
                    help="batched path: persistent blocks per CU (0 = occupancy)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the batched path's scheduling knobs on this batch first")
    ap.add_argument("--no-gather", action="store_true",
                    help="skip the random-gather reference measurement")
    ap.add_argument("--gather-sweep", action="store_true",
                    help="only measure the random-line rate over buffers of 1 MiB to 8 GiB")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(across GPUs, reads would be sharded over replicated filters)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from biobloom_amd import device as D
    from biobloom_amd import lib as B

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "read_bytes": n, "write_bytes": n}))
        return
    B.load()
    if args.gather_sweep:
        sweep = []
        for e in range(20, 34):
            lps, ms = B.measure_gather(1 << e, 64, 0)
            sweep.append({"bytes": 1 << e, "lines_per_s": lps, "ms": ms})
            log(f"random gather over {(1 << e) >> 20} MiB: {lps / 1e9:.1f} G lines/s")
        print(json.dumps({"random_gather_sweep": sweep}), flush=True)
        return
    if args.workload == "mibf":
        return bench_mibf(args, torch, D, B)
    if args.workload == "serve":
        return bench_serve(args, torch, D, B)
    if args.workload == "e2e":
        return bench_e2e(args, torch, D, B)
    if args.workload == "mibf-e2e":
        return bench_mibf_e2e(args, torch, D, B)

    w, buf, off = make_workload(args.workload, args.scale)
    paired = bool(w["paired"])
    t = time.time()
    fs, _ = D.build_filters(w["genomes"], w["k"], w["fpr"])
    torch.cuda.synchronize()
    log(f"filters built on device ({time.time() - t:.2f}s): "
        + ", ".join(f"m={fs.info(i).m} h={fs.info(i).h}" for i in range(min(2, len(fs)))))
    # reads stay in input order: the batched path buckets them by length on the
    # device, the plain path walks them as they come
    dr = D.DeviceReads(buf, off)
    nreads = dr.nreads
    nrec = nreads // 2 if paired else nreads
    out = D.Outputs(nreads, len(fs), paired, False)
    stream = int(torch.cuda.current_stream().cuda_stream)
    rs = dr.struct()
    os_ = out.struct()

    def step(path, flags=0):
        out.zero_counts()
        fs.eval_batch(rs, B.params(paired=paired, path=path, flags=flags,
                                   wave_min_len=args.wave_min), os_, stream)

    tuning = None
    if args.refill_min or args.blocks_per_cu:
        fs.set_tuning(wave_min_len=args.wave_min, refill_min=args.refill_min,
                      blocks_per_cu=args.blocks_per_cu)
        tuning = fs.get_tuning()
    if args.autotune:
        t = time.time()
        tuning = fs.autotune(rs, B.params(paired=paired), os_, stream)
        log(f"autotune ({time.time() - t:.1f}s): {tuning}")

    results = {}
    paths = {"batched": B.PATH_BATCHED, "plain": B.PATH_PLAIN}
    for name, path in paths.items():
        if args.path not in ("both", name):
            continue
        ms = time_steps(lambda: step(path), args.steps, args.warmup, torch)
        walk = []
        for _ in range(3):
            step(path, B.FLAG_TIME_WALK)
            walk.append(fs.last_walk_ms())
        torch.cuda.synchronize()
        host = out.host()
        results[name] = {"ms": ms, "walk_ms": float(np.mean(walk)),
                         "probes": int(host["counters"][0]),
                         "lookups": int(host["counters"][1]),
                         "issued": int(host["counters"][2]), "hits": host["hits"],
                         "summary": host["summary"]}
        assert int(host["counters"][3]) == 0, "device error flag set"
        log(f"{name}: {ms:.3f} ms/step, walk kernel {np.mean(walk):.3f} ms, "
            f"{nreads / ms * 1e3 / 1e6:.2f} M reads/s")
    main_name = "batched" if "batched" in results else "plain"
    R = results[main_name]
    if len(results) == 2:
        a, b = results["batched"], results["plain"]
        assert np.array_equal(a["hits"], b["hits"]), "batched and plain paths disagree"
        assert a["lookups"] == b["lookups"]
    reads_per_s = nreads / (R["ms"] / 1e3)
    probes_per_s = R["probes"] / (R["ms"] / 1e3)
    alg_bytes = R["lookups"] * LINE_BYTES
    achieved = alg_bytes / (R["walk_ms"] / 1e3) / 1e9
    traffic = load_traffic(args.workload)
    gather = {}
    if not args.no_gather:
        fs_bytes = sum(fs.info(i).m // 8 for i in range(len(fs)))
        fs_pow2 = 1 << max(12, fs_bytes.bit_length() - 1)
        for label, nbytes in (("hbm_8GiB", 8 << 30), ("filter_set_size", fs_pow2)):
            lps, _ = B.measure_gather(nbytes, 64, stream)
            gather[label] = {"bytes": nbytes, "lines_per_s": lps,
                             "gbs_at_64B": lps * LINE_BYTES / 1e9}
            log(f"random gather over {nbytes >> 20} MiB: {lps / 1e9:.1f} G lines/s")

    # the same step with the reads starting in pinned host memory and the hits
    # and classes returned to it (PCIe-inclusive; not the headline value)
    pcie = None
    if main_name in results:
        hseq = torch.from_numpy(np.ascontiguousarray(buf[: int(off[-1])])).pin_memory()
        hoff = torch.from_numpy(np.ascontiguousarray(off).view(np.int64)).pin_memory()
        hhits = torch.empty(out.hits.numel(), dtype=torch.uint8).pin_memory()
        hcls = torch.empty(out.cls.numel(), dtype=torch.int32).pin_memory()
        path = paths[main_name]

        def step_pcie():
            dr.seq[: hseq.numel()].copy_(hseq, non_blocking=True)
            dr.off.copy_(hoff, non_blocking=True)
            step(path)
            hhits.copy_(out.hits, non_blocking=True)
            hcls.copy_(out.cls, non_blocking=True)

        ms_p = time_steps(step_pcie, args.steps, 1, torch)
        pcie = {"reads_per_s": nreads / (ms_p / 1e3), "ms_per_step": ms_p,
                "host_to_device_bytes": int(hseq.numel() + hoff.numel() * 8),
                "device_to_host_bytes": int(hhits.numel() + hcls.numel() * 4)}
        log(f"{main_name} with host copies: {ms_p:.3f} ms/step, "
            f"{pcie['reads_per_s'] / 1e6:.2f} M reads/s")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        sample = min(args.cpu_sample, nrec)
        v, dt, parity = cpu_baseline(fs, buf, off, paired, sample, threads, B, R["hits"])
        sample1 = max(1, sample // 10)
        v1, dt1, parity1 = cpu_baseline(fs, buf, off, paired, sample1, 1, B, R["hits"])
        model = None
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f
                              if ln.startswith("model name")), None)
        except OSError:
            pass
        cpu = {"value": v, "unit": "reads/s", "cores": threads, "kind": "port",
               "sample": f"first {sample} {'pairs' if paired else 'reads'} of the "
                         f"workload, all filters, {dt:.2f}s",
               "parity_with_gpu": parity and parity1,
               "single_thread": {"value": v1, "sample": f"first {sample1} "
                                 f"{'pairs' if paired else 'reads'}, {dt1:.2f}s"},
               "nproc": os.cpu_count(), "cpu_model": model}
        log(f"cpu oracle: {v / 1e6:.3f} M reads/s on {threads} threads, "
            f"{v1 / 1e6:.4f} M reads/s on 1 thread, parity={parity and parity1}")

    line = {
        "metric": METRIC,
        "value": reads_per_s,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": {"multi": "multi-filter batch: 10 filters x 20 Mbp, k=25 FPR 0.075, "
                                  "2M pairs 2x150 bp, paired intersection, SIMPLE -s 0.15",
                         "single": "single filter 5 Mbp k=25 FPR 0.075, 100k SE 150 bp",
                         "uneven": "uneven lengths: 4 filters x 100 Mbp k=25 FPR 0.05, "
                                   "200k log-normal long reads + 2M x 150 bp"}[args.workload],
            "scale": args.scale,
            "reads": nreads,
            "filters": len(fs),
            "path": main_name,
        },
        "probes_per_s": probes_per_s,
        "pairs_per_s": reads_per_s / 2 if paired else None,
        "probes_per_step": R["probes"],
        "lookups_per_step": R["lookups"],
        "issued_lookups_per_step": R["issued"],
        "plain_reads_per_s": (nreads / (results["plain"]["ms"] / 1e3)
                              if "plain" in results else None),
        "speedup_vs_plain": (results["plain"]["ms"] / results["batched"]["ms"]
                             if len(results) == 2 else None),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_walk_persistent" if main_name == "batched" else "k_walk_plain",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            # measured line traffic (scaled FETCH_SIZE + WRITE_SIZE of this kernel on
            # this workload) over this run's kernel time: each 64 B lookup pulls a
            # 128 B line, so this is the HBM-side figure
            "traffic_gbs": (traffic["bytes_per_launch"] / (R["walk_ms"] / 1e3) / 1e9
                            if traffic else None),
            "traffic_frac": (traffic["bytes_per_launch"] / (R["walk_ms"] / 1e3) / 1e9
                             / HBM_PEAK_GBS if traffic else None),
            "walk_ms": R["walk_ms"],
            "algorithmic_bytes": alg_bytes,
            "random_gather": gather or None,
            "frac_of_random_gather": (achieved / gather["filter_set_size"]["gbs_at_64B"]
                                      if gather else None),
        },
        "tuning": tuning,
        "pcie_inclusive": pcie,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
