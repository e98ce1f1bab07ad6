"""Benchmark of the MI355X read-classification path.

One step = one bbt_eval_batch over the whole batch of reads already resident
in HBM: every (read, filter) SeqEval walk, the ResultsManager classes and the
summary counters.

Default run: the headline workload ("multi": 10 filters from 20 Mbp genomes,
k=25, FPR 0.075, h=3; 2M read pairs 2x150 bp, paired-end intersection, SIMPLE
-s 0.15 -r 3), then, in the same JSON line under "configs", the other four
workloads of the benchmark plan, each with its own roofline and CPU baseline:
  single  1 filter from a 5 Mbp genome, 100k single-end 150 bp reads
  uneven  4 filters from 100 Mbp genomes (FPR 0.05), 200k log-normal long
          reads interleaved with 2M x 150 bp reads: batched vs plain
  mibf    multi-index Bloom filter of 200 x 5 Mbp references, 2M x 150 bp
  serve   3 Gbp filter + 8 contaminant filters resident; Poisson requests of
          1k-20k reads; p50/p99 latency of continuous batching and of the
          plain per-request path
The three large workloads are generated in worker processes while the
headline runs; every CPU baseline runs after generation has finished, so the
host cores are not shared while they are timed.  --no-suite: headline only.

The headline value is reads/s of the batched persistent path replayed as one
captured hipGraph, with the knobs bbt_autotune picked on the batch; the
eager launch of the same path and the plain per-batch path are beside it.

--workload {single,uneven,mibf,serve}: one workload alone, as its own line.
--workload e2e / mibf-e2e: the CLIs end to end on FASTQ files.
--workload dust: -D (the sdust pre-pass) against the same batches without it.
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

WORKLOAD_TEXT = {
    "multi": "multi-filter batch: 10 filters x 20 Mbp, k=25 FPR 0.075, 2M pairs 2x150 bp, "
             "paired intersection, SIMPLE -s 0.15",
    "single": "single filter 5 Mbp k=25 FPR 0.075, 100k SE 150 bp (0.001 N), SIMPLE -s 0.15",
    "uneven": "uneven lengths: 4 filters x 100 Mbp k=25 FPR 0.05, 200k log-normal long reads "
              "(median 2 kbp, 100 bp-50 kbp, 8% errors) + 2M x 150 bp",
    "mibf": "multi-index Bloom filter: 200 x 5 Mbp references (5% shared 10 kbp segments), "
            "3 spaced seeds span 25 weight 20, occupancy 0.5, 32-bit IDs, 2M x 150 bp reads",
    "serve": "serving: 3 Gbp genome (50% 300 bp repeats, 10% divergence) + 8 x 20 Mbp "
             "contaminant filters, k=25 FPR 0.075; Poisson requests of 1k-20k reads, "
             "80% 150 bp / 20% log-normal long",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def generate(name, scale):
    """One workload as host arrays (run in a worker process for the large
    ones); reads are returned packed: buf (uint8) and off (uint64)."""
    from biobloom_amd import synth
    t = time.time()
    if name in ("multi", "single", "mibf"):
        w = getattr(synth, f"workload_{name}")(scale=scale)
        reads = w["reads"]
        w["buf"] = np.ascontiguousarray(reads).reshape(-1)
        w["off"] = np.arange(reads.shape[0] + 1, dtype=np.uint64) * np.uint64(reads.shape[1])
    elif name == "uneven":
        w = synth.workload_uneven(scale=scale)
        w["buf"], w["off"] = synth.pack(w.pop("reads_list"))
    elif name == "serve":
        w = synth.workload_serve(scale=scale)
    else:
        raise SystemExit(f"unknown workload {name}")
    w["gen_s"] = time.time() - t
    return w


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f
                         if ln.startswith("model name")), None)
    except OSError:
        return None


def cpu_threads():
    return min(16, os.cpu_count() or 1)


def oracle_filters(fs):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ofs = []
    for i in range(len(fs)):
        inf = fs.info(i)
        f = O.Filter(inf.m, inf.h, inf.k)
        fs.download(i, f.bits)
        ofs.append(f)
    return O, ofs


def cpu_baseline_walk(fs, buf, off, paired, target_s, gpu_hits):
    """The oracle (OpenMP over reads, every filter per read, as
    BioBloomClassifier.cpp:59-89) on the first records of the workload: a
    probe run sizes a sample of about target_s seconds, which is then timed
    on cpu_threads() threads and a tenth of it on one thread."""
    O, ofs = oracle_filters(fs)
    per = 2 if paired else 1
    nrec = (len(off) - 1) // per
    threads = cpu_threads()

    def run(nr, th):
        n = nr * per
        sbuf, soff = buf[: int(off[n])], off[: n + 1]
        t = time.time()
        hits, _, _, _ = O.classify(ofs, sbuf, soff, paired=paired, nthreads=th)
        return time.time() - t, bool(np.array_equal(hits, gpu_hits[:n]))

    probe = min(nrec, 4000)
    run(probe, threads)  # warm: first touch of the filter copies, thread start
    dt, _ = run(probe, threads)
    sample = int(min(nrec, max(probe, probe * target_s / max(dt, 1e-4))))
    dt, parity = run(sample, threads)
    reps = 1
    if sample == nrec and dt < target_s / 4:  # the whole workload is small: repeat it
        reps = int(min(100, max(1, target_s / 2 / max(dt, 1e-4))))
        t = time.time()
        for _ in range(reps):
            run(sample, threads)
        dt = (time.time() - t) / reps
    s1 = max(1, min(sample, int(sample * 0.6 / max(dt * threads, 1e-4))))
    dt1, parity1 = run(s1, 1)
    unit = "pairs" if paired else "reads"
    return {"value": sample * per / dt, "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} {unit} of the workload, all filters, {dt:.3f}s"
                      + (f" (mean of {reps} runs)" if reps > 1 else ""),
            "parity_with_gpu": parity and parity1,
            "single_thread": {"value": s1 * per / dt1,
                              "sample": f"first {s1} {unit}, {dt1:.2f}s"},
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


class GatherCache:
    """Random-line rates (bbt_measure_gather) by buffer size, measured once."""

    def __init__(self, B, enabled):
        self.B, self.enabled, self.rates = B, enabled, {}

    def rate(self, nbytes):
        nbytes = 1 << max(12, int(nbytes).bit_length() - 1)  # largest power of two <= size
        if not self.enabled:
            return nbytes, None
        if nbytes not in self.rates:
            self.rates[nbytes], _ = self.B.measure_gather(nbytes, 64, 0)
            log(f"random gather over {nbytes >> 20} MiB: {self.rates[nbytes] / 1e9:.1f} "
                f"G lines/s")
        return nbytes, self.rates[nbytes]


def load_traffic(workload):
    name = "traffic_mibf.json" if workload == "mibf" else f"traffic_walk_{workload}.json"
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def bench_walk(args, torch, D, B, name, w, steps, warmup, gathers, autotune=True):
    """One Bloom-filter workload: the batched persistent path (autotuned,
    eager and graph-replayed) and the plain per-batch path, timed on the
    resident batch; returns (result dict, deferred CPU-baseline function)."""
    buf, off = w["buf"], w["off"]
    paired = bool(w["paired"])
    t = time.time()
    fs, _ = D.build_filters(w["genomes"], w["k"], w["fpr"])
    torch.cuda.synchronize()
    log(f"{name}: {len(fs)} filters built on device ({time.time() - t:.2f}s): "
        + ", ".join(f"m={fs.info(i).m} h={fs.info(i).h}" for i in range(min(2, len(fs)))))
    dr = D.DeviceReads(buf, off)
    nreads = dr.nreads
    nrec = nreads // 2 if paired else nreads
    out = D.Outputs(nreads, len(fs), paired, False)
    stream = int(torch.cuda.current_stream().cuda_stream)
    rs, os_ = dr.struct(), out.struct()

    def step(path, flags=0):
        out.zero_counts()
        fs.eval_batch(rs, B.params(paired=paired, path=path, flags=flags,
                                   wave_min_len=args.wave_min), os_, stream)

    tuning = None
    if args.refill_min or args.blocks_per_cu or args.probe_ahead >= 0:
        fs.set_tuning(wave_min_len=args.wave_min, refill_min=args.refill_min,
                      blocks_per_cu=args.blocks_per_cu, probe_ahead=max(0, args.probe_ahead))
        tuning = fs.get_tuning()
    elif autotune and not args.no_autotune:
        t = time.time()
        tuning = fs.autotune(rs, B.params(paired=paired), os_, stream)
        tuning["tune_s"] = time.time() - t
        log(f"{name}: autotune ({tuning['tune_s']:.1f}s): {tuning}")

    res = {}
    runs = [("batched", B.PATH_BATCHED, 0), ("batched_graph", B.PATH_BATCHED, B.FLAG_GRAPH),
            ("plain", B.PATH_PLAIN, 0)]
    for label, path, flags in runs:
        if args.path == "plain" and path != B.PATH_PLAIN:
            continue
        if args.path == "batched" and path != B.PATH_BATCHED:
            continue
        ms = time_steps(lambda: step(path, flags), steps, warmup, torch)
        host = out.host()
        walk = []
        if not flags:
            for _ in range(3):
                step(path, B.FLAG_TIME_WALK)
                walk.append(fs.last_walk_ms())
        torch.cuda.synchronize()
        res[label] = {"ms": ms, "walk_ms": float(np.mean(walk)) if walk else None,
                      "kernel": fs.last_walk_kernel(),
                      "probes": int(host["counters"][0]), "lookups": int(host["counters"][1]),
                      "issued": int(host["counters"][2]), "hits": host["hits"],
                      "summary": host["summary"]}
        assert int(host["counters"][3]) == 0, "device error flag set"
        log(f"{name} {label}: {ms:.3f} ms/step"
            + (f", walk kernel {np.mean(walk):.3f} ms" if walk else "")
            + f", {nreads / ms * 1e3 / 1e6:.2f} M reads/s")
    base = res.get("batched") or res["plain"]
    for label, r in res.items():
        assert np.array_equal(r["hits"], base["hits"]), f"{label} and the first path disagree"
        assert r["lookups"] == base["lookups"] and np.array_equal(r["summary"], base["summary"])
    main_name = "batched_graph" if "batched_graph" in res else "plain"
    R = res[main_name]
    walk_ms = (res.get("batched") or res["plain"])["walk_ms"]
    alg_bytes = base["lookups"] * LINE_BYTES
    achieved = alg_bytes / (walk_ms / 1e3) / 1e9
    fs_bytes = sum(fs.info(i).m // 8 for i in range(len(fs)))
    gbytes, glps = gathers.rate(fs_bytes)
    _, g8 = gathers.rate(8 << 30)
    traffic = load_traffic(name)
    rl = {"bound": "hbm", "kernel": (res.get("batched") or res["plain"])["kernel"],
          "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": achieved / HBM_PEAK_GBS,
          "traffic": traffic["bytes_per_launch"] if traffic else None,
          "traffic_source": (f"profiles/traffic_walk_{name}.json: rocprofv3 --pmc FETCH_SIZE + "
                             "WRITE_SIZE per launch, scaled by a device copy of known size; "
                             "L2-miss line traffic, served by HBM or the MALL"
                             if traffic else None),
          "walk_ms": walk_ms, "algorithmic_bytes": alg_bytes,
          "filter_set_bytes": fs_bytes,
          "random_gather": {"filter_set_pow2": {"bytes": gbytes, "lines_per_s": glps},
                            "hbm_8GiB": {"bytes": 8 << 30, "lines_per_s": g8}},
          "frac_of_random_gather": (base["lookups"] / (walk_ms / 1e3) / glps if glps else None)}
    if traffic:
        rl["traffic_gbs"] = traffic["bytes_per_launch"] / (walk_ms / 1e3) / 1e9
        rl["traffic_frac"] = rl["traffic_gbs"] / HBM_PEAK_GBS
    sub = {
        "workload": WORKLOAD_TEXT[name], "reads": nreads, "filters": len(fs),
        "value": nreads / (R["ms"] / 1e3), "unit": "reads/s", "path": main_name,
        "ms_per_step": R["ms"], "steps": steps, "warmup": warmup,
        "probes_per_s": base["probes"] / (R["ms"] / 1e3),
        "pairs_per_s": nreads / 2 / (R["ms"] / 1e3) if paired else None,
        "probes_per_step": base["probes"], "lookups_per_step": base["lookups"],
        "issued_lookups_per_step": base["issued"],
        "batched_eager_reads_per_s": (nreads / (res["batched"]["ms"] / 1e3)
                                      if "batched" in res else None),
        "batched_graph_reads_per_s": (nreads / (res["batched_graph"]["ms"] / 1e3)
                                      if "batched_graph" in res else None),
        "plain_reads_per_s": (nreads / (res["plain"]["ms"] / 1e3) if "plain" in res else None),
        "speedup_vs_plain": (res["plain"]["ms"] / R["ms"] if "plain" in res and R is not
                             res["plain"] else None),
        "roofline": rl, "tuning": tuning,
    }
    gpu_hits = base["hits"]

    def cpu():
        return cpu_baseline_walk(fs, buf, off, paired, args.cpu_seconds, gpu_hits)

    ctx = {"fs": fs, "dr": dr, "out": out, "rs": rs, "os": os_, "step": step,
           "nreads": nreads, "nrec": nrec, "paired": paired, "stream": stream}
    return sub, cpu, ctx


def pcie_inclusive(torch, D, B, buf, off, ctx, steps):
    """The batched step with the reads starting in pinned host memory and the
    hits and classes returned to it (PCIe-inclusive; not the headline): reads
    copied as text, and 2-bit packed (codes + N mask, unpacked on the
    device) as the CLIs send them."""
    dr, out, fs = ctx["dr"], ctx["out"], ctx["fs"]
    paired, stream, os_ = ctx["paired"], ctx["stream"], ctx["os"]
    hoff = torch.from_numpy(np.ascontiguousarray(off).view(np.int64)).pin_memory()
    hhits = torch.empty(out.hits.numel(), dtype=torch.uint8).pin_memory()
    hcls = torch.empty(out.cls.numel(), dtype=torch.int32).pin_memory()
    packed, cb = D.pack_2bit(buf, off)
    res = {}
    for name in ("ascii", "2bit"):
        if name == "ascii":
            hseq = torch.from_numpy(np.ascontiguousarray(buf[: int(off[-1])])).pin_memory()
            dseq = dr.seq
            rs = dr.struct()
        else:
            hseq = torch.from_numpy(packed).pin_memory()
            dseq = torch.empty(packed.size, dtype=torch.uint8, device="cuda")
            rs = B.Reads(dseq.data_ptr(), cb, dr.off.data_ptr(), dr.nreads, dr.max_len,
                         B.READS_2BIT, None)

        def step_pcie():
            dseq[: hseq.numel()].copy_(hseq, non_blocking=True)
            dr.off.copy_(hoff, non_blocking=True)
            out.zero_counts()
            fs.eval_batch(rs, B.params(paired=paired, path=B.PATH_BATCHED, flags=B.FLAG_GRAPH),
                          os_, stream)
            hhits.copy_(out.hits, non_blocking=True)
            hcls.copy_(out.cls, non_blocking=True)

        ms_p = time_steps(step_pcie, steps, 1, torch)
        res[name] = {"reads_per_s": ctx["nreads"] / (ms_p / 1e3), "ms_per_step": ms_p,
                     "host_to_device_bytes": int(hseq.numel() + hoff.numel() * 8),
                     "device_to_host_bytes": int(hhits.numel() + hcls.numel() * 4)}
        log(f"batched with host copies ({name} reads): {ms_p:.3f} ms/step, "
            f"{res[name]['reads_per_s'] / 1e6:.2f} M reads/s, "
            f"{res[name]['host_to_device_bytes'] / 1e6:.0f} MB in")
    r = dict(res["2bit"])
    r["reads"] = "2-bit codes + N mask (BBT_READS_2BIT)"
    r["ascii"] = res["ascii"]
    return r


def cpu_baseline_mibf(mf, buf, off, target_s, params, gpu_rows, gpu_sat):
    """Oracle MIBF query (OpenMP over reads) on the first reads, against
    host copies of the device-built bit vector and ID array; the sample is
    sized by a probe run to about target_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    info = mf.info()
    bv = np.zeros(info.m // 64, np.uint64)
    data = np.zeros(max(1, info.pop), np.uint32)
    mf.download(bv, data)
    seeds = mf.seeds() or O.Kmer(info.span, info.nseeds)
    om = O.Mibf.from_arrays(bv, data, info.m, seeds, info.nids)
    nreads = len(off) - 1
    kw = dict(score=params.score, allowed_miss=params.allowed_miss,
              multi=params.multi_thresh, min_nonsat=params.min_nonsat,
              max_hits=params.max_hits)

    def reads(n):
        return [bytes(buf[int(off[i]):int(off[i + 1])]) for i in range(n)]

    # every oracle call first builds what the reference builds once at load
    # (rank support, frame probabilities, minCount rows): timing two sample
    # sizes and taking the difference leaves the per-read query cost
    probe = min(nreads, 5000)
    rp = reads(probe)
    t = time.time()
    om.query(rp, **kw)
    t_probe = time.time() - t
    rr = reads(min(nreads, 4 * probe))
    t = time.time()
    om.query(rr, **kw)
    per_read = max(1e-7, (time.time() - t - t_probe) / max(1, len(rr) - probe))
    n = int(min(nreads, max(2 * probe, target_s / per_read)))
    rr = reads(n)
    t = time.time()
    rows, sat, _ = om.query(rr, **kw)
    t_n = time.time() - t
    parity = rows == gpu_rows[:n] and sat == gpu_sat[:n]
    t = time.time()
    om.query(rp, **kw)
    t_probe = time.time() - t
    dt = max(1e-6, t_n - t_probe)
    return {"value": (n - probe) / dt, "unit": "reads/s", "cores": cpu_threads(),
            "kind": "port",
            "sample": f"first {n} reads of the workload in {t_n:.2f}s, less the first {probe} "
                      f"in {t_probe:.2f}s (the per-call setup the reference does once at load)",
            "parity_with_gpu": bool(parity), "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def bench_mibf(args, torch, D, B, w, steps, warmup, gathers):
    buf, off = w["buf"], w["off"]
    nreads = len(off) - 1
    t = time.time()
    mf = D.build_mibf(w["refs"], list(range(1, len(w["refs"]) + 1)), w["names"], w["seeds"],
                      occupancy=0.5)
    info = mf.info()
    log(f"mibf: built on device ({time.time() - t:.2f}s): m={info.m} pop={info.pop} "
        f"saturated={info.pop_saturated} ids={info.nids}")
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

    ms = time_steps(step, steps, warmup, torch)
    # the same queries through the lines and the ID array (two dependent
    # random loads per set bit) instead of the direct-mapped table
    lines_ms = None
    if info.direct:
        lp = B.mibf_params(score=1e-10, multi_thresh=1.0, allowed_miss=info.nseeds - 1,
                           min_nonsat=0, max_hits=8, lines=True)
        lines_ms = time_steps(lambda: mf.query_batch(rs, lp, os_, stream), steps, 1, torch)
        log(f"mibf query through lines + ID array: {lines_ms:.3f} ms/step")
    out.counters.zero_()
    step()
    torch.cuda.synchronize()
    rows, sat, ctr, _ = out.host()
    lookups, gathers_n = int(ctr[0]), int(ctr[1])
    assert int(ctr[3]) == 0 or all(x != 0xFFFFFFFF for x in out.nhits.cpu().numpy()), \
        "records not evaluated"
    classified = sum(1 for r in rows if r)
    alg_bytes = (lookups + gathers_n) * LINE_BYTES
    achieved = alg_bytes / (ms / 1e3) / 1e9
    log(f"mibf query: {ms:.3f} ms/step, {nreads / ms * 1e3 / 1e6:.2f} M reads/s, "
        f"{lookups / (ms / 1e3) / 1e9:.2f} G lookups/s, classified {classified}/{nreads}")
    _, g8 = gathers.rate(8 << 30)
    traffic = load_traffic("mibf")
    rl = {"bound": "hbm", "kernel": "k_mibf_query", "achieved": achieved,
          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
          "traffic": traffic["bytes_per_launch"] if traffic else None,
          "traffic_source": ("profiles/traffic_mibf.json: rocprofv3 --pmc FETCH_SIZE + "
                             "WRITE_SIZE per launch, scaled by a device copy of known size"
                             if traffic else None),
          "algorithmic_bytes": alg_bytes, "query_ms": ms,
          "layout": "direct-mapped table" if info.direct else "lines + ID array",
          "random_loads_per_step": lookups if info.direct else lookups + gathers_n,
          "lines_layout_query_ms": lines_ms,
          "structure_bytes": (int(info.m // 64 // 7 + 1) * 64 + int(info.pop) * 4
                              + (int(info.m) * 4 if info.direct else 0)),
          "random_gather": {"hbm_8GiB": {"bytes": 8 << 30, "lines_per_s": g8}},
          "frac_of_random_gather": ((lookups + gathers_n) / (ms / 1e3) / g8 if g8 else None)}
    if traffic:
        rl["traffic_gbs"] = traffic["bytes_per_launch"] / (ms / 1e3) / 1e9
        rl["traffic_frac"] = rl["traffic_gbs"] / HBM_PEAK_GBS
    sub = {"workload": WORKLOAD_TEXT["mibf"], "reads": nreads, "ids": info.nids, "m": info.m,
           "pop": info.pop, "value": nreads / (ms / 1e3), "unit": "reads/s",
           "ms_per_step": ms, "steps": steps, "warmup": warmup,
           "probes_per_s": lookups / (ms / 1e3), "probes_per_step": lookups,
           "id_gathers_per_step": gathers_n, "classified_reads": classified,
           "roofline": rl}

    def cpu():
        return cpu_baseline_mibf(mf, buf, off, args.cpu_seconds, params, rows, sat)

    return sub, cpu


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


def cpu_baseline_serve(fs, pool, target_s, want):
    """The oracle on the reads of the first pool requests (about target_s
    seconds of them), every filter per read, on cpu_threads() threads: the
    reference's categorizer classifying the same requests on the host."""
    O, ofs = oracle_filters(fs)
    threads = cpu_threads()
    done, dt_all, parity = 0, 0.0, True
    for i, (buf, off) in enumerate(pool):
        t = time.time()
        hits, _, _, _ = O.classify(ofs, buf, off, paired=False, nthreads=threads)
        dt_all += time.time() - t
        done += len(off) - 1
        parity &= bool(np.array_equal(hits, want[i][0]))
        if dt_all >= target_s:
            break
    return {"value": done / dt_all, "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"the reads of the first {i + 1} requests of the pool ({done} reads, "
                      f"mixed lengths), all 9 filters, {dt_all:.2f}s",
            "parity_with_gpu": parity, "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def serve_resident(torch, D, B, fs, pool, want, steps=3):
    """The serving pool's reads as one batch already in HBM, through the
    batched and the plain path with the library's default knobs: the walk
    kernel on the HBM-resident 3 Gbp filter without host copies or queueing.
    Each path's hits and classes are checked against the requests alone."""
    bufs, offs, base = [], [], 0
    for buf, off in pool:
        n = int(off[-1])
        bufs.append(np.asarray(buf[:n]))
        offs.append(np.asarray(off[:-1], np.uint64) + np.uint64(base))
        base += n
    rbuf = np.concatenate(bufs)
    roff = np.concatenate(offs + [np.array([base], np.uint64)])
    want_hits = np.concatenate([h for h, _ in want])
    want_cls = np.concatenate([c for _, c in want])
    dr = D.DeviceReads(rbuf, roff)
    out = D.Outputs(dr.nreads, len(fs), False, False)
    rs, os_ = dr.struct(), out.struct()
    stream = int(torch.cuda.current_stream().cuda_stream)
    res = {"reads": dr.nreads, "bases": base, "max_len": dr.max_len,
           "knobs": "library defaults"}
    for label, path in (("batched", B.PATH_BATCHED), ("plain", B.PATH_PLAIN)):
        def step(flags=0):
            out.zero_counts()
            fs.eval_batch(rs, B.params(path=path, flags=flags), os_, stream)

        ms = time_steps(step, steps, 1, torch)
        walk = []
        for _ in range(3):
            step(B.FLAG_TIME_WALK)
            walk.append(fs.last_walk_ms())
        torch.cuda.synchronize()
        host = out.host()
        assert int(host["counters"][3]) == 0, "device error flag set"
        res[label] = {"ms": ms, "walk_ms": float(np.mean(walk)), "kernel": fs.last_walk_kernel(),
                      "reads_per_s": dr.nreads / (ms / 1e3),
                      "probes": int(host["counters"][0]), "lookups": int(host["counters"][1]),
                      "issued": int(host["counters"][2]),
                      "equal_to_alone": bool(np.array_equal(host["hits"], want_hits)
                                             and np.array_equal(host["cls"], want_cls))}
        log(f"serve resident {label}: {ms:.3f} ms/step, walk {np.mean(walk):.3f} ms, "
            f"{res[label]['reads_per_s'] / 1e6:.2f} M reads/s, "
            f"{res[label]['lookups'] / (np.mean(walk) / 1e3) / 1e9:.1f} G lookups/s, "
            f"equal={res[label]['equal_to_alone']}")
    res["speedup_vs_plain"] = res["plain"]["ms"] / res["batched"]["ms"]
    return res


def bench_serve(args, torch, D, B, w, gathers):
    pool = w["requests"]
    t = time.time()
    fs, _ = D.build_filters(w["genomes"], w["k"], w["fpr"])
    torch.cuda.synchronize()
    big = fs.info(0)
    log(f"serve: filters built on device ({time.time() - t:.1f}s): big m={big.m} "
        f"({big.m / 8 / 2**30:.2f} GiB), h={big.h}; pool of {len(pool)} requests")
    if args.probe_ahead >= 0:
        fs.set_tuning(probe_ahead=args.probe_ahead)
    params = B.params()
    # each pool request alone through bbt_eval_batch: the results every
    # served copy must reproduce
    want = []
    t = time.time()
    for buf, off in pool:
        r = D.classify(fs, D.DeviceReads(buf, off), params)
        want.append((r["hits"], r["cls"]))
    log(f"serve: pool classified alone ({time.time() - t:.1f}s)")
    runs = {"plain": [], "batched": []}
    for mode, qps_list in (("plain", args.plain_qps), ("batched", args.qps)):
        for i, qps in enumerate(qps_list):
            r = serve_run(B, fs, params, pool, mode, qps, args.serve_seconds,
                          args.max_requests, 1000 + i, want)
            runs[mode].append(r)
            log(f"serve {mode} qps={qps:g}: {r['reads_per_s'] / 1e6:.2f} M reads/s, p50 "
                f"{r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, waves {r['waves']}, "
                f"{r['mean_wave_requests']:.2f} req/wave, equal={r['equal_to_alone']}; "
                f"span {r['span_ms']:.0f} ms: submit {r['submit_ms']:.0f}, pack "
                f"{r['pack_ms']:.0f}, device {r['device_ms']:.0f} ms")
    top = max(runs["batched"], key=lambda r: r["reads_per_s"])
    plain_top = max(runs["plain"], key=lambda r: r["reads_per_s"]) if runs["plain"] else None
    # latency at the lowest offered load both paths ran at
    low = min(runs["batched"], key=lambda r: r["qps"])
    low_plain = (min(runs["plain"], key=lambda r: r["qps"]) if runs["plain"] else None)
    # the random-line rate over a buffer of the big filter's size: what the
    # lookups (one random 64 B line each) are measured against
    gbytes, glps = gathers.rate(max(big.m // 8, 4096) * 2 - 1)
    served = {"achieved": top["lookups_per_s"] * LINE_BYTES / 1e9,
              "frac": top["lookups_per_s"] * LINE_BYTES / 1e9 / HBM_PEAK_GBS,
              "frac_of_random_gather": top["lookups_per_s"] / glps if glps else None,
              "note": "lookups x 64 B of the saturated batched run over its wall span "
                      "(host copies included)"}
    res = serve_resident(torch, D, B, fs, pool, want)
    walk_ms = res["batched"]["walk_ms"]
    lookups = res["batched"]["lookups"]
    achieved = lookups * LINE_BYTES / (walk_ms / 1e3) / 1e9
    traffic = load_traffic("serve")
    rl = {"bound": "hbm", "kernel": res["batched"]["kernel"],
          "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": achieved / HBM_PEAK_GBS,
          "traffic": traffic["bytes_per_launch"] if traffic else None,
          "traffic_source": ("profiles/traffic_walk_serve.json: rocprofv3 --pmc FETCH_SIZE + "
                             "WRITE_SIZE per launch, scaled by a device copy of known size"
                             if traffic else None),
          "note": "the request pool as one batch resident in HBM: reference lookups x 64 B "
                  "over the walk kernel's event-timed duration",
          "walk_ms": walk_ms, "algorithmic_bytes": lookups * LINE_BYTES,
          "filter_set_bytes": sum(fs.info(i).m // 8 for i in range(len(fs))),
          "random_gather": {"big_filter_pow2": {"bytes": gbytes, "lines_per_s": glps}},
          "frac_of_random_gather": (lookups / (walk_ms / 1e3) / glps if glps else None),
          "served": served}
    if traffic:
        rl["traffic_gbs"] = traffic["bytes_per_launch"] / (walk_ms / 1e3) / 1e9
        rl["traffic_frac"] = rl["traffic_gbs"] / HBM_PEAK_GBS
    sub = {"workload": WORKLOAD_TEXT["serve"], "filters": len(fs), "big_filter_bits": big.m,
           "value": top["reads_per_s"], "unit": "reads/s", "qps_at_value": top["qps"],
           "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
           "low_load": {"qps": low["qps"], "p50_ms": low["p50_ms"], "p99_ms": low["p99_ms"],
                        "plain_qps": low_plain["qps"] if low_plain else None,
                        "plain_p50_ms": low_plain["p50_ms"] if low_plain else None,
                        "plain_p99_ms": low_plain["p99_ms"] if low_plain else None},
           "plain_reads_per_s": plain_top["reads_per_s"] if plain_top else None,
           "speedup_vs_plain": (top["reads_per_s"] / plain_top["reads_per_s"]
                                if plain_top else None),
           "saturated_lookups_per_s": top["lookups_per_s"], "roofline": rl,
           "resident": res, "serving": runs,
           "alone_equals_mixed": all(r["equal_to_alone"] for m in runs.values() for r in m)}

    def cpu():
        return cpu_baseline_serve(fs, pool, args.cpu_seconds, want)

    return sub, cpu


def bench_dust(args, torch, D, B):
    """-D (sdust pre-pass, T = 20, W = 64) against the same batch without it:
    the single-filter workload and a quarter of the multi-filter reads
    (1M x 150 bp), batched path, steps timed by events around the call."""
    from biobloom_amd import synth
    res = {}
    for name, scale in (("single", 1.0), ("multi", 0.25)):
        w = getattr(synth, f"workload_{name}")(scale=scale)
        fs, _ = D.build_filters(w["genomes"], w["k"], w["fpr"])
        dr = D.DeviceReads.from_matrix(w["reads"])
        paired = bool(w["paired"])
        out = D.Outputs(dr.nreads, len(fs), paired, False)
        stream = int(torch.cuda.current_stream().cuda_stream)
        rs, os_ = dr.struct(), out.struct()
        r = {"reads": dr.nreads}
        for dust in (False, True):
            def step(flags=0):
                out.zero_counts()
                fs.eval_batch(rs, B.params(paired=paired, dust=dust, flags=flags), os_, stream)
            ms = time_steps(step, args.steps, args.warmup, torch)
            step(B.FLAG_TIME_WALK)
            torch.cuda.synchronize()
            key = "dust" if dust else "plain"
            r[key] = {"ms_per_step": ms, "walk_ms": fs.last_walk_ms(),
                      "reads_per_s": dr.nreads / (ms / 1e3)}
            log(f"dust {name} dust={dust}: {ms:.3f} ms/step, walk {fs.last_walk_ms():.3f} ms")
        r["dust_ms_per_million_reads"] = ((r["dust"]["ms_per_step"] - r["plain"]["ms_per_step"])
                                          * 1e6 / dr.nreads)
        res[name] = r
    line = {"metric": "-D step time (sdust pre-pass + walks)", "unit": "ms",
            "higher_is_better": False, "n_gpus": 1, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "single (100k reads) and multi (1M reads), T=20 W=64"},
            "results": res}
    print(json.dumps(line), flush=True)


def as_line(sub, steps, warmup, dtype="u64"):
    """A sub-object as a full bench line of its own (--workload runs)."""
    line = {"metric": METRIC, "value": sub["value"], "unit": sub["unit"], "n_gpus": 1,
            "steps": sub.get("steps", steps), "warmup": sub.get("warmup", warmup),
            "ms_per_step": sub.get("ms_per_step"), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": {"workload": sub["workload"]}}
    line.update({k: v for k, v in sub.items() if k not in line and k != "workload"})
    return line


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
    w = generate("multi", args.scale)
    buf, off = w["buf"], w["off"]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="multi",
                    choices=["multi", "single", "uneven", "mibf", "serve", "e2e", "mibf-e2e",
                             "dust"])
    ap.add_argument("--no-suite", action="store_true",
                    help="multi: the headline workload only, without the other configs")
    ap.add_argument("--suite-steps", type=int, default=3,
                    help="timed steps of each config of the suite")
    ap.add_argument("--e2e-args", default="",
                    help="e2e: extra biobloomcategorizer options, e.g. --batch_reads=524288")
    ap.add_argument("--e2e-keep", default=None, metavar="DIR",
                    help="e2e: write the filters and FASTQ files into DIR and keep them "
                         "(to run the CLI again, e.g. under a profiler)")
    ap.add_argument("--e2e-threads", type=lambda v: [int(x) for x in v.split(",")],
                    default=[16, 4, 1], help="e2e: CLI parse-thread counts to run")
    ap.add_argument("--e2e-gz", action="store_true", help="e2e: also gzip inputs")
    ap.add_argument("--qps", type=lambda v: [float(x) for x in v.split(",")],
                    default=[5, 50, 400, 1600],
                    help="serve: arrival rates of the batched server")
    ap.add_argument("--plain-qps", type=lambda v: [float(x) for x in v.split(",")],
                    default=[5, 20], help="serve: arrival rates of the plain per-request path")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="serve: arrival window per rate")
    ap.add_argument("--max-requests", type=int, default=400)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baselines: seconds of oracle work per sample (about)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--path", default="both", choices=["both", "batched", "plain"])
    ap.add_argument("--wave-min", type=int, default=0,
                    help="batched path: reads this long are walked by a whole wave")
    ap.add_argument("--refill-min", type=int, default=0,
                    help="batched path: idle lanes that trigger a queue refill")
    ap.add_argument("--blocks-per-cu", type=int, default=0,
                    help="batched path: persistent blocks per CU (0 = occupancy)")
    ap.add_argument("--probe-ahead", type=int, default=-1,
                    help="batched path: lane-walk load scheduling (bit 0: a call's lookups "
                         "4 at a time; bit 1: the next window's issued ahead)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched path: keep the library's default knobs")
    ap.add_argument("--no-gather", action="store_true",
                    help="skip the random-gather reference measurements")
    ap.add_argument("--gather-sweep", action="store_true",
                    help="only measure the random-line rate over buffers of 1 MiB to 32 GiB")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(across GPUs, reads would be sharded over replicated filters)",
              file=sys.stderr)
        sys.exit(2)

    # the large workloads of the suite are generated in worker processes
    # (started before the GPU is touched) while the headline runs
    suite = args.workload == "multi" and not args.no_suite
    pool, futures = None, {}
    if suite:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        pool = ProcessPoolExecutor(max_workers=3, mp_context=mp.get_context("spawn"))
        for name in ("mibf", "uneven", "serve"):
            futures[name] = pool.submit(generate, name, args.scale)

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
        for e in range(20, 36):
            lps, ms = B.measure_gather(1 << e, 64, 0)
            sweep.append({"bytes": 1 << e, "lines_per_s": lps, "ms": ms})
            log(f"random gather over {(1 << e) >> 20} MiB: {lps / 1e9:.1f} G lines/s")
        print(json.dumps({"random_gather_sweep": sweep}), flush=True)
        return
    if args.workload == "e2e":
        return bench_e2e(args, torch, D, B)
    if args.workload == "mibf-e2e":
        return bench_mibf_e2e(args, torch, D, B)
    if args.workload == "dust":
        return bench_dust(args, torch, D, B)
    gathers = GatherCache(B, not args.no_gather)
    if args.workload in ("single", "uneven", "mibf", "serve"):
        w = generate(args.workload, args.scale)
        log(f"workload {args.workload} generated ({w['gen_s']:.1f}s)")
        if args.workload == "mibf":
            sub, cpu = bench_mibf(args, torch, D, B, w, args.steps, args.warmup, gathers)
        elif args.workload == "serve":
            sub, cpu = bench_serve(args, torch, D, B, w, gathers)
        else:
            sub, cpu, _ = bench_walk(args, torch, D, B, args.workload, w, args.steps,
                                     args.warmup, gathers)
        sub["cpu_baseline"] = None if args.no_cpu else cpu()
        print(json.dumps(as_line(sub, args.steps, args.warmup)), flush=True)
        return

    # ---- headline: the multi-filter batch
    t0 = time.time()
    w = generate("multi", args.scale)
    log(f"workload multi: {len(w['off']) - 1} reads, {int(w['off'][-1])} bases, "
        f"{len(w['genomes'])} genomes ({w['gen_s']:.1f}s)")
    head, head_cpu, ctx = bench_walk(args, torch, D, B, "multi", w, args.steps, args.warmup,
                                     gathers)
    pcie = (pcie_inclusive(torch, D, B, w["buf"], w["off"], ctx, args.steps)
            if "batched" in head["path"] else None)
    del ctx
    deferred = [("multi", head, head_cpu)]
    configs = {}
    if suite:
        ws = generate("single", args.scale)
        sub, cpu, _ = bench_walk(args, torch, D, B, "single", ws, args.suite_steps, 1, gathers)
        configs["single"] = sub
        deferred.append(("single", sub, cpu))
        for name in ("mibf", "uneven", "serve"):
            t = time.time()
            wn = futures[name].result()
            log(f"workload {name} generated in a worker ({wn['gen_s']:.1f}s; waited "
                f"{time.time() - t:.1f}s)")
            if name == "mibf":
                sub, cpu = bench_mibf(args, torch, D, B, wn, args.suite_steps, 1, gathers)
            elif name == "serve":
                sub, cpu = bench_serve(args, torch, D, B, wn, gathers)
            else:
                sub, cpu, _ = bench_walk(args, torch, D, B, name, wn, args.suite_steps, 1,
                                         gathers)
            configs[name] = sub
            deferred.append((name, sub, cpu))
            futures[name] = None
        pool.shutdown()
    # CPU baselines last: no worker shares the host cores while they are timed
    for name, sub, cpu in deferred:
        if args.no_cpu:
            sub["cpu_baseline"] = None
            continue
        t = time.time()
        sub["cpu_baseline"] = cpu()
        c = sub["cpu_baseline"]
        log(f"{name} cpu oracle: {c['value'] / 1e6:.4f} M reads/s on {c['cores']} threads "
            f"({c['sample']}; {time.time() - t:.1f}s)")
    line = {
        "metric": METRIC, "value": head["value"], "unit": "reads/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT["multi"], "scale": args.scale,
                   "reads": head["reads"], "filters": head["filters"], "path": head["path"]},
    }
    for k, v in head.items():
        if k not in line and k not in ("workload", "reads", "filters", "path", "steps",
                                       "warmup", "ms_per_step", "value", "unit"):
            line[k] = v
    line["pcie_inclusive"] = pcie
    if configs:
        line["configs"] = configs
    line["bench_wall_s"] = time.time() - t0
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
