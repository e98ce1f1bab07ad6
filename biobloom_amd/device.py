"""Device-side batches for libbbt_hip: reads staged in HBM, filters built on
the device, classification outputs.  torch provides device memory and
streams only; every computation runs in libbbt_hip's kernels."""
import numpy as np
import torch

from . import lib as B


def _stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


class DeviceReads:
    """Reads packed back to back in device memory (bbt_reads)."""

    def __init__(self, buf, off, device="cuda", order=None):
        off = np.ascontiguousarray(off, np.uint64)
        total = int(off[-1])
        nbytes = ((total + 7) // 8) * 8 + 8
        self.seq = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        if total:
            self.seq[:total].copy_(torch.from_numpy(np.ascontiguousarray(buf[:total])))
        self.off = torch.from_numpy(off.view(np.int64)).to(device)
        self.nreads = len(off) - 1
        lens = np.diff(off)
        self.max_len = int(lens.max()) if len(lens) else 0
        self.lens = lens
        self.order = None
        if order is not None:
            self.order = torch.from_numpy(np.ascontiguousarray(order, np.int32)).to(device)
        self.format, self.seq_bytes = B.READS_ASCII, self.seq.numel()

    @classmethod
    def packed(cls, buf, off, device="cuda", order=None):
        """The same reads in the BBT_READS_2BIT layout (3/8 byte per base)."""
        data, cb = pack_2bit(buf, off)
        self = cls(np.zeros(0, np.uint8), np.zeros(1, np.uint64), device, order)
        off = np.ascontiguousarray(off, np.uint64)
        self.seq = torch.from_numpy(data).to(device)
        self.off = torch.from_numpy(off.view(np.int64)).to(device)
        self.nreads = len(off) - 1
        self.lens = np.diff(off)
        self.max_len = int(self.lens.max()) if len(self.lens) else 0
        self.format, self.seq_bytes = B.READS_2BIT, cb
        return self

    @classmethod
    def from_matrix(cls, reads2d, device="cuda"):
        n, L = reads2d.shape
        off = np.arange(n + 1, dtype=np.uint64) * np.uint64(L)
        return cls(reads2d.reshape(-1), off, device)

    def struct(self):
        return B.Reads(self.seq.data_ptr(), self.seq_bytes, self.off.data_ptr(),
                       self.nreads, self.max_len, self.format,
                       self.order.data_ptr() if self.order is not None else None)


def pack_2bit(buf, off):
    """Reads (uint8 buffer, nreads + 1 offsets) in the BBT_READS_2BIT layout
    of include/bbt_hip.h: (codes then N mask as one uint8 array, code bytes)."""
    n = int(off[-1])
    cb = (n + 31) // 32 * 8
    padded = np.full(cb * 4, ord("A"), np.uint8)
    padded[:n] = buf[:n]
    c = ((padded >> 1) & 3).reshape(-1, 4)
    codes = (c[:, 0] | (c[:, 1] << 2) | (c[:, 2] << 4) | (c[:, 3] << 6)).astype(np.uint8)
    up = padded & 0xDF
    bad = ~((up == ord("A")) | (up == ord("C")) | (up == ord("G")) | (up == ord("T")))
    return np.concatenate([codes, np.packbits(bad, bitorder="little")]), cb


def length_order(lens):
    """Processing order of reads: longest first (LPT), stable."""
    return np.argsort(-np.asarray(lens, np.int64), kind="stable").astype(np.int32)


def build_filters(genomes, k, fpr, ids=None, device="cuda", stream=None):
    """biobloommaker on the device: one filter per genome (list of uint8
    arrays, records separated by a non-ACGT byte), sized as the reference's
    BloomFilterInfo (expected entries = sum of len - k + 1)."""
    L = B.load()
    h = L.bbt_calc_optimal_hashnum(fpr)
    ms = []
    for g in genomes:
        n = max(1, len(g) - k + 1)
        ms.append(L.bbt_calc_optimal_size(n, fpr, h))
    fs = B.FilterSet.create(ms, [h] * len(ms), [k] * len(ms))
    st = _stream_handle(stream)
    counts = []
    for i, g in enumerate(genomes):
        dg = torch.from_numpy(np.ascontiguousarray(g)).to(device)
        counts.append(fs.insert(i, dg.data_ptr(), dg.numel(), st))
        fs.refresh(i, st)
        fs.set_id(i, ids[i] if ids else f"g{i}")
        del dg
    return fs, counts


class Outputs:
    def __init__(self, nreads, nf, paired, want_scores, device="cuda"):
        nrec = nreads // 2 if paired else nreads
        self.hits = torch.zeros(max(1, nreads * nf), dtype=torch.uint8, device=device)
        self.scores = (torch.zeros(max(1, nreads * nf), dtype=torch.float64, device=device)
                       if want_scores else None)
        self.cls = torch.zeros(max(1, nrec), dtype=torch.int32, device=device)
        self.summary = torch.zeros(2 * nf + 2, dtype=torch.int64, device=device)
        self.counters = torch.zeros(4, dtype=torch.int64, device=device)
        self.nreads, self.nf, self.nrec = nreads, nf, nrec

    def struct(self):
        return B.EvalOut(self.hits.data_ptr(),
                         self.scores.data_ptr() if self.scores is not None else None,
                         self.cls.data_ptr(), self.summary.data_ptr(),
                         self.counters.data_ptr())

    def zero_counts(self):
        self.summary.zero_()
        self.counters.zero_()

    def host(self):
        return {
            "hits": self.hits[: self.nreads * self.nf].cpu().numpy().reshape(self.nreads, self.nf),
            "scores": (None if self.scores is None else
                       self.scores[: self.nreads * self.nf].cpu().numpy().reshape(self.nreads, self.nf)),
            "cls": self.cls[: self.nrec].cpu().numpy().astype(np.uint32),
            "summary": self.summary.cpu().numpy().astype(np.uint64),
            "counters": self.counters.cpu().numpy().astype(np.uint64),
        }


def classify(fs, dreads, params, want_scores=None, stream=None, out=None):
    """Runs one bbt_eval_batch and returns host copies of every output."""
    if want_scores is None:
        want_scores = params.mode in (B.BESTHIT, B.SCORES)
    if out is None:
        out = Outputs(dreads.nreads, len(fs), bool(params.paired), want_scores,
                      dreads.seq.device)
    out.zero_counts()
    fs.eval_batch(dreads.struct(), params, out.struct(), _stream_handle(stream))
    torch.cuda.synchronize()
    return out.host()


def pack_records(records, sep=b"N"):
    """Reference records (bytes/str) packed back to back, each followed by one
    separator byte, as bbt_mibf_build takes them: (uint8 buffer, offsets)."""
    parts, off = [], [0]
    for r in records:
        b = r.encode() if isinstance(r, str) else bytes(r)
        parts.append(b + sep)
        off.append(off[-1] + len(b) + 1)
    buf = np.frombuffer(b"".join(parts), np.uint8) if parts else np.zeros(0, np.uint8)
    return buf, np.asarray(off, np.uint64)


def build_mibf(records, ids, names, seeds, occupancy=0.5, m=0, device="cuda", stream=None,
               kmer=None):
    """biobloommimaker on the device: records (str/bytes) with their IDs
    (1..len(names)), one ID per record or several records per ID.  seeds:
    spaced seeds, or None with kmer = (k, hash values) for k-mer mode."""
    buf, off = pack_records(records)
    dseq = torch.from_numpy(buf.copy()).to(device)
    doff = torch.from_numpy(off.view(np.int64)).to(device)
    did = torch.from_numpy(np.asarray(ids, np.uint32).view(np.int32)).to(device)
    mf = B.Mibf.build(dseq.data_ptr(), doff.data_ptr(), did.data_ptr(), off.tolist(), names,
                      seeds, occupancy=occupancy, m=m, stream=_stream_handle(stream),
                      kmer=kmer)
    torch.cuda.synchronize()
    return mf


class MibfOutputs:
    """Per-record outputs of bbt_mibf_query_batch (a record is a read, or a
    pair in paired mode)."""

    def __init__(self, nrec, max_hits, device="cuda"):
        n = max(1, nrec)
        nreads = nrec
        self.nhits = torch.zeros(n, dtype=torch.int32, device=device)
        self.hits = torch.zeros(n * max_hits * B.MIBF_HIT_WORDS, dtype=torch.int32,
                                device=device)
        self.sat = torch.zeros(n, dtype=torch.int32, device=device)
        self.counters = torch.zeros(4, dtype=torch.int64, device=device)
        self.nreads, self.max_hits = nreads, max_hits

    def struct(self):
        return B.MibfOut(self.nhits.data_ptr(), self.hits.data_ptr(), self.sat.data_ptr(),
                         self.counters.data_ptr())

    def host(self):
        """(per read list of hit tuples in rank order, per read satCount,
        counters, raw nhits)"""
        n, mh, w = self.nreads, self.max_hits, B.MIBF_HIT_WORDS
        nh = self.nhits[:n].cpu().numpy().view(np.uint32)
        hits = self.hits[: n * mh * w].cpu().numpy().view(np.uint32).reshape(n, mh, w)
        rows = []
        for r in range(n):
            if nh[r] == 0xFFFFFFFF:
                raise RuntimeError(f"MIBF query: record {r} was not evaluated "
                                   "(its frame records overflowed the per-record store)")
            c = int(nh[r])
            rows.append([tuple(int(x) for x in hits[r, j]) for j in range(min(c, mh))])
        return (rows, self.sat[:n].cpu().numpy().view(np.uint32).tolist(),
                self.counters.cpu().numpy().astype(np.uint64), nh.copy())


def mibf_classify(mf, dreads, params, stream=None, out=None):
    if out is None:
        nrec = dreads.nreads // 2 if params.flags & B.MIBF_PAIRED else dreads.nreads
        out = MibfOutputs(nrec, params.max_hits, dreads.seq.device)
    out.counters.zero_()
    mf.query_batch(dreads.struct(), params, out.struct(), _stream_handle(stream))
    torch.cuda.synchronize()
    return out.host()
