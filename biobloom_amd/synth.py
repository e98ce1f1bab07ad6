"""Deterministic synthetic genomes and reads for the classification workloads.

Sequences are numpy uint8 arrays of ASCII bases.  Randomness is numpy's
PCG64 seeded per workload (master seed 20261203 + workload number), so a
given workload is identical on every machine with the same numpy.

Workloads (reference flags in parentheses):
  single   1 x 5 Mbp genome, k=25 FPR 0.075; 100k SE 150 bp reads, half from
           the genome (1% substitutions), half random, each base N w.p. 0.001
  multi    10 x 20 Mbp genomes, GC ~ U(0.35, 0.65), k=25 FPR 0.075; 2M pairs
           2x150, 70% from the genomes (1% substitutions), 30% random
  uneven   4 x 100 Mbp genomes, k=25 FPR 0.05; 200k log-normal long reads
           (median 2 kbp, 100 bp-50 kbp, 5% subst 1.5% ins 1.5% del)
           interleaved with 2M x 150 bp reads
  serve    one 3 Gbp genome (24 records, half of it copies of 1000 random
           300 bp repeat consensuses with 10% divergence) and 8 x 20 Mbp
           contaminant genomes, k=25 FPR 0.075; a pool of requests of
           1k-20k reads each, 80% 150 bp / 20% log-normal long reads, from the
           big genome (50%), the contaminants (10%) or random (40%)
  mibf     200 x 5 Mbp references ref000..ref199, each with 5% of its sequence
           overwritten by 10 kbp segments copied from other references'
           original sequence; 3 spaced seeds of span 25 and weight 20 closed
           under reversal; 2M x 150 bp reads, 90% from the references (1%
           substitutions), 10% random
"""
import numpy as np

MASTER_SEED = 20261203
ASCII = np.frombuffer(b"ACGT", np.uint8)
COMP_ASCII = np.zeros(256, np.uint8)
for a, b in zip(b"ACGTNacgtn", b"TGCANtgcan"):
    COMP_ASCII[a] = b


def rng_for(workload, sub=0):
    return np.random.Generator(np.random.PCG64([MASTER_SEED + workload, sub]))


def random_codes(rng, n, gc=0.5):
    """n bases as codes 0..3 (A C G T) with P(G)=P(C)=gc/2."""
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return rng.choice(4, size=n, p=p).astype(np.uint8)


def codes_to_ascii(codes):
    return ASCII[codes]


def revcomp(seq):
    return COMP_ASCII[seq[::-1]]


def pack(reads_list):
    """list of uint8 arrays -> (buffer, offsets)"""
    lens = np.array([len(r) for r in reads_list], np.uint64)
    off = np.zeros(len(reads_list) + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    buf = np.concatenate(reads_list) if reads_list else np.zeros(0, np.uint8)
    return buf, off


def fixed_len_reads(rng, genomes, n, L, subst, from_frac, n_prob=0.0, chunk=250_000):
    """n reads of length L: from_frac drawn from a uniform genome (uniform
    start and strand, i.i.d. substitutions), the rest uniform random."""
    out = np.empty((n, L), np.uint8)
    src = rng.random(n) < from_frac
    gsel = rng.integers(0, len(genomes), size=n)
    strand = rng.random(n) < 0.5
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = b - a
        block = codes_to_ascii(rng.integers(0, 4, size=(m, L), dtype=np.uint8))
        for g, gen in enumerate(genomes):
            idx = np.flatnonzero(src[a:b] & (gsel[a:b] == g))
            if idx.size == 0:
                continue
            start = rng.integers(0, len(gen) - L + 1, size=idx.size)
            take = gen[start[:, None] + np.arange(L)[None, :]]
            rv = strand[a:b][idx]
            take[rv] = COMP_ASCII[take[rv][:, ::-1]]
            block[idx] = take
        if subst > 0:
            mask = (rng.random((m, L)) < subst) & src[a:b, None]
            cur = np.searchsorted(ASCII, block[mask])
            block[mask] = ASCII[(cur + rng.integers(1, 4, size=cur.size)) % 4]
        if n_prob > 0:
            block[rng.random((m, L)) < n_prob] = ord("N")
        out[a:b] = block
    return out, src, gsel


def paired_reads(rng, genomes, npairs, L, subst, from_frac, frag_mean=300, frag_sd=30,
                 chunk=200_000):
    """npairs pairs; mate 2 is the reverse complement of the fragment end."""
    r1 = np.empty((npairs, L), np.uint8)
    r2 = np.empty((npairs, L), np.uint8)
    src = rng.random(npairs) < from_frac
    gsel = rng.integers(0, len(genomes), size=npairs)
    for a in range(0, npairs, chunk):
        b = min(npairs, a + chunk)
        m = b - a
        b1 = codes_to_ascii(rng.integers(0, 4, size=(m, L), dtype=np.uint8))
        b2 = codes_to_ascii(rng.integers(0, 4, size=(m, L), dtype=np.uint8))
        frag = np.maximum(L, np.round(rng.normal(frag_mean, frag_sd, size=m))).astype(np.int64)
        strand = rng.random(m) < 0.5
        for g, gen in enumerate(genomes):
            idx = np.flatnonzero(src[a:b] & (gsel[a:b] == g))
            if idx.size == 0:
                continue
            fl = frag[idx]
            start = rng.integers(0, len(gen) - fl + 1)
            ar = np.arange(L)[None, :]
            fw_first = gen[start[:, None] + ar]
            fw_last = gen[(start + fl - L)[:, None] + ar]
            # forward fragment: mate1 = first L, mate2 = revcomp(last L)
            m1 = fw_first.copy()
            m2 = COMP_ASCII[fw_last[:, ::-1]]
            # reverse-strand fragment: the roles swap
            rv = strand[idx]
            m1[rv] = COMP_ASCII[fw_last[rv][:, ::-1]]
            m2[rv] = fw_first[rv]
            b1[idx] = m1
            b2[idx] = m2
        if subst > 0:
            for blk in (b1, b2):
                mask = (rng.random((m, L)) < subst) & src[a:b, None]
                cur = np.searchsorted(ASCII, blk[mask])
                blk[mask] = ASCII[(cur + rng.integers(1, 4, size=cur.size)) % 4]
        r1[a:b] = b1
        r2[a:b] = b2
    return r1, r2, src, gsel


def long_reads(rng, genomes, n, median=2000, sigma=0.75, lo=100, hi=50_000,
               sub=0.05, ins=0.015, dele=0.015):
    """n log-normal-length reads from the genomes with substitutions and
    indels; returns a list of uint8 arrays."""
    lens = np.clip(np.round(rng.lognormal(np.log(median), sigma, size=n)), lo, hi)
    lens = lens.astype(np.int64)
    gsel = rng.integers(0, len(genomes), size=n)
    strand = rng.random(n) < 0.5
    out = []
    for i in range(n):
        gen = genomes[gsel[i]]
        L = int(lens[i])
        tl = L + L // 10 + 16  # template long enough for deletions
        st = int(rng.integers(0, len(gen) - tl))
        t = gen[st:st + tl].copy()
        u = rng.random(tl)
        is_del = u < dele
        is_ins = (u >= dele) & (u < dele + ins)
        is_sub = (u >= dele + ins) & (u < dele + ins + sub)
        cur = np.searchsorted(ASCII, t[is_sub])
        t[is_sub] = ASCII[(cur + rng.integers(1, 4, size=cur.size)) % 4]
        keep = t[~is_del]
        ins_at = np.flatnonzero(is_ins[~is_del])
        if ins_at.size:
            keep = np.insert(keep, ins_at, codes_to_ascii(
                rng.integers(0, 4, size=ins_at.size, dtype=np.uint8)))
        r = keep[:L]
        if strand[i]:
            r = revcomp(r)
        out.append(np.ascontiguousarray(r))
    return out, gsel


# ---------------------------------------------------------------- workloads

def workload_single(scale=1.0):
    rng = rng_for(1)
    g = codes_to_ascii(random_codes(rng, int(5_000_000 * scale) if scale < 1 else 5_000_000))
    n = max(10, int(100_000 * scale))
    reads, src, _ = fixed_len_reads(rng, [g], n, 150, 0.01, 0.5, n_prob=0.001)
    return {"genomes": [g], "k": 25, "fpr": 0.075, "reads": reads, "paired": False,
            "src": src}


def workload_multi(scale=1.0, nfilters=10, genome_len=20_000_000):
    rng = rng_for(2)
    gcs = rng.uniform(0.35, 0.65, size=nfilters)
    glen = max(10_000, int(genome_len * min(1.0, scale)))
    genomes = [codes_to_ascii(random_codes(rng_for(2, 1000 + g), glen, gcs[g]))
               for g in range(nfilters)]
    npairs = max(10, int(2_000_000 * scale))
    r1, r2, src, gsel = paired_reads(rng, genomes, npairs, 150, 0.01, 0.7)
    reads = np.empty((2 * npairs, 150), np.uint8)
    reads[0::2] = r1
    reads[1::2] = r2
    return {"genomes": genomes, "k": 25, "fpr": 0.075, "reads": reads, "paired": True,
            "src": src, "gsel": gsel}


def workload_uneven(scale=1.0):
    rng = rng_for(3)
    glen = max(100_000, int(100_000_000 * min(1.0, scale)))
    genomes = [codes_to_ascii(random_codes(rng_for(3, 1000 + g), glen)) for g in range(4)]
    nlong = max(4, int(200_000 * scale))
    nshort = max(10, int(2_000_000 * scale))
    longs, _ = long_reads(rng, genomes, nlong)
    shorts, _, _ = fixed_len_reads(rng, genomes, nshort, 150, 0.01, 0.7)
    items = longs + list(shorts)
    perm = rng.permutation(len(items))
    reads = [items[i] for i in perm]
    return {"genomes": genomes, "k": 25, "fpr": 0.05, "reads_list": reads,
            "paired": False}


def mibf_seeds(rng, span=25, weight=20):
    """Three spaced seeds closed under reversal: s0 random with both ends
    set, s2 = reverse(s0), s1 a palindrome with the centre unset and
    weight/2 of the mirrored pairs (the end pair included) set."""
    inner = np.zeros(span - 2, np.uint8)
    inner[rng.choice(span - 2, size=weight - 2, replace=False)] = 1
    s0 = "1" + "".join(map(str, inner)) + "1"
    half = span // 2
    pairs = np.zeros(half, np.uint8)
    pairs[0] = 1
    pairs[1 + rng.choice(half - 1, size=weight // 2 - 1, replace=False)] = 1
    left = "".join(map(str, pairs))
    s1 = left + ("0" if span % 2 else "") + left[::-1]
    return [s0, s1, s0[::-1]]


def workload_mibf(scale=1.0, nrefs=200, ref_len=5_000_000, seg=10_000):
    rng = rng_for(4)
    rlen = max(4 * seg, int(ref_len * min(1.0, scale)))
    rlen -= rlen % seg
    orig = [codes_to_ascii(rng_for(4, 1000 + g).integers(0, 4, size=rlen, dtype=np.uint8))
            for g in range(nrefs)]
    nseg = max(1, int(round(0.05 * rlen / seg)))
    refs = []
    for g in range(nrefs):
        r = orig[g].copy()
        slots = rng.choice(rlen // seg, size=nseg, replace=False)
        for sl in slots:
            other = int(rng.integers(0, nrefs - 1))
            other += other >= g
            src = int(rng.integers(0, rlen // seg)) * seg
            r[sl * seg:(sl + 1) * seg] = orig[other][src:src + seg]
        refs.append(r)
    del orig
    seeds = mibf_seeds(rng)
    n = max(10, int(2_000_000 * scale))
    reads, src, gsel = fixed_len_reads(rng, refs, n, 150, 0.01, 0.9)
    names = [f"ref{g:03d}" for g in range(nrefs)]
    return {"refs": refs, "names": names, "seeds": seeds, "reads": reads, "src": src,
            "gsel": gsel, "paired": False, "genomes": refs}


def repeat_genome(rng, n, nrep=1000, rep_len=300, frac=0.5, div=0.10, records=24,
                  chunk=1 << 20):
    """n bases: uniform random, then a `frac` share of the rep_len slots
    overwritten by copies of nrep random consensuses with i.i.d. `div`
    substitutions per copy; records - 1 'N' bytes split it into records."""
    g = codes_to_ascii(rng.integers(0, 4, size=n, dtype=np.uint8))
    cons = codes_to_ascii(rng.integers(0, 4, size=(nrep, rep_len), dtype=np.uint8))
    nslots = n // rep_len
    view = g[: nslots * rep_len].reshape(nslots, rep_len)
    chosen = np.flatnonzero(rng.random(nslots) < frac)
    for a in range(0, chosen.size, chunk):
        idx = chosen[a:a + chunk]
        cp = cons[rng.integers(0, nrep, size=idx.size)]
        mask = rng.random(cp.shape) < div
        cur = np.searchsorted(ASCII, cp[mask])
        cp[mask] = ASCII[(cur + rng.integers(1, 4, size=cur.size)) % 4]
        view[idx] = cp
    for r in range(1, records):
        g[r * (n // records)] = ord("N")
    return g


def workload_serve(scale=1.0, npool=48, min_reads=1000, max_reads=20_000):
    rng = rng_for(5)
    big = repeat_genome(rng_for(5, 1), max(2_000_000, int(3_000_000_000 * min(1.0, scale))))
    clen = max(200_000, int(20_000_000 * min(1.0, scale)))
    contams = [codes_to_ascii(rng_for(5, 100 + c).integers(0, 4, size=clen, dtype=np.uint8))
               for c in range(8)]
    pool = []
    for q in range(npool):
        n = int(rng.integers(min_reads, max_reads + 1))
        nlong = int(rng.binomial(n, 0.2))
        nshort = n - nlong
        origin = rng.choice(3, size=n, p=[0.5, 0.1, 0.4])
        items = []
        # short reads: big genome, contaminants, random
        s_or = origin[:nshort]
        for src, gens in ((0, [big]), (1, contams)):
            k = int(np.count_nonzero(s_or == src))
            if k:
                r, _, _ = fixed_len_reads(rng, gens, k, 150, 0.01, 1.0)
                items.extend(list(r))
        k = int(np.count_nonzero(s_or == 2))
        if k:
            items.extend(list(codes_to_ascii(rng.integers(0, 4, size=(k, 150), dtype=np.uint8))))
        l_or = origin[nshort:]
        for src, gens in ((0, [big]), (1, contams)):
            k = int(np.count_nonzero(l_or == src))
            if k:
                r, _ = long_reads(rng, gens, k)
                items.extend(r)
        k = int(np.count_nonzero(l_or == 2))
        if k:
            lens = np.clip(np.round(rng.lognormal(np.log(2000), 0.75, size=k)), 100, 50_000)
            items.extend(codes_to_ascii(rng.integers(0, 4, size=int(L), dtype=np.uint8))
                         for L in lens)
        perm = rng.permutation(len(items))
        pool.append(pack([items[i] for i in perm]))
    return {"genomes": [big] + contams, "k": 25, "fpr": 0.075, "requests": pool,
            "paired": False}
