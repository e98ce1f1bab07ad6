"""Pure-Python model of the MIBF build and query as specified in
oracle/bbt_mibf.cpp (an independent second implementation, small inputs
only).  Test infrastructure."""
import math

import oracle_lib as O

SAT = 0x80000000
MASK = 0x7FFFFFFF


def mix32(x):
    M = (1 << 64) - 1
    x ^= x >> 33
    x = (x * 0xff51afd7ed558ccd) & M
    x ^= x >> 33
    x = (x * 0xc4ceb9fe1a85ec53) & M
    x ^= x >> 33
    return x >> 32


def frames(seq, seeds):
    if isinstance(seeds, O.Kmer):
        pos, hv = O.nthash(seq, seeds.k, seeds.h)
    else:
        pos, hv = O.sthash(seq, seeds, len(seeds[0]))
    return [int(p) for p in pos], [[int(v) for v in row] for row in hv]


def build(records, seeds, m):
    """records: list of (id, seq). Returns (set-bit list, data list)."""
    bits = set()
    fr = {}
    for rid, s in records:
        pos, hv = frames(s, seeds)
        fr[rid] = hv
        for row in hv:
            for h in row:
                bits.add(h % m)
    order = sorted(bits)
    slot = {p: i for i, p in enumerate(order)}
    best = [0] * len(order)
    for rid, s in records:
        for row in fr[rid]:
            for h in row:
                key = (mix32(h ^ ((rid * 0x9E3779B97F4A7C15) & ((1 << 64) - 1))) << 32) | rid
                i = slot[h % m]
                best[i] = max(best[i], key)
    data = [b & MASK for b in best]
    for rid, s in records:
        for row in fr[rid]:
            idx = [slot[h % m] for h in row]
            if not any((data[i] & MASK) == rid for i in idx):
                for i in idx:
                    data[i] |= SAT
    return order, data


def frame_probs(data, pop, m, S, a, nids):
    occ = pop / m
    n = [0] * (nids + 1)
    for d in data:
        if (d & MASK) <= nids:
            n[d & MASK] += 1
    out = [0.0]
    for i in range(1, nids + 1):
        f = n[i] / pop if pop else 0.0
        P = 0.0
        for j in range(max(0, S - a), S + 1):
            P += math.comb(S, j) * occ ** j * (1 - occ) ** (S - j) * (1 - (1 - f) ** j)
        out.append(P)
    return out


def query(order, data, m, seeds, probs, nids, score, a, multi, min_nonsat, seq, seq2=None,
          best_agree=False):
    """Hits of one read (or of the pair seq, seq2) and its satCount."""
    S, k, _ = O.seed_shape(seeds)
    if seq2 is None:
        fc = len(seq) - k + 1 if len(seq) >= k else 0
        mates = [seq]
    else:
        # the precedence of MIBFClassifier.hpp:804-805
        fc = (len(seq) - k + 1 if len(seq) >= k else
              (len(seq2) - k + 1 if len(seq2) >= k else 0))
        mates = [seq, seq2]
    if fc == 0:
        return [], 0
    slot = {p: i for i, p in enumerate(order)}
    acc = {}
    sat = 0
    for s in mates:
        pos, hv = frames(s, seeds)
        for row in hv:
            raw = [data[slot[h % m]] for h in row if (h % m) in slot]
            if len(raw) + a < S:
                continue
            if raw and all(r & SAT for r in raw):
                sat += 1
            clean = not any(r & SAT for r in raw)
            for iid in {r & MASK for r in raw}:
                occ = [r for r in raw if (r & MASK) == iid]
                ns = sum(1 for r in occ if not r & SAT)
                h = acc.setdefault(iid, [0, 0, 0, 0, 0, 0])
                h[0] += 1
                h[1] += 1 if ns else 0
                h[2] += len(occ)
                h[3] += ns
                h[4] += 1 if len(occ) == S else 0
                h[5] += 1 if clean else 0
    cand = []
    for iid, (cnt, nsat, tot, tns, solid, nsf) in acc.items():
        if not 1 <= iid <= nids:  # ID 0 (no build writes it) is never a candidate
            continue
        mc = max(1, int(O.lib().orc_binom_isf_up(score / nids, fc, probs[iid])))
        if cnt >= mc and nsat >= min_nonsat:
            cand.append((iid, cnt, nsat, tot, tns, solid, nsf))
    cand.sort(key=lambda c: (-c[2], -c[1], -c[5], -c[4], -c[3], c[0]))
    if not cand:
        return [], sat
    if best_agree:
        top = cand[0]
        if any(c[f] > top[f] for c in cand for f in range(1, 7)):
            return [], sat
    bar = multi * cand[0][2]
    return [c for c in cand if c[2] >= bar], sat
