"""Pure-Python model of the SeqEval walks, written from Common/SeqEval.h as an
independent second implementation.  It consumes the list of N-free window
starts and a per-window contains() answer, so it checks the walk logic apart
from hashing.  Test infrastructure only.
"""
import math


def frames(length, k):
    # size_t arithmetic of seqLen - kmerSize + 1 (SeqEval.h:28-35)
    return (length - k + 1) % (1 << 64)


class Walk:
    """Replays windows in order; probe(i) answers contains() for window i and
    counts the calls."""

    def __init__(self, positions, hits):
        self.positions = list(positions)
        self.hits = list(hits)
        self.probes = []

    def probe(self, i):
        self.probes.append(i)
        return self.hits[i]


def simple(length, k, positions, hits, t, streak_thr=3, harmonic=False, sub=None):
    """evalSimple (SeqEval.h:37-114) / evalHarmonic (:116-193); sub[i] says
    whether the subtract filter contains window i (a hit it holds scores
    nothing, SeqEval.h:55, 81, 84)."""
    w = Walk(positions, hits)
    sub = sub or [False] * len(w.positions)
    n = float(frames(length, k))
    thres = t * n
    anti_thres = math.floor((1.0 - t) * n)
    score, anti, streak = 0.0, 0, 0
    i, npos = 0, len(w.positions)
    prev = 0
    if npos:
        if w.probe(0):
            if not sub[0]:
                score += 0.5
            if thres <= score:
                return True, w
            streak += 1
        else:
            anti += 1
            if anti_thres <= anti:
                return False, w
        prev = w.positions[0]
        i = 1
    while i < npos:
        p = w.positions[i]
        if p != prev + 1:
            anti += p - prev - 1
            if anti_thres <= anti:
                return False, w
            streak = 0
        if w.probe(i):
            if sub[i]:
                pass
            elif streak == 0:
                score += 0.5
            elif harmonic:
                score += 1.0 - 1.0 / (1.0 + float(streak))
            else:
                score += 1
            if thres <= score:
                return True, w
            prev = p
            i += 1
            streak += 1
        else:
            if streak < streak_thr:
                anti += 1
                if anti_thres <= anti:
                    return False, w
                prev = p
                i += 1
            else:
                skip_end = p + k
                while i < npos and w.positions[i] < skip_end:
                    anti += 1
                    if anti_thres <= anti:
                        return False, w
                    prev = w.positions[i]
                    i += 1
            streak = 0
    return False, w


def binomial(length, k, positions, hits, thres, streak_thr=3):
    """evalBinomial (SeqEval.h:218-290) with a precomputed integer threshold."""
    w = Walk(positions, hits)
    if length < k:
        return False, w
    frame_len = length - k + 1
    anti_thres = (frame_len - thres) % (1 << 32)
    score = anti = streak = 0
    i, npos, prev = 0, len(w.positions), 0
    if npos:
        if w.probe(0):
            score += 1
            if thres <= score:
                return True, w
            streak += 1
        else:
            anti += 1
            if anti_thres <= anti:
                return False, w
        prev = w.positions[0]
        i = 1
    while i < npos:
        p = w.positions[i]
        if p != prev + 1:
            anti += p - prev - 1
            if anti_thres <= anti:
                return False, w
            streak = 0
        if w.probe(i):
            score += 1
            if thres <= score:
                return True, w
            prev = p
            i += 1
            streak += 1
        else:
            if streak < streak_thr:
                anti += 1
                if anti_thres <= anti:
                    return False, w
                prev = p
                i += 1
            else:
                skip_end = p + k
                while i < npos and w.positions[i] < skip_end:
                    anti += 1
                    if anti_thres <= anti:
                        return False, w
                    prev = w.positions[i]
                    i += 1
            streak = 0
    return False, w


def min_match_len(length, k, positions, hits, min_len, sub=None):
    """evalMinMatchLen (SeqEval.h:295-332); a hit the subtract filter holds
    (sub[i]) leaves the match length as it is (:316)."""
    w = Walk(positions, hits)
    sub = sub or [False] * len(w.positions)
    match, prev = 0, 0
    for i, p in enumerate(w.positions):
        if length - p + match < min_len:
            return False, w
        if p != prev + 1:
            match = 0
        if w.probe(i):
            if not sub[i]:
                match = k if match == 0 else match + 1
        else:
            match = 0
        if match >= min_len:
            return True, w
        prev = p
    return False, w


def tiling_score(length, k, positions, hits, streak_thr=3, kind="simple"):
    """evalSimpleScore / evalHarmonicScore / evalBinomialScore count
    (SeqEval.h:334-491).  Returns (normalised score, hit count, walk)."""
    w = Walk(positions, hits)
    score, count, streak = 0.0, 0, 0
    i, npos, prev = 0, len(w.positions), 0
    while i < npos:
        p = w.positions[i]
        if p != prev + 1:
            streak = 0
        if w.probe(i):
            count += 1
            if streak == 0:
                score += 0.5
            elif kind == "harmonic":
                score += 1.0 - 1.0 / (1.0 + float(streak))
            else:
                score += 1
            prev = p
            i += 1
            streak += 1
        else:
            if streak < streak_thr:
                prev = p
                i += 1
            else:
                skip_end = p + k
                while i < npos and w.positions[i] < skip_end:
                    prev = w.positions[i]
                    i += 1
            streak = 0
    n = float(frames(length, k))
    norm = score / n if n != 0 else (math.nan if score == 0 else math.inf)
    return norm, count, w


def min_match_len_score(length, k, positions, hits):
    w = Walk(positions, hits)
    match, prev = 0, 0
    for i, p in enumerate(w.positions):
        if p != prev + 1:
            match = 0
        match = (k if match == 0 else match + 1) if w.probe(i) else 0
        prev = p
    return match, w
