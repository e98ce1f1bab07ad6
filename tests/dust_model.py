"""Brute-force model of the symmetric DUST definition (Morgulis et al. 2006)
for A/C/G/T-only sequences, independent of the sdust procedure.  Test
infrastructure.

Words are the triplets starting at each base.  A run of words [a, b] has
counts c_t and sum r = sum c_t (c_t - 1) / 2; with l = b - a (words - 1) it is
above the threshold when 10 r > T l.  It is perfect when above the threshold
and its score r / l is at least that of every perfect run inside it.  The
mask is the union of the base intervals [a, b + 3) of the perfect runs that
fit in W bases (b + 3 - a <= W), merged into disjoint intervals with adjacent
ones joined.
"""
CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def intervals(seq, T=20, W=64):
    s = seq.upper()
    assert set(s) <= set("ACGT")
    n = len(s)
    nw = n - 2
    if nw < 2:
        return []
    words = [CODE[s[i]] * 16 + CODE[s[i + 1]] * 4 + CODE[s[i + 2]] for i in range(nw)]
    maxw = W - 2  # words in a window of W bases
    # best[(a, b)] = (r, l) of the best perfect run inside [a, b], or None
    best = {}
    masked = []
    for length in range(1, maxw + 1):
        for a in range(0, nw - length + 1):
            b = a + length - 1
            cnt = {}
            r = 0
            for t in words[a:b + 1]:
                r += cnt.get(t, 0)
                cnt[t] = cnt.get(t, 0) + 1
            l = b - a
            inner = None
            if length > 1:
                for y in (best.get((a + 1, b)), best.get((a, b - 1))):
                    if y is not None and (inner is None or y[0] * inner[1] > inner[0] * y[1]):
                        inner = y
            perfect = l > 0 and 10 * r > T * l and (inner is None or r * inner[1] >= inner[0] * l)
            if perfect:
                masked.append((a, b + 3))
                best[(a, b)] = (r, l)
            else:
                best[(a, b)] = inner
    masked.sort()
    out = []
    for a, f in masked:
        if out and a <= out[-1][1]:
            out[-1][1] = max(out[-1][1], f)
        else:
            out.append([a, f])
    return [tuple(x) for x in out]


def low_complexity(iv, pos):
    return any(a <= pos < f for a, f in iv)
