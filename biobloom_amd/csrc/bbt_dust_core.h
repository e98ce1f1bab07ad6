/*
 * bbt_dust_core.h -- sdust's low-complexity intervals of one sequence, as one
 * thread's sequential pass, over a storage view chosen by the caller.
 *
 * Shared by the gfx950 kernels (bbt_dust.h: window state in LDS, or in a
 * per-thread device-memory region) and by the host-side self test, which
 * compiles the same code for the CPU.
 *
 * The steps are sdust_core's (Common/sdust.c:139-165): shift the word window
 * (:75-95), report the perfect intervals that leave the window, merging
 * adjacent ones (:97-111), collect the perfect suffixes (:113-137); an N
 * resets only the word and the run length (:156-160), so the window carries
 * the words before it.  Intervals are clipped to the sequence: SDust::isLowComp
 * (Common/SDust.hpp:38-48) is only asked about window starts inside it.
 *
 * A view V provides
 *   cv(t), cw(t), c(t)  references to the word counts (64 words): from the
 *                       last perfect boundary, over the whole window, and a
 *                       work copy of cv
 *   ring(i)             reference to slot i of the window's words, i < ring
 *                       capacity (a power of two, at least W - 2)
 *   rmask()             ring capacity - 1
 *   pcap()              entries the perfect-interval list can hold
 *   P(j), setP(j, x)    entry j of that list
 *   kSepNew             true: a separate list N(k), setN(k, x) of at least
 *                       W - 2 entries holds one perfect() call's new entries
 * The source S gives byte i of the sequence through at(i), called with
 * i = 0, 1, 2, ... in order.  emit(a, f) receives each merged interval
 * [a, f) of window starts, already clipped to [0, n).
 */
#ifndef BBT_DUST_CORE_H
#define BBT_DUST_CORE_H

#include <stdint.h>

#if defined(__HIPCC__)
#define BBT_DHD __host__ __device__ __forceinline__
#else
#define BBT_DHD inline
#endif

namespace bbt {

struct DustPerf {
    int32_t start, finish, r, l;
};

/* largest -W of the LDS kernel: its ring holds 64 words and its 8-bit counts
 * stay below W */
constexpr int kDustFastW = 64;
/* perfect intervals an LDS lane keeps (sequences that need more go to the
 * device-memory kernel) */
constexpr uint32_t kDustFastP = 16;

BBT_DHD int dust_code(unsigned c) {
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

/* Returns false, at once, when the perfect-interval list would need more
 * than v.pcap() entries (the intervals emitted so far are then incomplete);
 * true when the whole sequence was taken. */
template <class V, class S, class E>
BBT_DHD bool dust_core(S &src, uint32_t n, int T, int W, V &v, E &emit) {
    for (unsigned t = 0; t < 64; ++t) {
        v.cv(t) = 0;
        v.cw(t) = 0;
    }
    const uint32_t rmask = v.rmask();
    const uint32_t pcap = v.pcap();
    uint32_t head = 0, cnt = 0, np = 0;
    int rv = 0, rw = 0, L = 0;
    int res_s = -1, res_f = -1; /* the last reported interval, still mergeable */
    const auto report = [&](int a, int f) {
        a = a < 0 ? 0 : a;
        f = f > int(n) ? int(n) : f;
        if (a < f)
            emit(a, f);
    };
    const auto flush = [&](int start) {
        if (np == 0 || v.P(np - 1).start >= start)
            return;
        const DustPerf p = v.P(np - 1);
        if (res_s >= 0 && p.start <= res_f) {
            res_f = res_f > p.finish ? res_f : p.finish;
        } else {
            if (res_s >= 0)
                report(res_s, res_f);
            res_s = p.start;
            res_f = p.finish;
        }
        while (np && v.P(np - 1).start < start)
            --np;
    };
    /* false: the list is full.  The entries that start at or after
     * i + start (a prefix of the list, which is kept in descending order of
     * start) are folded into (max_r, max_l) once each, as the prefix grows
     * with falling i: folding an entry again cannot change the maximum
     * ratio, which is all the insertion test reads.  An insertion goes after
     * every entry with the same or a later start, as the reference's scan
     * puts it.  With V::kSepNew the call's new entries (each starting before
     * the one before it) are kept apart and merged into the list at the
     * end; otherwise each is shifted into place at once. */
    const auto perfect = [&](int start) -> bool {
        for (unsigned t = 0; t < 64; ++t)
            v.c(t) = v.cv(t);
        int r = rv, max_r = 0, max_l = 0;
        const int wn = int(cnt);
        uint32_t jo = 0, nn = 0; /* old entries folded, new entries */
        for (int i = wn - L - 1; i >= 0; --i) {
            const unsigned tt = v.ring((head + uint32_t(i)) & rmask);
            r += v.c(tt)++;
            const int new_r = r, new_l = wn - i - 1;
            if (new_r * 10 > T * new_l) {
                for (; jo < np; ++jo) {
                    const DustPerf q = v.P(V::kSepNew ? jo : jo + nn);
                    if (q.start < i + start)
                        break;
                    if (max_r == 0 || q.r * max_l > max_r * q.l) {
                        max_r = q.r;
                        max_l = q.l;
                    }
                }
                if (max_r == 0 || new_r * max_l >= max_r * new_l) {
                    max_r = new_r;
                    max_l = new_l;
                    if (np + nn >= pcap)
                        return false;
                    const DustPerf e{i + start, wn + 2 + start, new_r, new_l};
                    if constexpr (V::kSepNew) {
                        v.setN(nn, e);
                    } else {
                        const uint32_t j = jo + nn;
                        for (uint32_t x = np + nn; x > j; --x)
                            v.setP(x, v.P(x - 1));
                        v.setP(j, e);
                    }
                    ++nn;
                }
            }
        }
        if constexpr (V::kSepNew) {
            /* merge from the back: the smaller start goes last, and of equal
             * starts the new entry goes after the old one */
            int o = int(np) - 1, k = int(nn) - 1;
            for (uint32_t w = np + nn; k >= 0;) {
                --w;
                const DustPerf e = v.N(uint32_t(k));
                if (o < 0 || e.start <= v.P(uint32_t(o)).start) {
                    v.setP(w, e);
                    --k;
                } else {
                    v.setP(w, v.P(uint32_t(o)));
                    --o;
                }
            }
        }
        np += nn;
        return true;
    };
    int l = 0;
    unsigned t = 0;
    for (uint32_t i = 0; i <= n; ++i) {
        const int b = i < n ? dust_code(src.at(i)) : 4;
        if (b < 4) {
            ++l;
            t = ((t << 2) | unsigned(b)) & 63u;
            if (l >= 3) {
                const int start = (l - W > 0 ? l - W : 0) + (int(i) + 1 - l);
                flush(start);
                if (int(cnt) >= W - 2) {
                    const unsigned s0 = v.ring(head);
                    head = (head + 1) & rmask;
                    --cnt;
                    rw -= --v.cw(s0);
                    if (L > int(cnt)) {
                        --L;
                        rv -= --v.cv(s0);
                    }
                }
                v.ring((head + cnt) & rmask) = t;
                ++cnt;
                ++L;
                rw += v.cw(t)++;
                rv += v.cv(t)++;
                if (v.cv(t) * 10 > T * 2) {
                    unsigned s0;
                    do {
                        s0 = v.ring((head + cnt - uint32_t(L)) & rmask);
                        rv -= --v.cv(s0);
                        --L;
                    } while (s0 != t);
                }
                if (rw * 10 > L * T && !perfect(start))
                    return false;
            }
        } else {
            int start = (l - W + 1 > 0 ? l - W + 1 : 0) + (int(i) + 1 - l);
            while (np)
                flush(start++);
            l = 0;
            t = 0;
        }
    }
    if (res_s >= 0)
        report(res_s, res_f);
    return true;
}

/* Entries the perfect-interval list can hold at once, at most.  A call of
 * perfect() inserts at most W - 2 entries, with starts in
 * [start, start + W - 3], and flush(start) removes every entry that starts
 * before start.  Within a run of A/C/G/T, start stays at the run's first
 * word for at most W - 2 steps and then moves by one per step; an N empties
 * the list.  So the live entries come from at most W - 2 steps of the fixed
 * start and the last W - 2 moving steps: 2 (W - 2)^2.  (After an N the
 * window keeps its words, so one step can insert W - 2 entries, some with
 * the same (start, finish) as earlier ones.) */
BBT_DHD uint64_t dust_pcap_bound(int W) {
    return 2 * uint64_t(W - 2) * uint64_t(W - 2) + 64;
}

} // namespace bbt

#endif
