/*
 * bbt_dust.h -- the -D pre-pass on gfx950: sdust low-complexity intervals of
 * every read of a batch, as a bitmap of window starts that the walks consult
 * (bit off[read] + p set: window p is treated as a miss without a
 * contains() call, SeqEval.h:53 and the same test in every scorer).
 *
 * The steps are sdust_core's (Common/sdust.c:139-165): shift the word window
 * (:75-95), report the perfect intervals that leave the window, merging
 * adjacent ones (:97-111), collect the perfect suffixes (:113-137); an N
 * resets only the word and the run length (:156-160), so the window carries
 * the words before it.  Intervals are clipped to the read: SDust::isLowComp
 * (Common/SDust.hpp:38-48) is only asked about window starts inside it.
 */
#ifndef BBT_DUST_H
#define BBT_DUST_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bbt {

/* largest -W: word counts are 16-bit, and the perfect-interval list of one
 * window grows as W^2 (below) */
constexpr int kDustMaxW = 4096;
/* threads of the pre-pass at most (each owns one scratch region) */
constexpr uint32_t kDustThreads = 8192;
/* the pre-pass's scratch for all threads at most (large W: fewer threads) */
constexpr uint64_t kDustScratchBudget = uint64_t(2) << 30;

struct DustPerf {
    int32_t start, finish, r, l;
};

/* Per-thread scratch of one sdust pass, in device memory: counts of the 64
 * words in the window (cv: from the last perfect boundary, cw: whole
 * window) and a work copy c, the window's words as a ring of ring_cap >=
 * W - 2 entries, then the perfect-interval list, which holds distinct
 * (start, finish) word pairs of one window: at most (W-2)(W-1)/2 of them. */
struct DustLayout {
    uint32_t ring_cap; /* a power of two */
    uint32_t pcap;
    uint64_t perf_at;  /* byte offset of the list in a thread's region */
    uint64_t stride;   /* bytes per thread */
    uint32_t threads;  /* threads of the pre-pass */
};

inline DustLayout dust_layout(int W) {
    DustLayout d;
    d.ring_cap = 64;
    while (int(d.ring_cap) < W - 2)
        d.ring_cap <<= 1;
    d.pcap = uint32_t(W - 2) * uint32_t(W - 1) / 2 + 64;
    d.perf_at = ((3 * 64 + uint64_t(d.ring_cap)) * 2 + 15) / 16 * 16;
    d.stride = (d.perf_at + uint64_t(d.pcap) * sizeof(DustPerf) + 255) / 256 * 256;
    const uint64_t fit = kDustScratchBudget / d.stride / 64 * 64;
    d.threads = uint32_t(fit < 64 ? 64 : (fit > kDustThreads ? kDustThreads : fit));
    return d;
}

__device__ __forceinline__ int dust_nt4(unsigned c) {
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

__device__ inline void dust_read(const uint8_t *s, uint32_t n, int T, int W, uint8_t *region,
                                 const DustLayout &lay, uint32_t *mask, uint64_t gbase,
                                 bool &overflow) {
    struct View {
        uint16_t *cv, *cw, *c, *ring;
        DustPerf *P;
    } S;
    S.cv = reinterpret_cast<uint16_t *>(region);
    S.cw = S.cv + 64;
    S.c = S.cw + 64;
    S.ring = S.c + 64;
    S.P = reinterpret_cast<DustPerf *>(region + lay.perf_at);
    const uint32_t rmask = lay.ring_cap - 1;
    const uint32_t kDustPCap = lay.pcap;
    for (int t = 0; t < 64; ++t) {
        S.cv[t] = 0;
        S.cw[t] = 0;
    }
    uint32_t head = 0, cnt = 0, np = 0;
    int rv = 0, rw = 0, L = 0;
    int res_s = -1, res_f = -1; /* the last reported interval, still mergeable */
    const auto emit = [&](int a, int f) {
        a = a < 0 ? 0 : a;
        f = f > int(n) ? int(n) : f;
        for (int p = a; p < f;) {
            const uint64_t g = gbase + uint64_t(p);
            const uint32_t bit = uint32_t(g & 31);
            const int take = min(32 - int(bit), f - p);
            const uint32_t m = take == 32 ? 0xFFFFFFFFu : (((1u << take) - 1u) << bit);
            atomicOr(mask + (g >> 5), m);
            p += take;
        }
    };
    const auto flush = [&](int start) {
        if (np == 0 || S.P[np - 1].start >= start)
            return;
        const DustPerf p = S.P[np - 1];
        if (res_s >= 0 && p.start <= res_f) {
            res_f = res_f > p.finish ? res_f : p.finish;
        } else {
            if (res_s >= 0)
                emit(res_s, res_f);
            res_s = p.start;
            res_f = p.finish;
        }
        while (np && S.P[np - 1].start < start)
            --np;
    };
    const auto perfect = [&](int start) {
        for (int t = 0; t < 64; ++t)
            S.c[t] = S.cv[t];
        int r = rv, max_r = 0, max_l = 0;
        const int wn = int(cnt);
        for (int i = wn - L - 1; i >= 0; --i) {
            const unsigned tt = S.ring[(head + uint32_t(i)) & rmask];
            r += S.c[tt]++;
            const int new_r = r, new_l = wn - i - 1;
            if (new_r * 10 > T * new_l) {
                uint32_t j = 0;
                for (; j < np && S.P[j].start >= i + start; ++j) {
                    const int pr = S.P[j].r, pl = S.P[j].l;
                    if (max_r == 0 || pr * max_l > max_r * pl) {
                        max_r = pr;
                        max_l = pl;
                    }
                }
                if (max_r == 0 || new_r * max_l >= max_r * new_l) {
                    max_r = new_r;
                    max_l = new_l;
                    if (np >= kDustPCap) {
                        overflow = true;
                        continue;
                    }
                    for (uint32_t x = np; x > j; --x)
                        S.P[x] = S.P[x - 1];
                    S.P[j] = DustPerf{i + start, wn + 2 + start, new_r, new_l};
                    ++np;
                }
            }
        }
    };
    int l = 0;
    unsigned t = 0;
    for (uint32_t i = 0; i <= n; ++i) {
        const int b = i < n ? dust_nt4(s[i]) : 4;
        if (b < 4) {
            ++l;
            t = ((t << 2) | unsigned(b)) & 63u;
            if (l >= 3) {
                const int start = (l - W > 0 ? l - W : 0) + (int(i) + 1 - l);
                flush(start);
                if (int(cnt) >= W - 2) {
                    const unsigned s0 = S.ring[head];
                    head = (head + 1) & rmask;
                    --cnt;
                    rw -= --S.cw[s0];
                    if (L > int(cnt)) {
                        --L;
                        rv -= --S.cv[s0];
                    }
                }
                S.ring[(head + cnt) & rmask] = uint16_t(t);
                ++cnt;
                ++L;
                rw += S.cw[t]++;
                rv += S.cv[t]++;
                if (S.cv[t] * 10 > T * 2) {
                    unsigned s0;
                    do {
                        s0 = S.ring[(head + cnt - uint32_t(L)) & rmask];
                        rv -= --S.cv[s0];
                        --L;
                    } while (s0 != t);
                }
                if (rw * 10 > L * T)
                    perfect(start);
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
        emit(res_s, res_f);
}

/* One thread per read from a queue; each thread owns lay.stride bytes of
 * scratch. */
__global__ __launch_bounds__(64) void k_dust(const uint8_t *seq, const uint64_t *off,
                                             uint64_t nreads, int T, int W, uint8_t *scratch,
                                             DustLayout lay, uint32_t *mask,
                                             unsigned long long *queue,
                                             unsigned long long *err) {
    uint8_t *S = scratch + (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * lay.stride;
    bool overflow = false;
    while (true) {
        const unsigned long long r = atomicAdd(queue, 1ull);
        if (r >= nreads)
            break;
        const uint64_t b = off[r];
        dust_read(seq + b, uint32_t(off[r + 1] - b), T, W, S, lay, mask, b, overflow);
    }
    if (overflow && err)
        atomicOr(err, 1ull);
}

} // namespace bbt

#endif
