/*
 * bbt_dust.h -- the -D pre-pass on gfx950: sdust low-complexity intervals of
 * every read of a batch, as a bitmap of window starts that the walks consult
 * (bit off[read] + p set: window p is treated as a miss without a
 * contains() call, SeqEval.h:53 and the same test in every scorer).
 *
 * The per-read pass is dust_core (bbt_dust_core.h, the steps of
 * Common/sdust.c:75-165).  Two kernels run it:
 *   k_dust_lds  (-W up to kDustFastW) one lane per read, the window state of
 *               each lane in LDS (word counts, the ring of words, up to
 *               kDustFastP perfect intervals); a read whose interval list
 *               needs more is cleared from the bitmap and listed
 *   k_dust_big  the listed reads, one per workgroup, the whole state of the
 *               read in that workgroup's LDS (the list sized for the
 *               longest it can get)
 *   k_dust      one thread per read (all reads, or the listed ones where the
 *               LDS cannot hold the list), the state in a per-thread
 *               device-memory region sized for the longest list
 */
#ifndef BBT_DUST_H
#define BBT_DUST_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bbt_dust_core.h"

namespace bbt {

/* largest -W: word counts are 16-bit, and the perfect-interval list of one
 * window grows as W^2 (below) */
constexpr int kDustMaxW = 4096;
/* threads of the device-memory pass at most (each owns one scratch region) */
constexpr uint32_t kDustThreads = 4096;
/* the device-memory pass's scratch for all threads at most (large W: fewer
 * threads, at least one) */
constexpr uint64_t kDustScratchBudget = uint64_t(2) << 30;
/* LDS of one k_dust_lds workgroup (64 lanes) */
constexpr uint32_t kDustFastLds = 4 * 64 * 64 + kDustFastP * 64 * 12;

/* Per-thread region of the device-memory pass: counts of the 64 words in
 * the window (cv, cw) and a work copy c, the window's words as a ring of
 * ring_cap >= W - 2 entries, the new entries of one perfect() call (W), then
 * the perfect-interval list (dust_pcap_bound(W) entries, what it can need). */
struct DustLayout {
    uint32_t ring_cap; /* a power of two */
    uint32_t pcap;
    uint64_t new_at;   /* byte offset of the new entries in a thread's region */
    uint64_t perf_at;  /* byte offset of the list */
    uint64_t stride;   /* bytes per thread */
    uint32_t threads;  /* threads of the pass (the grid is rounded up to 64) */
};

inline DustLayout dust_layout(int W) {
    DustLayout d;
    d.ring_cap = 64;
    while (int(d.ring_cap) < W - 2)
        d.ring_cap <<= 1;
    d.pcap = uint32_t(dust_pcap_bound(W));
    d.new_at = ((3 * 64 + uint64_t(d.ring_cap)) * 2 + 15) / 16 * 16;
    d.perf_at = d.new_at + uint64_t(W) * sizeof(DustPerf);
    d.stride = (d.perf_at + uint64_t(d.pcap) * sizeof(DustPerf) + 255) / 256 * 256;
    const uint64_t fit = kDustScratchBudget / d.stride;
    d.threads = uint32_t(fit < 1 ? 1 : (fit > kDustThreads ? kDustThreads : fit));
    return d;
}

/* View of one sequence's state in plain arrays: counts of type C, the ring,
 * a call's new entries and the list (a thread's device-memory region, or the
 * LDS of a workgroup). */
template <typename C>
struct DustFlatView {
    static constexpr bool kSepNew = true;
    C *cv_, *cw_, *c_, *ring_;
    DustPerf *N_, *P_;
    uint32_t rm, pc;
    __device__ DustFlatView(uint8_t *region, const DustLayout &lay) {
        cv_ = reinterpret_cast<C *>(region);
        cw_ = cv_ + 64;
        c_ = cw_ + 64;
        ring_ = c_ + 64;
        N_ = reinterpret_cast<DustPerf *>(region + lay.new_at);
        P_ = reinterpret_cast<DustPerf *>(region + lay.perf_at);
        rm = lay.ring_cap - 1;
        pc = lay.pcap;
    }
    __device__ C &cv(unsigned t) { return cv_[t]; }
    __device__ C &cw(unsigned t) { return cw_[t]; }
    __device__ C &c(unsigned t) { return c_[t]; }
    __device__ C &ring(uint32_t i) { return ring_[i]; }
    __device__ uint32_t rmask() const { return rm; }
    __device__ uint32_t pcap() const { return pc; }
    __device__ DustPerf P(uint32_t j) const { return P_[j]; }
    __device__ void setP(uint32_t j, const DustPerf &x) { P_[j] = x; }
    __device__ DustPerf N(uint32_t k) const { return N_[k]; }
    __device__ void setN(uint32_t k, const DustPerf &x) { N_[k] = x; }
};

/* View of lane ln's share of the LDS arrays of k_dust_lds: element i of a
 * lane's array at [i][ln], so the 64 lanes of an access touch 64 different
 * bytes of one 64-byte row when their indices agree.  The counts stay below
 * 2 (W - 2) < 256. */
struct DustLdsView {
    static constexpr bool kSepNew = false;
    uint8_t (*cnt)[64][64]; /* [0] cv, [1] cw, [2] c */
    uint8_t (*rg)[64];
    int32_t (*ps)[64];
    int32_t (*pf)[64];
    uint32_t (*prl)[64]; /* r << 16 | l (r < 2^16 for W <= 64) */
    uint32_t ln;
    __device__ uint8_t &cv(unsigned t) { return cnt[0][t][ln]; }
    __device__ uint8_t &cw(unsigned t) { return cnt[1][t][ln]; }
    __device__ uint8_t &c(unsigned t) { return cnt[2][t][ln]; }
    __device__ uint8_t &ring(uint32_t i) { return rg[i][ln]; }
    __device__ uint32_t rmask() const { return 63; }
    __device__ uint32_t pcap() const { return kDustFastP; }
    __device__ DustPerf P(uint32_t j) const {
        const uint32_t x = prl[j][ln];
        return DustPerf{ps[j][ln], pf[j][ln], int32_t(x >> 16), int32_t(x & 0xFFFFu)};
    }
    __device__ void setP(uint32_t j, const DustPerf &x) {
        ps[j][ln] = x.start;
        pf[j][ln] = x.finish;
        prl[j][ln] = (uint32_t(x.r) << 16) | uint32_t(x.l);
    }
};

/* Bytes b .. b + n - 1 of the batch through an 8-byte aligned register
 * window, the next word loaded one word ahead of its use: every load is an
 * aligned word holding at least one byte of the read, so it stays inside a
 * sequence buffer padded to a multiple of 8 bytes. */
struct DustWordSrc {
    const uint64_t *w; /* the batch as aligned words */
    uint64_t b, wi, last;
    uint64_t word, next;
    __device__ DustWordSrc(const uint8_t *seq, uint64_t b_, uint32_t n)
        : w(reinterpret_cast<const uint64_t *>(seq)), b(b_), wi(b_ >> 3),
          last((b_ + (n ? n : 1) - 1) >> 3) {
        word = n ? w[wi] : 0;
        next = wi < last ? w[wi + 1] : 0;
    }
    __device__ unsigned at(uint32_t i) {
        const uint64_t a = b + i;
        if ((a >> 3) != wi) {
            wi = a >> 3;
            word = next;
            next = wi < last ? w[wi + 1] : 0;
        }
        return unsigned(word >> ((a & 7) * 8)) & 0xFFu;
    }
};

/* Sets the bits of window starts [a, f) of the read at gbase. */
struct DustEmit {
    uint32_t *mask;
    uint64_t gbase;
    __device__ void operator()(int a, int f) const {
        for (int p = a; p < f;) {
            const uint64_t g = gbase + uint64_t(p);
            const uint32_t bit = uint32_t(g & 31);
            const int take = min(32 - int(bit), f - p);
            const uint32_t m = take == 32 ? 0xFFFFFFFFu : (((1u << take) - 1u) << bit);
            atomicOr(mask + (g >> 5), m);
            p += take;
        }
    }
};

/* Clears the bits of window starts [0, n) of the read at gbase (the other
 * reads' bits in the shared words are left as they are). */
__device__ inline void dust_clear(uint32_t *mask, uint64_t gbase, uint32_t n) {
    for (uint32_t p = 0; p < n;) {
        const uint64_t g = gbase + p;
        const uint32_t bit = uint32_t(g & 31);
        const uint32_t take = min(32u - bit, n - p);
        const uint32_t m = take == 32 ? 0xFFFFFFFFu : (((1u << take) - 1u) << bit);
        atomicAnd(mask + (g >> 5), ~m);
        p += take;
    }
}

/* One lane per read, 64 reads per queue claim, window state in LDS.  A read
 * whose list outgrows kDustFastP entries is cleared and appended to
 * fall[*nfall]. */
__global__ __launch_bounds__(64) void k_dust_lds(const uint8_t *seq, const uint64_t *off,
                                                 uint64_t nreads, int T, int W, uint32_t *mask,
                                                 unsigned long long *queue, uint32_t *fall,
                                                 unsigned long long *nfall) {
    __shared__ uint8_t cnt[3][64][64];
    __shared__ uint8_t rg[64][64];
    __shared__ int32_t ps[kDustFastP][64];
    __shared__ int32_t pf[kDustFastP][64];
    __shared__ uint32_t prl[kDustFastP][64];
    const uint32_t ln = threadIdx.x;
    DustLdsView v{cnt, rg, ps, pf, prl, ln};
    while (true) {
        unsigned long long base = 0;
        if (ln == 0)
            base = atomicAdd(queue, 64ull);
        base = __shfl(base, 0);
        if (base >= nreads)
            break;
        const uint64_t r = base + ln;
        if (r < nreads) {
            const uint64_t b = off[r];
            const uint32_t n = uint32_t(off[r + 1] - b);
            DustWordSrc src(seq, b, n);
            DustEmit em{mask, b};
            if (!dust_core(src, n, T, W, v, em)) {
                dust_clear(mask, b, n);
                fall[atomicAdd(nfall, 1ull)] = uint32_t(r);
            }
        }
    }
}

/* The reads k_dust_lds handed on, one per workgroup from a queue over the
 * *nlist reads of list, with the whole state of the read (list included,
 * laid out as lay, dust_layout of the window) in the workgroup's dynamic
 * LDS; lane 0 works.  Used when the device lets one workgroup hold
 * lay.stride bytes of LDS. */
__global__ __launch_bounds__(64) void k_dust_big(const uint8_t *seq, const uint64_t *off, int T,
                                                 int W, DustLayout lay, uint32_t *mask,
                                                 unsigned long long *queue, const uint32_t *list,
                                                 const unsigned long long *nlist,
                                                 unsigned long long *err) {
    extern __shared__ __align__(16) uint8_t dust_lds[];
    if (threadIdx.x != 0)
        return;
    DustFlatView<uint16_t> v(dust_lds, lay);
    const uint64_t total = *nlist;
    bool overflow = false;
    while (true) {
        const unsigned long long q = atomicAdd(queue, 1ull);
        if (q >= total)
            break;
        const uint64_t r = list[q];
        const uint64_t b = off[r];
        const uint32_t n = uint32_t(off[r + 1] - b);
        DustWordSrc src(seq, b, n);
        DustEmit em{mask, b};
        if (!dust_core(src, n, T, W, v, em))
            overflow = true; /* not reached: pcap is what the list can need */
    }
    if (overflow && err)
        atomicOr(err, 1ull);
}

/* One thread per read from a queue over every read (list null) or over the
 * *nlist reads of list; each thread owns lay.stride bytes of scratch. */
__global__ __launch_bounds__(64) void k_dust(const uint8_t *seq, const uint64_t *off,
                                             uint64_t nreads, int T, int W, uint8_t *scratch,
                                             DustLayout lay, uint32_t *mask,
                                             unsigned long long *queue, const uint32_t *list,
                                             const unsigned long long *nlist,
                                             unsigned long long *err) {
    const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= lay.threads)
        return;
    DustFlatView<uint16_t> v(scratch + tid * lay.stride, lay);
    const uint64_t total = list ? *nlist : nreads;
    bool overflow = false;
    while (true) {
        const unsigned long long q = atomicAdd(queue, 1ull);
        if (q >= total)
            break;
        const uint64_t r = list ? list[q] : q;
        const uint64_t b = off[r];
        const uint32_t n = uint32_t(off[r + 1] - b);
        DustWordSrc src(seq, b, n);
        DustEmit em{mask, b};
        if (!dust_core(src, n, T, W, v, em))
            overflow = true; /* not reached: pcap is what the list can need */
    }
    if (overflow && err)
        atomicOr(err, 1ull);
}

} // namespace bbt

#endif
