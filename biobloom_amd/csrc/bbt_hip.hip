/*
 * bbt_hip.hip -- gfx950 kernels and the C ABI of libbbt_hip.
 *
 * Kernels
 *   k_walk_plain       one lane per (read, filter) walk, one launch per batch
 *   k_walk_persistent  persistent waves pulling (read, filter) walks from a
 *                      device work queue; idle lanes are refilled per wave so
 *                      every lane keeps one filter load in flight
 *   k_ordered          ORDERED mode: one lane per record, filters in order
 *   k_classify         hit lists -> ResultsManager class + summary counters
 *   k_insert           filter construction: every k-mer of a sequence buffer
 *
 * Reference behaviour is restated in bbt_device.h (the walk) and below
 * (modes: BioBloomClassifier.cpp:1145-1237; classes: ResultsManager.hpp:
 * 40-189).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bbt_hip.h"
#include "bbt_device.h"
#include "bbt_dust.h"
#include "bbt_reads.h"
#include "bbt_host.h"

using namespace bbt;

/* ======================================================================= */
/* error reporting                                                          */

static bbt_status fail(bbt_status code, const std::string &msg) {
    return set_error(code, msg);
}

#define HIP_OK(expr)                                                            \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess)                                                   \
            return fail(BBT_ERR_HIP, std::string(#expr) + ": " +                \
                                         hipGetErrorString(e_));                \
    } while (0)

/* ======================================================================= */
/* kernels                                                                  */

namespace {

constexpr int kBlock = 256;
/* default length from which the batched path walks a read with a whole wave */
constexpr uint32_t kDefaultWaveMin = 256;
/* idle lanes of a wave that trigger a refill from the queue */
constexpr uint32_t kDefaultRefillMin = 32;
/* probe unroll: lookups of one contains() call issued together */
constexpr uint32_t kAhead = 4;
/* LDS of one CU (gfx950) */
constexpr uint32_t kLdsPerCu = 160 * 1024;

/* A (read, filter) item whose BINOMIAL evalScore the host computes: the
 * walk's hit count over its frame count (the read has more frames than the
 * score table); the host writes the score's bits back over frames/count. */
struct ScoreFix {
    uint64_t idx; /* read * nf + filter */
    uint32_t frames;
    uint32_t count;
};

struct EvalArgs {
    const uint8_t *seq;
    const uint64_t *off;
    const uint32_t *order;
    const FilterDev *filters;
    uint8_t *hits;
    double *scores;
    unsigned long long *counters;
    unsigned long long *queue;  /* lane-item queue */
    unsigned long long *lqueue; /* wave-item queue */
    const uint32_t *long_list;  /* reads walked by a whole wave, longest first */
    const uint32_t *short_list; /* reads walked one per lane */
    const unsigned long long *list_counts; /* [0] long reads, [1] short reads */
    /* length buckets of the wave list: bucket b's reads are at
     * long_list[bucket_base[b] ..) and number bucket_cnt[b] (b = 0..31) */
    const unsigned long long *bucket_base;
    const unsigned long long *bucket_cnt;
    const uint32_t *dust;       /* -D: bit (off[read] + p) = window p is low complexity */
    uint64_t nitems;
    uint32_t nf;  /* filters of the set: the stride of hits and scores */
    /* filters this launch walks: nw of them, walk index i is filter fmap[i]
     * (fmap null: all nf, in set order) */
    const uint32_t *fmap;
    uint32_t nw;
    uint32_t refill_min;
    uint32_t probe_ahead; /* lane walks: issue a contains() call's lookups together */
    /* lane items read by read (item = read * nf + filter): the lanes of one
     * read walk its filters side by side over interleaved bits */
    uint32_t read_major;
    /* BINOMIAL evalScore past the score table: the walk stores the hit count
     * and lists the item; the host turns the counts into scores */
    unsigned long long *fix_count;
    ScoreFix *fix_list;
    WalkParams wp;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ void add_counters(unsigned long long *ctr, uint32_t probes,
                                             uint32_t lookups) {
    if (!ctr)
        return;
    unsigned long long a = probes, b = lookups;
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    if (lane_id() == 0) {
        atomicAdd(ctr + 0, a);
        atomicAdd(ctr + 1, b);
    }
}

__device__ __forceinline__ void add_issued(unsigned long long *ctr, uint32_t issued) {
    if (!ctr)
        return;
    unsigned long long a = issued;
    for (int o = 32; o > 0; o >>= 1)
        a += __shfl_xor(a, o);
    if (lane_id() == 0)
        atomicAdd(ctr + 2, a);
}

template <int SC, bool EXH, int K>
__device__ __forceinline__ void store_walk(const EvalArgs &a, uint64_t read, uint32_t f,
                                           const Walker<SC, EXH, K> &w) {
    const uint64_t idx = read * a.nf + f;
    bool bad = w.bad != 0;
    if (EXH) {
        double v = w.score_value();
        if (SC == BINOMIAL) {
            /* evalBinomialScore then log10(.) * -10 (SeqEval.h:450-491, 558):
             * a read shorter than k has tail 1, i.e. -0 */
            const FilterDev &F = a.filters[f];
            const uint32_t kv = w.kk();
            if (w.len < kv) {
                v = -0.0;
            } else {
                const uint64_t fl = uint64_t(w.len) - kv + 1;
                if (F.bscore && fl <= F.bscore_frames) {
                    v = F.bscore[fl * (fl + 1) / 2 + w.score2];
                } else if (a.fix_list) {
                    const unsigned long long slot = atomicAdd(a.fix_count, 1ull);
                    ScoreFix fx;
                    fx.idx = idx;
                    fx.frames = uint32_t(fl);
                    fx.count = w.score2;
                    a.fix_list[slot] = fx;
                    v = double(w.score2);
                } else {
                    bad = true;
                }
            }
        }
        a.scores[idx] = v;
    } else {
        a.hits[idx] = w.accepted() ? 1 : 0;
    }
    if (bad && a.counters)
        atomicOr(a.counters + 3, 1ull);
}

/* Item -> (read, filter).  The batched path's lane walks are filter-major
 * (item = f * nr + r), so the walks in flight at any moment probe one or two
 * filters, whose lines then stay in each XCD's L2: random 128 B lines come
 * back about 1.5x faster from a 16 MiB working set than from hundreds of MiB.
 * The plain path keeps the input order, read-major (nr = 0). */
/* The roll terms of a launch whose filters share a compile-time k, copied
 * into LDS: the lane walks read one entry per base, and LDS keeps those reads
 * off the vector memory path that carries the filter loads.  With a runtime
 * k every walk reads its own filter's table (FilterDev::roll). */
template <int K>
__device__ __forceinline__ const uint64_t *roll_lds(const EvalArgs &a, uint64_t *s_roll) {
    if (K) {
        for (int t = threadIdx.x; t < int(2 * kRollEntries); t += blockDim.x)
            s_roll[t] = a.filters[0].roll[t];
    }
    return s_roll;
}

template <int SC, bool EXH, int K>
__device__ __forceinline__ void start_walk(const EvalArgs &a, uint64_t item, uint64_t nr,
                                           Walker<SC, EXH, K> &w, uint64_t &read,
                                           uint32_t &f, const uint64_t *s_roll) {
    uint64_t r;
    if (nr) {
        f = uint32_t(item / nr);
        r = item - uint64_t(f) * nr;
    } else {
        r = item / a.nw;
        f = uint32_t(item - r * a.nw);
    }
    if (a.fmap)
        f = a.fmap[f];
    read = a.short_list ? a.short_list[r] : (a.order ? a.order[r] : r);
    const uint64_t b = a.off[read];
    w.init(a.seq + b, uint32_t(a.off[read + 1] - b), a.filters + f, a.wp);
    if (K)
        w.rt = s_roll;
    w.dm = a.dust;
    w.dbase = b;
}

template <int SC, bool EXH, int K>
__global__ __launch_bounds__(kBlock) void k_walk_plain(EvalArgs a) {
    __shared__ uint64_t s_roll[2 * kRollEntries];
    const uint64_t *roll = roll_lds<K>(a, s_roll);
    __syncthreads();
    const uint64_t item = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    Walker<SC, EXH, K> w;
    w.probes = 0;
    w.lookups = 0;
    if (item < a.nitems) {
        uint64_t read;
        uint32_t f;
        start_walk(a, item, 0, w, read, f, roll);
        while (w.advance())
            w.consume(ld_global(w.addr));
        store_walk(a, read, f, w);
    }
    add_counters(a.counters, w.probes, w.lookups);
    add_issued(a.counters, w.lookups);
}

/* The windows of a read for a whole wavefront, 64 per step: lane L of the
 * chunk that starts at base b0 takes the window ending at base b0 + L, whose
 * hash comes from prefix-XOR scans across lanes (bbt_device.h, pfx_*) and
 * the prefixes carried from the previous chunk.  next() returns whether the
 * window is N-free and complete, with its start and canonical hash. */
struct ChunkHash {
    uint64_t Hc = 0, Gc = 0, Hp = 0, Gp = 0;
    int lastN = -1;

    __device__ __forceinline__ bool next(const uint8_t *s, uint32_t len, uint32_t b0,
                                         unsigned kv, uint32_t lane, uint32_t &pw_out,
                                         uint64_t &h0) {
        const uint32_t i = b0 + lane;
        const unsigned c = i < len ? unsigned(s[i]) : 0u;
        const bool ok = i < len && base_valid(c);
        const unsigned code = base_code(c);
        uint64_t H = ok ? pfx_fwd(code, i) : 0;
        uint64_t G = ok ? pfx_rc(code, i) : 0;
        int ln = ok ? -1 : int(i);
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t th = __shfl_up(H, unsigned(d));
            const uint64_t tg = __shfl_up(G, unsigned(d));
            const int tl = __shfl_up(ln, unsigned(d));
            if (int(lane) >= d) {
                H ^= th;
                G ^= tg;
                ln = ln > tl ? ln : tl;
            }
        }
        H ^= Hc;
        G ^= Gc;
        ln = ln > lastN ? ln : lastN;
        /* prefix just before the window: base i - k, here or in the previous chunk */
        const int src = int(lane) - int(kv);
        const uint64_t hb_cur = __shfl(H, src & 63), gb_cur = __shfl(G, src & 63);
        const uint64_t hb_prev = __shfl(Hp, (src + 64) & 63), gb_prev = __shfl(Gp, (src + 64) & 63);
        const uint64_t Hb = src >= 0 ? hb_cur : hb_prev;
        const uint64_t Gb = src >= 0 ? gb_cur : gb_prev;
        const int64_t pw = int64_t(i) - int64_t(kv) + 1;
        const bool valid = pw >= 0 && i < len && int64_t(ln) < pw;
        const uint64_t wf = pfx_win_fwd(H, Hb, i);
        const uint64_t wr = pfx_win_rc(G, Gb, uint32_t(pw));
        h0 = wr < wf ? wr : wf;
        Hc = __shfl(H, 63);
        Gc = __shfl(G, 63);
        lastN = __shfl(ln, 63);
        Hp = H;
        Gp = G;
        pw_out = uint32_t(pw);
        return valid;
    }
};

/*
 * One (read, filter) walk by a whole wavefront, 64 windows per step: the
 * window hashes come from prefix-XOR scans across lanes (ChunkHash), every
 * window the walk may probe is probed at once (each lane
 * stops at its first clear bit), and the walk is then replayed in window
 * order over the chunk's valid / hit masks with the same window_pre /
 * on_result steps as the lane walk.  The replay counts contains() calls and
 * lookups exactly as the reference would make them; lookups of windows the
 * walk skips or never reaches are counted separately as issued-only.
 */
template <int SC, bool EXH, int K>
__device__ void wave_walk(const EvalArgs &a, uint64_t read, uint32_t f,
                          Walker<SC, EXH, K> &w, uint32_t &issued) {
    const uint32_t lane = lane_id();
    const uint64_t b = a.off[read];
    const uint32_t len = uint32_t(a.off[read + 1] - b);
    const uint8_t *s = a.seq + b;
    w.init(s, len, a.filters + f, a.wp);
    w.dm = a.dust;
    w.dbase = b;
    w.probes = 0;
    w.lookups = 0;
    const unsigned kv = w.kk();
    ChunkHash ch;
    for (uint32_t b0 = 0; b0 < len && w.state == W_RUNNING; b0 += 64) {
        uint32_t pw;
        uint64_t h0;
        const bool valid = ch.next(s, len, b0, kv, lane, pw, h0);
        const uint64_t vmask = __ballot(valid);
        if (!vmask)
            continue;
        /* windows inside the current jump-skip are consumed without a probe,
         * and so are -D's low-complexity windows */
        const bool low = valid && w.low_complexity(pw);
        const uint64_t dmask = __ballot(low);
        bool need = valid && !low && !(w.skipping && pw < w.skipEnd);
        bool allset = need;
        uint32_t lk = 0;
        for (uint32_t jj = 0; __any(need); ++jj) {
            if (need) {
                const uint64_t hv = multi_hash(h0, jj, w.kseed());
                const uint64_t pos = fastmod(hv, w.m, w.magic, w.sh1, w.sh2);
                const unsigned byte = ld_global(w.bits + ((pos >> 3) << w.bshift));
                ++lk;
                if (!(byte & (1u << unsigned(pos & 7)))) {
                    allset = false;
                    need = false;
                } else if (jj + 1 >= w.h) {
                    need = false;
                }
            }
        }
        issued += lk;
        const uint64_t hmask = __ballot(allset);
        /* lookups of lanes [x, y) from an inclusive prefix sum across lanes */
        uint32_t lkp = lk;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(lkp, unsigned(d));
            if (int(lane) >= d)
                lkp += t;
        }
        const auto lk_range = [&](uint32_t x, uint32_t y) -> uint32_t {
            const uint32_t hi = uint32_t(__shfl(int(lkp), int(y - 1)));
            const uint32_t lo = x ? uint32_t(__shfl(int(lkp), int(x - 1))) : 0u;
            return hi - lo;
        };
        w.replay(b0, vmask, hmask, dmask, lk_range);
    }
    if (w.state == W_RUNNING)
        w.finish_end();
}

template <int SC, bool EXH, int K>
__global__ __launch_bounds__(kBlock) void k_walk_persistent(EvalArgs a) {
    const uint32_t lane = lane_id();
    const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
    Walker<SC, EXH, K> w;
    w.probes = 0;
    w.lookups = 0;
    uint32_t issued = 0;
    uint32_t wave_probes = 0, wave_lookups = 0; /* lane 0 only */
    const uint64_t nl = a.list_counts ? a.list_counts[0] : 0;
    const uint64_t nr = a.list_counts ? a.list_counts[1] : (a.nw ? a.nitems / a.nw : 0);
    const uint64_t nlong = nl * a.nw;
    const uint64_t nitems = nr * a.nw;

    /* phase 1: long reads, a whole wave per (read, filter): length bucket by
     * length bucket, longest first, and filter-major inside a bucket.  The
     * longest walks of every filter start first (none is left for the end of
     * a small launch), and the walks in flight probe one or two filters at a
     * time, as the lane walks do (uneven workload: 88.5 -> 85.4 ms against
     * read-major order) */
    __shared__ unsigned long long s_bbase[32], s_bcnt[32];
    __shared__ uint64_t s_roll[2 * kRollEntries];
    const uint64_t *roll = roll_lds<K>(a, s_roll);
    if (nlong) {
        for (int t = threadIdx.x; t < 32; t += blockDim.x) {
            s_bbase[t] = a.bucket_base[t];
            s_bcnt[t] = a.bucket_cnt[t];
        }
    }
    __syncthreads();
    while (nlong) {
        unsigned long long it = 0, fr = 0;
        if (lane == 0) {
            it = atomicAdd(a.lqueue, 1ull);
            if (it < nlong) {
                /* the bucket whose item range [base nf, (base + cnt) nf) holds it */
                for (int b = 31; b >= 0; --b) {
                    const unsigned long long c = s_bcnt[b];
                    const unsigned long long lo = s_bbase[b] * a.nw;
                    if (c && it < lo + c * a.nw) {
                        const unsigned long long q = it - lo;
                        if (a.read_major) {
                            /* the filters of one read side by side */
                            const unsigned long long rb = q / a.nw;
                            fr = ((q - rb * a.nw) << 40) | (s_bbase[b] + rb);
                        } else {
                            const unsigned long long fb = q / c;
                            fr = (fb << 40) | (s_bbase[b] + (q - fb * c));
                        }
                        break;
                    }
                }
            }
        }
        it = __shfl(it, 0);
        if (it >= nlong)
            break;
        fr = __shfl(fr, 0);
        const uint32_t f = a.fmap ? a.fmap[fr >> 40] : uint32_t(fr >> 40);
        const uint64_t r = fr & ((uint64_t(1) << 40) - 1);
        const uint64_t read = a.long_list[r];
        Walker<SC, EXH, K> ww;
        wave_walk(a, read, f, ww, issued);
        if (lane == 0) {
            store_walk(a, read, f, ww);
            wave_probes += ww.probes;
            wave_lookups += ww.lookups;
        }
    }

    /* phase 2: one walk per lane, idle lanes refilled from the queue */
    bool active = false;
    bool drained = false;
    uint64_t read = 0;
    uint32_t f = 0;
    uint32_t ahead_extra = 0; /* probe unroll: lookups issued past a call's end */
    /* probe_ahead bit 0: a call's bits kAhead at a time; bit 1: lookahead */
    const uint32_t na = (a.probe_ahead & 1u) ? kAhead : 1u;
    const bool look = (a.probe_ahead & 2u) != 0;
    unsigned pbyte[kAhead] = {0, 0, 0, 0}; /* lookahead: the pending call's loaded bytes */
    uint32_t psh = 0, pn = 0;              /* their bit positions (3 bits each), count */
    while (true) {
        const uint64_t idle = __ballot(!active);
        const uint32_t nidle = uint32_t(__popcll(idle));
        if (!drained && nidle != 0 && (nidle >= a.refill_min || nidle == 64)) {
            const uint32_t leader = uint32_t(__ffsll((unsigned long long)idle) - 1);
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(a.queue, (unsigned long long)nidle);
            base = __shfl(base, int(leader));
            if (base + nidle >= nitems)
                drained = true;
            if (!active) {
                const uint64_t item = base + uint64_t(__popcll(idle & lt_mask));
                if (item < nitems) {
                    start_walk(a, item, a.read_major ? 0 : nr, w, read, f, roll);
                    active = true;
                }
            }
        }
        if (!__any(active)) {
            if (drained)
                break;
            continue;
        }
        if (active) {
            const uint64_t ks = w.kseed();
            /* the rest of a pending contains() call: na bits per round (probe
             * unroll: kAhead independent loads), taken in hash order until the
             * call ends at a clear bit or the last hash; loads past that end
             * are issued-only */
            const auto finish_call = [&]() {
                while (w.pending) {
                    const uint32_t j0 = w.j;
                    bool bit[kAhead];
                    uint32_t n = 0;
#pragma unroll
                    for (uint32_t jj = 0; jj < kAhead; ++jj) {
                        bit[jj] = false;
                        if (jj < na && j0 + jj < w.h) {
                            const uint64_t pos = fastmod(multi_hash(w.h0, j0 + jj, ks), w.m,
                                                         w.magic, w.sh1, w.sh2);
                            bit[jj] = (ld_global(w.bits + ((pos >> 3) << w.bshift)) >> unsigned(pos & 7)) & 1u;
                            ++n;
                        }
                    }
                    const uint32_t before = w.lookups;
#pragma unroll
                    for (uint32_t jj = 0; jj < kAhead; ++jj)
                        if (jj < n && w.pending)
                            w.take(bit[jj]);
                    ahead_extra += n - (w.lookups - before);
                }
            };
            if (look) {
                /* lookahead: roll the next base and issue its window's first
                 * bits, then take the pending call's bits (issued a pass
                 * earlier, so their loads overlapped this pass's work), then
                 * decide the new window on the state they leave; its loads
                 * are issued-only if it is not probed */
                uint32_t cp = 0;
                uint64_t ch = 0;
                const bool cand = w.state == W_RUNNING && w.i < w.len && w.roll_step(cp, ch);
                unsigned cbyte[kAhead];
                uint32_t csh = 0, cn = 0;
#pragma unroll
                for (uint32_t jj = 0; jj < kAhead; ++jj) {
                    cbyte[jj] = 0;
                    if (cand && jj < na && jj < w.h) {
                        const uint64_t pos =
                            fastmod(multi_hash(ch, jj, ks), w.m, w.magic, w.sh1, w.sh2);
                        cbyte[jj] = ld_global(w.bits + ((pos >> 3) << w.bshift));
                        csh |= uint32_t(pos & 7) << (3 * jj);
                        ++cn;
                    }
                }
                if (w.pending) {
                    const uint32_t before = w.lookups;
#pragma unroll
                    for (uint32_t jj = 0; jj < kAhead; ++jj)
                        if (jj < pn && w.pending)
                            w.take((pbyte[jj] >> ((psh >> (3 * jj)) & 7u)) & 1u);
                    ahead_extra += pn - (w.lookups - before);
                    finish_call();
                }
                pn = 0;
                if (cand && w.state == W_RUNNING && w.decide(cp, ch)) {
#pragma unroll
                    for (uint32_t jj = 0; jj < kAhead; ++jj)
                        pbyte[jj] = cbyte[jj];
                    psh = csh;
                    pn = cn;
                } else {
                    ahead_extra += cn;
                }
                if (w.state == W_RUNNING && !w.pending && w.i >= w.len)
                    w.finish_end();
            } else if (w.step()) {
                /* one base per lane per pass */
                finish_call();
            }
            if (w.state != W_RUNNING) {
                store_walk(a, read, f, w);
                active = false;
            }
        }
    }
    issued += w.lookups + ahead_extra;
    add_counters(a.counters, w.probes + wave_probes, w.lookups + wave_lookups);
    add_issued(a.counters, issued);
}

/* Lookup counts of four filters packed 16 bits each: one more for every
 * filter whose bit is set in the low four bits of need. */
__device__ __forceinline__ uint64_t spread16(uint32_t need) {
    return uint64_t(need & 1u) | (uint64_t(need & 2u) << 15) | (uint64_t(need & 4u) << 30) |
           (uint64_t(need & 8u) << 45);
}

/*
 * One read against the 3 to 2^SH filters of an interleaved group (2^SH = 4
 * or 8 bytes per bit position; walk index g is filter fmap[g], a.nw of
 * them), by a whole wavefront, SIMPLE evalRead: the window hashes are
 * computed once (ChunkHash), each lane's
 * window is probed for every filter whose walk may probe it with one load
 * per hash function (the group's bytes; a filter stops at its first clear
 * bit, the lane when no filter needs the next one), and each filter's walk is
 * then replayed over its own hit mask (Walker::replay), so its contains()
 * calls and lookups are the reference's.  Only the walkers' automaton fields
 * are used.  Lane 0 stores the hits and returns the walks' probes and
 * lookups; issued counts per filter every bit loaded for it.
 */
template <int K, unsigned SH>
__device__ void wave_group_walk(const EvalArgs &a, uint64_t read, const uint8_t *ilv,
                                uint32_t &issued, uint32_t &probes, uint32_t &lookups) {
    constexpr uint32_t G = 1u << SH;
    const uint32_t lane = lane_id();
    const uint32_t nf = a.nw;
    const uint64_t b = a.off[read];
    const uint32_t len = uint32_t(a.off[read + 1] - b);
    const uint8_t *s = a.seq + b;
    const auto fidx = [&](uint32_t g) -> uint32_t { return a.fmap ? a.fmap[g] : g; };
    Walker<SIMPLE, false, K> w[G];
#pragma unroll
    for (uint32_t g = 0; g < G; ++g) {
        w[g].init(s, len, a.filters + fidx(g < nf ? g : 0), a.wp);
        w[g].dm = a.dust;
        w[g].dbase = b;
        w[g].probes = 0;
        w[g].lookups = 0;
        if (g >= nf)
            w[g].state = W_REJECT;
    }
    const FilterDev *F0 = a.filters + fidx(0);
    const uint64_t m = F0->m, magic = F0->magic;
    const uint32_t sh1 = F0->sh1, sh2 = F0->sh2, h = F0->h;
    const unsigned kv = w[0].kk();
    const uint64_t ks = w[0].kseed();
    const auto running = [&]() {
        bool r = false;
#pragma unroll
        for (uint32_t g = 0; g < G; ++g)
            r = r || w[g].state == W_RUNNING;
        return r;
    };
    ChunkHash ch;
    uint32_t np = 0, nl = 0;
    for (uint32_t b0 = 0; b0 < len && running(); b0 += 64) {
        uint32_t pw;
        uint64_t h0;
        const bool valid = ch.next(s, len, b0, kv, lane, pw, h0);
        const uint64_t vmask = __ballot(valid);
        if (!vmask)
            continue;
        /* -D's low-complexity windows are misses without a probe for all */
        const bool low = valid && w[0].low_complexity(pw);
        const uint64_t dmask = __ballot(low);
        uint32_t need = 0;
#pragma unroll
        for (uint32_t g = 0; g < G; ++g)
            if (valid && !low && w[g].state == W_RUNNING && !(w[g].skipping && pw < w[g].skipEnd))
                need |= 1u << g;
        uint32_t all = need;
        uint64_t lk = 0, lk2 = 0; /* filters 0-3 and 4-7, 16 bits each */
        /* takes the group's bits of hash jj for the filters in need */
        const auto take = [&](const uint32_t *word, unsigned b, uint32_t jj) {
            const uint32_t bit = group_bits<SH>(word, b);
            lk += spread16(need);
            if (SH == 3)
                lk2 += spread16(need >> 4);
            all &= bit | ~need;
            need &= bit;
            if (jj + 1 >= h)
                need = 0;
        };
        uint32_t jj = 0;
        if (SH == 3) {
            /* a group of 8: some filter almost always needs the second and
             * third hash of a window it probes, so the first kSpec loads are
             * issued together (one load latency instead of kSpec in a row) */
            constexpr uint32_t kSpec = 3;
            const uint32_t ns = h < kSpec ? h : kSpec;
            if (need) {
                uint32_t wv[kSpec][2];
                uint32_t bsh = 0;
#pragma unroll
                for (uint32_t q = 0; q < kSpec; ++q) {
                    if (q < ns) {
                        const uint64_t pos = fastmod(multi_hash(h0, q, ks), m, magic, sh1, sh2);
                        const uint2 v =
                            ld_global(reinterpret_cast<const uint2 *>(ilv + ((pos >> 3) << 3)));
                        wv[q][0] = v.x;
                        wv[q][1] = v.y;
                        bsh |= uint32_t(pos & 7) << (3 * q);
                    }
                }
                issued += uint32_t(__popc(need)) * ns;
#pragma unroll
                for (uint32_t q = 0; q < kSpec; ++q)
                    if (q < ns && need)
                        take(wv[q], (bsh >> (3 * q)) & 7u, q);
            }
            jj = ns;
        }
        for (; __any(need != 0); ++jj) {
            if (need) {
                const uint64_t pos = fastmod(multi_hash(h0, jj, ks), m, magic, sh1, sh2);
                uint32_t word[2];
                if (SH == 3) {
                    const uint2 v = ld_global(reinterpret_cast<const uint2 *>(ilv + ((pos >> 3) << 3)));
                    word[0] = v.x;
                    word[1] = v.y;
                } else {
                    word[0] = ld_global(reinterpret_cast<const uint32_t *>(ilv + ((pos >> 3) << 2)));
                }
                issued += uint32_t(__popc(need));
                take(word, unsigned(pos & 7), jj);
            }
        }
        /* per-filter lookups of lanes [x, y) from packed prefix sums */
        uint64_t lkp = lk, lkp2 = lk2;
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t t = __shfl_up(lkp, unsigned(d));
            const uint64_t t2 = SH == 3 ? __shfl_up(lkp2, unsigned(d)) : uint64_t(0);
            if (int(lane) >= d) {
                lkp += t;
                lkp2 += t2;
            }
        }
#pragma unroll
        for (uint32_t g = 0; g < G; ++g) {
            const uint64_t hmask = __ballot((all >> g) & 1u);
            if (w[g].state != W_RUNNING)
                continue;
            const auto lk_range = [&](uint32_t x, uint32_t y) -> uint32_t {
                const uint64_t src = g < 4 ? lkp : lkp2;
                const uint64_t hi = __shfl(src, int(y - 1));
                const uint64_t lo = x ? __shfl(src, int(x - 1)) : uint64_t(0);
                return uint32_t(((hi - lo) >> (16 * (g & 3))) & 0xFFFFu);
            };
            w[g].replay(b0, vmask, hmask, dmask, lk_range);
            /* the walks' counts are summed chunk by chunk, which keeps them
             * out of the state carried from one chunk to the next */
            np += w[g].probes;
            nl += w[g].lookups;
            w[g].probes = 0;
            w[g].lookups = 0;
        }
    }
#pragma unroll
    for (uint32_t g = 0; g < G; ++g) {
        if (g >= nf)
            continue;
        if (w[g].state == W_RUNNING)
            w[g].finish_end();
        if (lane == 0)
            store_walk(a, read, fidx(g), w[g]);
    }
    if (lane == 0) {
        probes += np;
        lookups += nl;
    }
}

/* The long reads of a batch with length buckets (a.long_list, longest bucket
 * first) against a group of 3 to 2^SH filters, a whole wave per read
 * (wave_group_walk); the short reads are k_walk_rows's. */
template <int K, unsigned SH>
__global__ __launch_bounds__(kBlock, 3) void k_walk_group_long(EvalArgs a, const uint8_t *ilv) {
    const uint32_t lane = lane_id();
    uint32_t probes = 0, lookups = 0, issued = 0;
    const uint64_t nl = a.list_counts[0];
    while (true) {
        unsigned long long r = 0;
        if (lane == 0)
            r = atomicAdd(a.lqueue, 1ull);
        r = __shfl(r, 0);
        if (r >= nl)
            break;
        wave_group_walk<K, SH>(a, a.long_list[r], ilv, issued, probes, lookups);
    }
    add_counters(a.counters, probes, lookups);
    add_issued(a.counters, issued);
}

/*
 * One read per lane against every filter of a set interleaved as one group,
 * SIMPLE evalRead (RowsWalk, bbt_device.h): the read's windows are hashed
 * once, each bit position is one load of the group's 2^SH bytes (16 for 9
 * to 16 filters), and each filter keeps its own automaton in registers.
 * Lanes take reads from the queue (with length buckets, the short list);
 * a read's hits are written for all its filters when the last of its walks
 * ends.  The group's filters are a.fmap[0 .. a.nw) (null: the set's first
 * a.nw); a.filters: the filters' own descriptors (m, h, k, roll table;
 * their bits are not read).
 */
template <int K, unsigned SH>
__global__ __launch_bounds__(kBlock) void k_walk_rows(EvalArgs a, const uint8_t *ilv) {
    __shared__ uint64_t s_roll[2 * kRollEntries];
    const uint64_t *roll = roll_lds<K>(a, s_roll);
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
    const FilterDev *F0 = a.filters + (a.fmap ? a.fmap[0] : 0u);
    const uint64_t m = F0->m, magic = F0->magic;
    const uint32_t sh1 = F0->sh1, sh2 = F0->sh2;
    const uint64_t ks = uint64_t(K ? K : F0->k) * kMultiSeed;
    const uint32_t nfmask = a.nw >= 32 ? ~0u : (1u << a.nw) - 1u;
    const auto fetch = [&](uint64_t hv, uint32_t j) -> uint32_t {
        const uint64_t pos = fastmod(multi_hash(hv, j, ks), m, magic, sh1, sh2);
        const uint8_t *g = ilv + ((pos >> 3) << SH);
        uint32_t w[4];
        if (SH == 4) {
            const uint4 v = ld_global(reinterpret_cast<const uint4 *>(g));
            w[0] = v.x;
            w[1] = v.y;
            w[2] = v.z;
            w[3] = v.w;
        } else if (SH == 3) {
            const uint2 v = ld_global(reinterpret_cast<const uint2 *>(g));
            w[0] = v.x;
            w[1] = v.y;
        } else {
            w[0] = ld_global(reinterpret_cast<const uint32_t *>(g));
        }
        return group_bits<SH>(w, unsigned(pos & 7)) & nfmask;
    };
    RowsWalk<K, 16> rw;
    uint32_t probes = 0, lookups = 0, issued = 0;
    const uint64_t nreads = a.list_counts ? a.list_counts[1] : a.nitems / a.nw;
    bool active = false, drained = false;
    uint64_t read = 0;
    while (true) {
        const uint64_t idle = __ballot(!active);
        const uint32_t nidle = uint32_t(__popcll(idle));
        if (!drained && nidle != 0 && (nidle >= a.refill_min || nidle == 64)) {
            const uint32_t leader = uint32_t(__ffsll((unsigned long long)idle) - 1);
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(a.queue, (unsigned long long)nidle);
            base = __shfl(base, int(leader));
            if (base + nidle >= nreads)
                drained = true;
            if (!active) {
                const uint64_t r = base + uint64_t(__popcll(idle & lt_mask));
                if (r < nreads) {
                    read = a.short_list ? a.short_list[r] : (a.order ? a.order[r] : r);
                    const uint64_t b = a.off[read];
                    rw.init(a.seq + b, uint32_t(a.off[read + 1] - b), F0, a.nw, a.wp);
                    if (K)
                        rw.w.rt = roll;
                    rw.w.dm = a.dust;
                    rw.w.dbase = b;
                    active = true;
                }
            }
        }
        if (!__any(active)) {
            if (drained)
                break;
            continue;
        }
        if (active && !rw.step(fetch)) {
            for (uint32_t g = 0; g < a.nw; ++g)
                a.hits[read * a.nf + (a.fmap ? a.fmap[g] : g)] = uint8_t((rw.acc >> g) & 1u);
            probes += rw.probes;
            lookups += rw.lookups;
            issued += rw.lookups;
            active = false;
        }
    }
    add_counters(a.counters, probes, lookups);
    add_issued(a.counters, issued);
}

/* Length buckets for the batched path: reads of at least wave_min bases go
 * to the wave list by floor(log2(len)), longest bucket first; the rest to
 * the lane list.  cnt[0..31] bucket sizes, cnt[32] short reads. */
__global__ __launch_bounds__(kBlock) void k_len_count(const uint64_t *off, uint64_t nreads,
                                                      uint32_t wave_min,
                                                      unsigned long long *cnt) {
    __shared__ unsigned int h[33];
    for (int t = threadIdx.x; t < 33; t += blockDim.x)
        h[t] = 0;
    __syncthreads();
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nreads;
         r += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t len = off[r + 1] - off[r];
        if (len >= wave_min && len > 0)
            atomicAdd(&h[min(31, 63 - __clzll((long long)len))], 1u);
        else
            atomicAdd(&h[32], 1u);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 33; t += blockDim.x)
        if (h[t])
            atomicAdd(cnt + t, (unsigned long long)h[t]);
}

__global__ void k_len_scan(unsigned long long *cnt, unsigned long long *base,
                           unsigned long long *list_counts) {
    if (threadIdx.x != 0)
        return;
    unsigned long long pos = 0;
    for (int b = 31; b >= 0; --b) {
        base[b] = pos;
        pos += cnt[b];
    }
    list_counts[0] = pos;
    list_counts[1] = cnt[32];
    base[32] = 0;
    for (int b = 0; b < 33; ++b)
        cnt[b] = 0; /* reused as scatter cursors */
}

/* Claims, for every lane with a bucket b >= 0, the next free slot of that
 * bucket's counter in cnt (LDS), one atomic per distinct bucket in the wave;
 * returns the lane's slot. */
__device__ __forceinline__ uint32_t wave_claim(int b, unsigned int *cnt) {
    const uint32_t lane = lane_id();
    const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
    uint64_t pending = __ballot(b >= 0);
    uint32_t mine = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int bl = __shfl(b, leader);
        const uint64_t m = __ballot(b == bl);
        unsigned int base = 0;
        if (int(lane) == leader)
            base = atomicAdd(cnt + bl, unsigned(__popcll(m)));
        base = __shfl(base, leader);
        if (b == bl)
            mine = base + uint32_t(__popcll(m & lt_mask));
        pending &= ~m;
    }
    return mine;
}

/* Each workgroup takes a contiguous share of the reads: it counts its reads
 * per bucket in LDS, reserves that many list slots per bucket with one
 * global atomic each, then writes every read into its bucket's slots (the
 * order inside a bucket is not defined).  Buckets 0..31 are the wave list by
 * floor(log2 len) at base[b]; 32 is the lane list. */
__global__ __launch_bounds__(kBlock) void k_len_scatter(const uint64_t *off, uint64_t nreads,
                                                        uint32_t wave_min,
                                                        const unsigned long long *base,
                                                        unsigned long long *cursor,
                                                        uint32_t *long_list,
                                                        uint32_t *short_list) {
    __shared__ unsigned int cnt[33];
    __shared__ unsigned long long at[33];
    const uint64_t share = (nreads + gridDim.x - 1) / gridDim.x;
    const uint64_t r_begin = uint64_t(blockIdx.x) * share;
    const uint64_t r_end = r_begin + share < nreads ? r_begin + share : nreads;
    const auto bucket = [&](uint64_t r) -> int {
        if (r >= r_end)
            return -1;
        const uint64_t len = off[r + 1] - off[r];
        return len >= wave_min && len > 0 ? min(31, 63 - __clzll((long long)len)) : 32;
    };
    for (int t = threadIdx.x; t < 33; t += blockDim.x)
        cnt[t] = 0;
    __syncthreads();
    for (uint64_t r0 = r_begin; r0 < r_end; r0 += blockDim.x)
        (void)wave_claim(bucket(r0 + threadIdx.x), cnt);
    __syncthreads();
    for (int t = threadIdx.x; t < 33; t += blockDim.x) {
        at[t] = cnt[t] ? atomicAdd(cursor + t, (unsigned long long)cnt[t]) : 0;
        cnt[t] = 0;
    }
    __syncthreads();
    for (uint64_t r0 = r_begin; r0 < r_end; r0 += blockDim.x) {
        const uint64_t r = r0 + threadIdx.x;
        const int b = bucket(r);
        const uint32_t slot = wave_claim(b, cnt);
        if (b == 32)
            short_list[at[32] + slot] = uint32_t(r);
        else if (b >= 0)
            long_list[base[b] + at[b] + slot] = uint32_t(r);
    }
}

template <int SC, int K>
__device__ bool walk_one(const EvalArgs &a, uint64_t read, uint32_t f,
                         Walker<SC, false, K> &w, const uint64_t *s_roll) {
    const uint64_t b = a.off[read];
    w.init(a.seq + b, uint32_t(a.off[read + 1] - b), a.filters + f, a.wp);
    if (K)
        w.rt = s_roll;
    w.dm = a.dust;
    w.dbase = b;
    /* flat form: one base per pass, a contains() call's bits one at a time */
    const uint64_t ks = w.kseed();
    while (w.state == W_RUNNING) {
        if (!w.step())
            continue;
        while (w.pending) {
            const uint64_t pos = fastmod(multi_hash(w.h0, w.j, ks), w.m, w.magic, w.sh1, w.sh2);
            w.take((ld_global(w.bits + ((pos >> 3) << w.bshift)) >> unsigned(pos & 7)) & 1u);
        }
    }
    return w.accepted();
}

/* evaluateReadOrdered / evaluateReadOrderedPair (BioBloomClassifier.cpp:
 * 1145-1180): the first filter that passes, with the same short-circuit. */
template <int SC, int K>
__global__ __launch_bounds__(kBlock) void k_ordered(EvalArgs a, uint64_t nrec,
                                                    uint32_t paired, uint32_t inclusive) {
    __shared__ uint64_t s_roll[2 * kRollEntries];
    const uint64_t *roll = roll_lds<K>(a, s_roll);
    __syncthreads();
    const uint64_t rec = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    Walker<SC, false, K> w;
    w.probes = 0;
    w.lookups = 0;
    if (rec < nrec) {
        const uint64_t r1 = paired ? 2 * rec : rec;
        for (uint32_t f = 0; f < a.nf; ++f) {
            bool pass;
            if (!paired)
                pass = walk_one(a, r1, f, w, roll);
            else if (inclusive)
                pass = walk_one(a, r1, f, w, roll) || walk_one(a, r1 + 1, f, w, roll);
            else
                pass = walk_one(a, r1, f, w, roll) && walk_one(a, r1 + 1, f, w, roll);
            if (pass) {
                a.hits[r1 * a.nf + f] = 1;
                if (paired)
                    a.hits[(r1 + 1) * a.nf + f] = 1;
                break;
            }
        }
    }
    add_counters(a.counters, w.probes, w.lookups);
    add_issued(a.counters, w.lookups);
}

__global__ void k_fill_u8(uint8_t *p, uint64_t n, uint8_t v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

/* the host-computed scores of the listed items (value bits in place of
 * frames/count) into their rows */
__global__ void k_fix_scores(const ScoreFix *fix, uint64_t n, double *scores) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const ScoreFix f = fix[i];
        const uint64_t bits = uint64_t(f.frames) | (uint64_t(f.count) << 32);
        scores[f.idx] = __longlong_as_double((long long)bits);
    }
}

/* The members of a group of filters of one geometry: filter set indices,
 * slot q = the q-th member. */
struct IlvMembers {
    uint32_t n;
    uint8_t f[64];
};

/* A group interleaved byte by byte: byte i of its q-th member at
 * dst[(i << sh) + q], the unused slots of each 2^sh group zero.  The
 * descriptors give the filters' own arrays. */
__global__ __launch_bounds__(kBlock) void k_interleave(const FilterDev *src, IlvMembers mem,
                                                       uint64_t nbytes, uint32_t sh,
                                                       uint8_t *dst) {
    const uint32_t S = 1u << sh;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nbytes;
         i += uint64_t(gridDim.x) * blockDim.x) {
        uint8_t *o = dst + (i << sh);
        if (S < 4) {
            for (uint32_t q = 0; q < S; ++q)
                o[q] = q < mem.n ? src[mem.f[q]].bits[i] : 0;
            continue;
        }
        for (uint32_t q0 = 0; q0 < S; q0 += 4) {
            uint32_t w = 0;
            for (uint32_t q = 0; q < 4; ++q)
                if (q0 + q < mem.n)
                    w |= uint32_t(src[mem.f[q0 + q]].bits[i]) << (8 * q);
            *reinterpret_cast<uint32_t *>(o + q0) = w;
        }
    }
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        p[i] = 0;
}

/* evaluateReadBestHit (BioBloomClassifier.cpp:1194-1218): the filters with
 * the maximal evalScore, kept only when that maximum is > 0. */
__device__ void best_hits(const double *srow, uint8_t *hrow, uint32_t nf) {
    double mx = 0;
    for (uint32_t i = 0; i < nf; ++i) {
        const double s = srow[i];
        if (mx < s)
            mx = s;
    }
    for (uint32_t i = 0; i < nf; ++i)
        hrow[i] = (mx > 0 && srow[i] == mx) ? 1 : 0;
}

/* ResultsManager::updateSummaryData (ResultsManager.hpp:66-189). */
__global__ __launch_bounds__(kBlock) void k_classify(uint8_t *hits, const double *scores,
                                                     uint32_t nf, uint64_t nrec,
                                                     uint32_t paired, uint32_t inclusive,
                                                     uint32_t besthit, uint32_t *cls,
                                                     unsigned long long *summary) {
    extern __shared__ unsigned int cnt[]; /* 2*nf + 2 */
    const uint32_t ncnt = 2 * nf + 2;
    if (summary)
        for (uint32_t i = threadIdx.x; i < ncnt; i += blockDim.x)
            cnt[i] = 0;
    __syncthreads();
    const uint64_t rec = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (rec < nrec) {
        const uint64_t r1 = paired ? 2 * rec : rec;
        uint8_t *h1 = hits + r1 * nf;
        uint8_t *h2 = paired ? h1 + nf : nullptr;
        if (besthit) {
            best_hits(scores + r1 * nf, h1, nf);
            if (paired)
                best_hits(scores + (r1 + 1) * nf, h2, nf);
        }
        uint32_t c = nf;
        for (uint32_t i = 0; i < nf; ++i) {
            bool in = paired ? (inclusive ? (h1[i] | h2[i]) != 0 : (h1[i] & h2[i]) != 0)
                             : h1[i] != 0;
            if (in) {
                c = (c == nf) ? i : nf + 1;
                if (summary)
                    atomicAdd(&cnt[i], 1u);
            }
        }
        if (cls)
            cls[rec] = c;
        if (summary) {
            if (c == nf)
                atomicAdd(&cnt[2 * nf + 1], 1u);
            else if (c == nf + 1)
                atomicAdd(&cnt[2 * nf], 1u);
            else
                atomicAdd(&cnt[nf + c], 1u);
        }
    }
    __syncthreads();
    if (summary)
        for (uint32_t i = threadIdx.x; i < ncnt; i += blockDim.x)
            if (cnt[i])
                atomicAdd(summary + i, (unsigned long long)cnt[i]);
}

/* BloomFilterGenerator::loadFilter (BloomFilterGenerator.h:143-188): insert
 * every N-free k-mer.  Thread t owns window starts [t*chunk, (t+1)*chunk).
 * SUB: loadFilterSubtract (BloomFilterGenerator.h:217-284), a k-mer that the
 * subtract filter contains is counted as removed and not inserted. */
template <int K, bool SUB>
__global__ __launch_bounds__(kBlock) void k_insert(const uint8_t *seq, uint64_t nbytes,
                                                   FilterDev f, FilterDev sub, uint32_t chunk,
                                                   unsigned long long *counts) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t p0 = t * chunk;
    uint32_t nnew = 0, nold = 0, nsub = 0;
    if (p0 < nbytes) {
        const unsigned kv = K ? unsigned(K) : f.k;
        const uint64_t last = std::min<uint64_t>(p0 + chunk + kv - 1, nbytes);
        unsigned int *words = reinterpret_cast<unsigned int *>(const_cast<uint8_t *>(f.bits));
        uint64_t fwd = 0, rc = 0, win = 0;
        uint32_t run = 0;
        for (uint64_t i = p0; i < last; ++i) {
            const unsigned c = seq[i];
            if (!base_valid(c)) {
                run = 0;
                fwd = rc = 0;
                continue;
            }
            const unsigned code = base_code(c);
            if (run >= kv) {
                const unsigned out = unsigned(win >> (2 * (kv - 1))) & 3u;
                fwd = rotl1(fwd) ^ rotl64(seed_fwd(out), kv) ^ seed_fwd(code);
                rc = rotr1(rc ^ seed_rc(out) ^ rotl64(seed_rc(code), kv));
            } else {
                fwd = rotl1(fwd) ^ seed_fwd(code);
                rc = rotr1(rc ^ rotl64(seed_rc(code), kv));
                ++run;
            }
            win = (win << 2) | code;
            if (run < kv)
                continue;
            const uint64_t h0 = rc < fwd ? rc : fwd;
            if (SUB) {
                /* BloomFilter::contains on the subtract filter (same h and k) */
                bool in = true;
                for (uint32_t j = 0; j < sub.h && in; ++j) {
                    const uint64_t pos = fastmod(multi_hash(h0, j, uint64_t(kv) * kMultiSeed),
                                                 sub.m, sub.magic, sub.sh1, sub.sh2);
                    in = (sub.bits[pos >> 3] >> (pos & 7)) & 1u;
                }
                if (in) {
                    ++nsub;
                    continue;
                }
            }
            bool found = true;
            for (uint32_t j = 0; j < f.h; ++j) {
                const uint64_t hv = multi_hash(h0, j, uint64_t(kv) * kMultiSeed);
                const uint64_t pos = fastmod(hv, f.m, f.magic, f.sh1, f.sh2);
                const unsigned bit = 1u << unsigned(pos & 31);
                const unsigned old = atomicOr(words + (pos >> 5), bit);
                found = found && (old & bit);
            }
            if (found)
                ++nold;
            else
                ++nnew;
        }
    }
    unsigned long long a = nnew, b = nold, c = nsub;
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        if (SUB)
            c += __shfl_xor(c, o);
    }
    if (lane_id() == 0 && counts) {
        atomicAdd(counts + 0, a);
        atomicAdd(counts + 1, b);
        if (SUB)
            atomicAdd(counts + 2, c);
    }
}

/* BloomFilter::contains for the canonical hash h0: the h bit tests, stopping
 * at the first clear bit. */
__device__ inline bool filter_contains(const FilterDev &f, uint64_t h0, uint64_t kseed) {
    for (uint32_t j = 0; j < f.h; ++j) {
        const uint64_t p = fastmod(multi_hash(h0, j, kseed), f.m, f.magic, f.sh1, f.sh2);
        if (!((f.bits[p >> 3] >> (p & 7)) & 1u))
            return false;
    }
    return true;
}

/* SeqEval::evalSimple (SeqEval.h:37-114) and evalMinMatchLen (:295-332)
 * with a subtract filter, one lane per read: the test the progressive
 * filter builders apply to each read (BloomFilterGenerator.cpp:176-262,
 * 404-470).  A k-mer that the filter and the subtract filter both contain
 * keeps a hit's control flow but adds nothing to the score (SIMPLE) and
 * leaves the match length as it is (LENGTH). */
struct SubEvalArgs {
    const uint8_t *seq;
    const uint64_t *off;
    uint64_t nreads;
    FilterDev f, sub;
    uint32_t has_sub, length_mode;
    double threshold;
    uint32_t streak, min_len;
    uint8_t *pass;
};

__global__ __launch_bounds__(kBlock) void k_eval_sub(SubEvalArgs a) {
    const uint32_t kv = a.f.k;
    const uint64_t kseed = uint64_t(kv) * kMultiSeed;
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < a.nreads;
         r += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t b = a.off[r], len = a.off[r + 1] - b;
        const uint8_t *s = a.seq + b;
        /* evalSimple: thres, antiThres from the size_t frame count */
        const double n = double(uint64_t(len - kv + 1));
        const double thres = a.threshold * n;
        const double anti_thres = floor((1.0 - a.threshold) * n);
        double score = 0;
        uint32_t anti = 0, streak = 0, match = 0;
        uint64_t prev = 0, skip_end = 0;
        bool first = true, skipping = false;
        int res = -1; /* undecided / false / true */
        uint64_t fwd = 0, rc = 0, win = 0;
        uint32_t run = 0;
        for (uint64_t i = 0; i < len && res < 0; ++i) {
            const unsigned c = s[i];
            if (!base_valid(c)) {
                run = 0;
                fwd = rc = 0;
                continue;
            }
            const unsigned code = base_code(c);
            if (run >= kv) {
                const unsigned out = unsigned(win >> (2 * (kv - 1))) & 3u;
                fwd = rotl1(fwd) ^ rotl64(seed_fwd(out), kv) ^ seed_fwd(code);
                rc = rotr1(rc ^ seed_rc(out) ^ rotl64(seed_rc(code), kv));
            } else {
                fwd = rotl1(fwd) ^ seed_fwd(code);
                rc = rotr1(rc ^ rotl64(seed_rc(code), kv));
                ++run;
            }
            win = (win << 2) | code;
            if (run < kv)
                continue;
            const uint64_t p = i + 1 - kv;
            const uint64_t h0 = rc < fwd ? rc : fwd;
            if (a.length_mode) {
                if (len - p + match < a.min_len) {
                    res = 0;
                    break;
                }
                if (p != prev + 1)
                    match = 0;
                if (filter_contains(a.f, h0, kseed)) {
                    if (!a.has_sub || !filter_contains(a.sub, h0, kseed))
                        match = match == 0 ? kv : match + 1;
                } else {
                    match = 0;
                }
                if (match >= a.min_len) {
                    res = 1;
                    break;
                }
                prev = p;
                continue;
            }
            if (skipping) {
                /* the skip after a miss that ends a streak: every window up
                 * to skipEnd counts against the read, gaps inside included */
                if (p < skip_end) {
                    if (anti_thres <= double(++anti)) {
                        res = 0;
                        break;
                    }
                    prev = p;
                    continue;
                }
                skipping = false;
            }
            if (first) {
                first = false;
                if (filter_contains(a.f, h0, kseed)) {
                    if (!a.has_sub || !filter_contains(a.sub, h0, kseed))
                        score += 0.5;
                    if (thres <= score) {
                        res = 1;
                        break;
                    }
                    ++streak;
                } else if (anti_thres <= double(++anti)) {
                    res = 0;
                    break;
                }
                prev = p;
                continue;
            }
            if (p != prev + 1) {
                anti += uint32_t(p - prev - 1);
                if (anti_thres <= double(anti)) {
                    res = 0;
                    break;
                }
                streak = 0;
            }
            if (filter_contains(a.f, h0, kseed)) {
                if (!a.has_sub || !filter_contains(a.sub, h0, kseed))
                    score += streak == 0 ? 0.5 : 1.0;
                if (thres <= score) {
                    res = 1;
                    break;
                }
                prev = p;
                ++streak;
            } else {
                if (anti_thres <= double(++anti)) {
                    res = 0;
                    break;
                }
                prev = p;
                if (streak >= a.streak) {
                    skipping = true;
                    skip_end = p + kv;
                }
                streak = 0;
            }
        }
        a.pass[r] = res == 1 ? 1 : 0;
    }
}

/* Independent random byte loads over a power-of-two buffer, 8 in flight per
 * lane: the random-line rate that bounds a Bloom probe. */
__global__ __launch_bounds__(kBlock) void k_rand_gather(const uint8_t *buf, uint64_t mask,
                                                        uint32_t rounds,
                                                        unsigned long long *sink) {
    uint64_t x = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 0x9E3779B97F4A7C15ull + 1;
    unsigned acc = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        unsigned v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            v[u] = buf[x & mask];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            acc += v[u];
    }
    if (acc == 0xFFFFFFFFu)
        sink[0] = acc;
}

__global__ void k_popcount(const uint64_t *w, uint64_t nwords, unsigned long long *out) {
    unsigned long long c = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nwords;
         i += uint64_t(gridDim.x) * blockDim.x)
        c += __popcll(w[i]);
    for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o);
    if (lane_id() == 0)
        atomicAdd(out, c);
}

} // namespace

/* ======================================================================= */
/* filter set                                                               */

/* Device work queues and replay state of one stream of bbt_eval_batch calls.
 * A filter set owns one (bbt_eval_batch); callers that evaluate batches on
 * several streams at once give each stream its own (bbt_eval_batch_ws). */
struct bbt_workspace {
    /* batched path: [0..1] lane queues, [2..3] wave queues, [4..5] list
     * sizes, [8..40] bucket counts / cursors, [48..80] bucket bases */
    unsigned long long *dctl = nullptr;
    uint32_t *dlists = nullptr; /* 2 * lists_cap read indices */
    uint64_t lists_cap = 0;
    /* replayed batch (BBT_FLAG_GRAPH) */
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;
    /* captures run on this stream: the caller's may be the null stream,
     * which cannot be captured; the graph is launched on the caller's */
    hipStream_t cap_stream = nullptr;
    /* walk-kernel timing (BBT_FLAG_TIME_WALK) */
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_walk_ms = -1;
    const char *last_walk_kernel = ""; /* the walk kernel of the last batch */
    /* -D: the batch's low-complexity bitmap and the pre-pass's scratch */
    uint32_t *dmask = nullptr;
    uint64_t dmask_words = 0;
    uint8_t *dscratch = nullptr; /* sized for dscratch_w */
    uint32_t dscratch_w = 0;
    uint32_t *dfall = nullptr; /* reads the LDS pass leaves to the scratch pass */
    uint64_t dfall_cap = 0;
    /* BBT_READS_2BIT: the batch unpacked to ASCII */
    uint8_t *dascii = nullptr;
    uint64_t ascii_cap = 0;
    /* BINOMIAL scores past the table: the listed items */
    ScoreFix *dfix = nullptr;
    uint64_t fix_cap = 0;
};

static void workspace_release(bbt_workspace *w) {
    if (w->dctl)
        (void)hipFree(w->dctl);
    if (w->dlists)
        (void)hipFree(w->dlists);
    if (w->graph_exec)
        (void)hipGraphExecDestroy(w->graph_exec);
    if (w->cap_stream)
        (void)hipStreamDestroy(w->cap_stream);
    if (w->ev0)
        (void)hipEventDestroy(w->ev0);
    if (w->ev1)
        (void)hipEventDestroy(w->ev1);
    if (w->dmask)
        (void)hipFree(w->dmask);
    if (w->dscratch)
        (void)hipFree(w->dscratch);
    if (w->dfall)
        (void)hipFree(w->dfall);
    if (w->dascii)
        (void)hipFree(w->dascii);
    if (w->dfix)
        (void)hipFree(w->dfix);
    *w = bbt_workspace();
}

struct bbt_filterset {
    int device = 0;
    std::vector<FilterDev> desc;          /* host copy, device pointers inside */
    std::vector<void *> dbits;
    std::vector<std::string> ids;
    std::vector<bbt_filter_info> info;
    FilterDev *ddesc = nullptr;
    bool dirty = true;
    /* BINOMIAL thresholds per filter: calcMinCount for every frame length */
    std::vector<uint32_t *> dbinom;
    double binom_min_fpr = -1;
    uint32_t binom_frames = 0;
    std::vector<double> binom_fpr; /* each filter's FPR the thresholds were made for */
    /* BINOMIAL evalScore tables per filter (and the FPR they were made for) */
    std::vector<double *> dbscore;
    std::vector<double> bscore_fpr;
    uint32_t bscore_frames = 0;
    /* work queue counter and per-call scratch */
    unsigned long long *dscratch = nullptr; /* [3] popcount, [4..6] insert counts */
    int cu_count = 0;
    /* scheduling knobs (bbt_set_tuning / bbt_autotune); without them
     * (tuned false) a batch takes the tuning table's entry for its geometry */
    bbt_tuning tune = {0, 0, 0, 0};
    bool tuned = false;
    bbt_workspace ws; /* of bbt_eval_batch */
    std::mutex mu;
    /* groups of filters of one geometry, each interleaved byte by byte
     * (probe_ahead bit 2): rebuilt from the filters' own arrays by every batch
     * that walks them; ddesc_ilv describes every filter, group members at
     * their slot of the group's copy (bits + q, bshift), the others as in
     * ddesc */
    struct IlvGroup {
        IlvMembers mem;
        uint32_t shift = 0;
        uint8_t *buf = nullptr;
        uint64_t bytes = 0;
    };
    std::vector<IlvGroup> ilv;
    FilterDev *ddesc_ilv = nullptr;
    bool ilv_desc_fresh = false;
    /* with one group: its members, then the other filters (walked-filter
     * maps of the read-per-lane walk and of the walk of the rest) */
    uint32_t *dmap = nullptr;
};

static void fill_desc(FilterDev &d, void *bits, uint64_t m, uint32_t h, uint32_t k) {
    std::memset(&d, 0, sizeof(d));
    d.bits = static_cast<const uint8_t *>(bits);
    d.m = m;
    FastMod fm = fastmod_make(m);
    d.magic = fm.magic;
    d.sh1 = fm.sh1;
    d.sh2 = fm.sh2;
    d.h = h;
    d.k = k;
    for (uint32_t j = 1; j < kMaxHash; ++j)
        d.mult[j] = uint64_t(j) ^ (uint64_t(k) * kMultiSeed);
    roll_fill(d.roll, k);
}

static bbt_status sync_desc(bbt_filterset *fs) {
    if (!fs->dirty)
        return BBT_OK;
    /* kernels of earlier batches on other streams may still read them */
    HIP_OK(hipDeviceSynchronize());
    const size_t n = fs->desc.size();
    if (!fs->ddesc) /* allocated with the set */
        return fail(BBT_ERR_ARG, "filter set without device descriptors");
    for (size_t i = 0; i < n; ++i) {
        fs->desc[i].fpr = fs->info[i].fpr;
        fs->desc[i].binom = i < fs->dbinom.size() ? fs->dbinom[i] : nullptr;
        fs->desc[i].binom_len = fs->dbinom.empty() ? 0 : fs->binom_frames + 1;
        fs->desc[i].bscore = i < fs->dbscore.size() ? fs->dbscore[i] : nullptr;
        fs->desc[i].bscore_frames = fs->dbscore.empty() ? 0 : fs->bscore_frames;
    }
    HIP_OK(hipMemcpy(fs->ddesc, fs->desc.data(), sizeof(FilterDev) * n,
                     hipMemcpyHostToDevice));
    fs->dirty = false;
    fs->ilv_desc_fresh = false;
    return BBT_OK;
}

/* The interleave plan of a filter set: its filters grouped by (m, h, k) in
 * set order, each group of at least 2 (up to 64 per group; more of one
 * geometry make further groups) whose copy, the member count rounded up to a
 * power of two times m/8 bytes, stays within kIlvMaxBytes. */
constexpr uint64_t kIlvMaxBytes = uint64_t(32) << 30;
constexpr uint32_t kDefaultIlvAhead = 7;
/* ... and with the read-per-lane walk, for batches of at least
 * kRowsMinReads reads: a lone long read is walked against all of a group's
 * filters by one wave, which sets the latency of a small launch (a serving
 * request of 10k reads against the serving set: 2.7 -> 7.4 ms) */
constexpr uint32_t kDefaultRowsAhead = 15;
constexpr uint64_t kRowsMinReads = 65536;

static std::vector<bbt_filterset::IlvGroup> ilv_plan(const bbt_filterset *fs) {
    std::vector<bbt_filterset::IlvGroup> out;
    const size_t n = fs->desc.size();
    if (n > 256) /* members are byte indices */
        return out;
    std::vector<bool> used(n, false);
    for (size_t a = 0; a < n; ++a) {
        if (used[a])
            continue;
        const FilterDev &da = fs->desc[a];
        bbt_filterset::IlvGroup g;
        g.mem.n = 0;
        for (size_t b = a; b < n && g.mem.n < 64; ++b) {
            const FilterDev &db = fs->desc[b];
            if (!used[b] && db.m == da.m && db.h == da.h && db.k == da.k) {
                used[b] = true;
                g.mem.f[g.mem.n++] = uint8_t(b);
            }
        }
        if (g.mem.n < 2)
            continue;
        uint32_t sh = 1;
        while ((1u << sh) < g.mem.n)
            ++sh;
        if (da.m / 8 > (kIlvMaxBytes >> sh))
            continue;
        g.shift = sh;
        g.bytes = (da.m / 8) << sh;
        out.push_back(g);
    }
    return out;
}

static bool ilv_any(const bbt_filterset *fs) { return !ilv_plan(fs).empty(); }

/* every filter of the set in one group */
static bool ilv_whole(const bbt_filterset *fs) {
    const auto plan = ilv_plan(fs);
    return plan.size() == 1 && plan[0].mem.n == fs->desc.size();
}

/* The read-per-lane walk (probe_ahead bit 3) applies to a batch: the set's
 * plan has one group, of 3 to 16 filters, SIMPLE evalRead in STD mode on the
 * batched path, and the lane-walked reads shorter than 32768 bases; when
 * some reads reach the whole-wave length the group has at most 8 filters
 * (the whole-wave walk of a long read against all of them).  Returns the
 * group's size, or 0. */
static uint32_t rows_group(const bbt_filterset *fs,
                           const std::vector<bbt_filterset::IlvGroup> &plan,
                           const bbt_eval_params *p, uint64_t max_len, uint32_t wave_min) {
    if (plan.size() != 1 || p->scorer != BBT_SIMPLE || p->mode != BBT_STD ||
        p->path != BBT_PATH_BATCHED)
        return 0;
    /* the long-read walk sums a chunk's per-filter lookups in 16-bit fields:
     * up to 64 h */
    if (fs->desc[plan[0].mem.f[0]].h > 1023)
        return 0;
    const uint32_t gn = plan[0].mem.n;
    const bool buckets = max_len >= wave_min;
    const bool ok = gn >= 3 && gn <= 16 &&
                    (buckets ? gn <= 8 && wave_min <= 32768 : max_len < 32768);
    return ok ? gn : 0;
}

/* The interleaved copies' memory and descriptors (not their contents: the
 * batch's k_interleave writes them); false when no group qualifies. */
static bbt_status ensure_ilv(bbt_filterset *fs, bool &ok) {
    ok = false;
    std::vector<bbt_filterset::IlvGroup> plan = ilv_plan(fs);
    if (plan.empty())
        return BBT_OK;
    const size_t n = fs->desc.size();
    bool same = plan.size() == fs->ilv.size();
    for (size_t g = 0; same && g < plan.size(); ++g)
        same = plan[g].bytes <= fs->ilv[g].bytes && plan[g].shift == fs->ilv[g].shift &&
               plan[g].mem.n == fs->ilv[g].mem.n &&
               std::memcmp(plan[g].mem.f, fs->ilv[g].mem.f, plan[g].mem.n) == 0;
    if (!same || !fs->ddesc_ilv) {
        HIP_OK(hipDeviceSynchronize()); /* batches of other streams may read them */
        for (auto &g : fs->ilv)
            if (g.buf)
                (void)hipFree(g.buf);
        fs->ilv.clear();
        for (auto &g : plan) {
            HIP_OK(hipMalloc(&g.buf, g.bytes));
            fs->ilv.push_back(g);
        }
        if (!fs->ddesc_ilv)
            HIP_OK(hipMalloc(&fs->ddesc_ilv, sizeof(FilterDev) * n));
        fs->ilv_desc_fresh = false;
    }
    if (!fs->ilv_desc_fresh) {
        HIP_OK(hipDeviceSynchronize());
        std::vector<FilterDev> v = fs->desc;
        for (const auto &g : fs->ilv)
            for (uint32_t q = 0; q < g.mem.n; ++q) {
                v[g.mem.f[q]].bits = g.buf + q;
                v[g.mem.f[q]].bshift = g.shift;
            }
        HIP_OK(hipMemcpy(fs->ddesc_ilv, v.data(), sizeof(FilterDev) * n, hipMemcpyHostToDevice));
        if (fs->ilv.size() == 1) {
            std::vector<uint32_t> map;
            std::vector<bool> in_group(n, false);
            for (uint32_t q = 0; q < fs->ilv[0].mem.n; ++q) {
                map.push_back(fs->ilv[0].mem.f[q]);
                in_group[fs->ilv[0].mem.f[q]] = true;
            }
            for (uint32_t f = 0; f < n; ++f)
                if (!in_group[f])
                    map.push_back(f);
            if (!fs->dmap)
                HIP_OK(hipMalloc(&fs->dmap, sizeof(uint32_t) * n));
            HIP_OK(hipMemcpy(fs->dmap, map.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
        }
        fs->ilv_desc_fresh = true;
    }
    ok = true;
    return BBT_OK;
}

static bbt_status alloc_set(uint32_t n, const uint64_t *m, const uint32_t *h,
                            const uint32_t *k, bbt_filterset **out) {
    if (!out || n == 0 || !m || !h || !k)
        return fail(BBT_ERR_ARG, "bbt_filters_create: empty filter list");
    for (uint32_t i = 0; i < n; ++i) {
        if (m[i] == 0 || m[i] % 64 != 0)
            return fail(BBT_ERR_ARG, "filter size must be a positive multiple of 64 bits");
        if (h[i] == 0 || h[i] > kMaxHash)
            return fail(BBT_ERR_ARG, "hash count must be in 1..16");
        if (k[i] == 0 || k[i] > 32)
            return fail(BBT_ERR_UNSUPPORTED, "k-mer size must be in 1..32");
    }
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    bbt_filterset *fs = new bbt_filterset();
    fs->device = dev;
    HIP_OK(hipDeviceGetAttribute(&fs->cu_count, hipDeviceAttributeMultiprocessorCount, dev));
    for (uint32_t i = 0; i < n; ++i) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, m[i] / 8);
        if (e != hipSuccess) {
            bbt_filters_free(fs);
            return fail(BBT_ERR_NOMEM, "device allocation for filter bits failed");
        }
        e = hipMemset(p, 0, m[i] / 8);
        if (e != hipSuccess) {
            (void)hipFree(p);
            bbt_filters_free(fs);
            return fail(BBT_ERR_HIP, hipGetErrorString(e));
        }
        fs->dbits.push_back(p);
        FilterDev d;
        fill_desc(d, p, m[i], h[i], k[i]);
        fs->desc.push_back(d);
        bbt_filter_info inf;
        std::memset(&inf, 0, sizeof(inf));
        inf.m = m[i];
        inf.h = h[i];
        inf.k = k[i];
        fs->info.push_back(inf);
        fs->ids.push_back("filter" + std::to_string(i));
    }
    if (hipMalloc(&fs->dscratch, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&fs->ddesc, sizeof(FilterDev) * n) != hipSuccess) {
        bbt_filters_free(fs);
        return fail(BBT_ERR_NOMEM, "device allocation for scratch failed");
    }
    *out = fs;
    return BBT_OK;
}

static void set_fpr(bbt_filterset *fs, uint32_t idx, uint64_t pop) {
    bbt_filter_info &inf = fs->info[idx];
    inf.popcount = pop;
    /* the FPR the binomial scorer uses: (popcount / m)^h */
    inf.fpr = std::pow(double(pop) / double(inf.m), double(inf.h));
    fs->dirty = true;
}

extern "C" {

const char *bbt_last_error(void) { return last_error(); }

bbt_status bbt_filters_create(uint32_t n, const uint64_t *m, const uint32_t *h,
                              const uint32_t *k, bbt_filterset **out) {
    return alloc_set(n, m, h, k, out);
}

bbt_status bbt_filters_load(const char *const *paths, uint32_t n, bbt_filterset **out) {
    if (!paths || n == 0 || !out)
        return fail(BBT_ERR_ARG, "Error: Need Filter File (-f)");
    std::vector<BfHeader> hdrs(n);
    std::vector<std::vector<uint8_t>> bits(n);
    std::vector<InfoFile> infos(n);
    std::string err;
    for (uint32_t i = 0; i < n; ++i) {
        std::string bf = paths[i] ? paths[i] : "";
        /* <x>.bf -> <x>.txt as BioBloomClassifier.cpp:957 */
        if (bf.size() < 2)
            return fail(BBT_ERR_ARG, "Error: " + bf + " File cannot be opened");
        std::string txt = bf.substr(0, bf.size() - 2) + "txt";
        if (!read_info_file(txt, infos[i], err))
            return fail(BBT_ERR_IO, err);
        if (!read_bf_file(bf, hdrs[i], bits[i], err))
            return fail(BBT_ERR_FORMAT, err);
    }
    std::vector<uint64_t> m(n);
    std::vector<uint32_t> h(n), k(n);
    for (uint32_t i = 0; i < n; ++i) {
        m[i] = hdrs[i].size;
        h[i] = hdrs[i].nhash;
        k[i] = hdrs[i].kmer;
    }
    bbt_filterset *fs = nullptr;
    bbt_status st = alloc_set(n, m.data(), h.data(), k.data(), &fs);
    if (st != BBT_OK)
        return st;
    for (uint32_t i = 0; i < n; ++i) {
        st = bbt_filter_upload(fs, i, bits[i].data(), bits[i].size());
        if (st != BBT_OK) {
            bbt_filters_free(fs);
            return st;
        }
        fs->ids[i] = infos[i].filter_id;
        fs->info[i].desired_fpr = infos[i].desired_fpr;
    }
    *out = fs;
    return BBT_OK;
}

bbt_status bbt_filters_free(bbt_filterset *fs) {
    if (!fs)
        return BBT_OK;
    for (void *p : fs->dbits)
        if (p)
            (void)hipFree(p);
    for (uint32_t *p : fs->dbinom)
        if (p)
            (void)hipFree(p);
    for (double *p : fs->dbscore)
        if (p)
            (void)hipFree(p);
    if (fs->ddesc)
        (void)hipFree(fs->ddesc);
    if (fs->dscratch)
        (void)hipFree(fs->dscratch);
    for (auto &g : fs->ilv)
        if (g.buf)
            (void)hipFree(g.buf);
    if (fs->ddesc_ilv)
        (void)hipFree(fs->ddesc_ilv);
    if (fs->dmap)
        (void)hipFree(fs->dmap);
    workspace_release(&fs->ws);
    delete fs;
    return BBT_OK;
}

uint32_t bbt_filterset_size(const bbt_filterset *fs) {
    return fs ? uint32_t(fs->desc.size()) : 0;
}

bbt_status bbt_filter_info_get(const bbt_filterset *fs, uint32_t idx, bbt_filter_info *out) {
    if (!fs || idx >= fs->info.size() || !out)
        return fail(BBT_ERR_ARG, "bbt_filter_info_get: bad index");
    *out = fs->info[idx];
    return BBT_OK;
}

const char *bbt_filter_id(const bbt_filterset *fs, uint32_t idx) {
    if (!fs || idx >= fs->ids.size())
        return nullptr;
    return fs->ids[idx].c_str();
}

bbt_status bbt_filter_set_id(bbt_filterset *fs, uint32_t idx, const char *id) {
    if (!fs || idx >= fs->ids.size() || !id)
        return fail(BBT_ERR_ARG, "bbt_filter_set_id: bad argument");
    fs->ids[idx] = id;
    return BBT_OK;
}

const void *bbt_filter_device_bits(const bbt_filterset *fs, uint32_t idx) {
    if (!fs || idx >= fs->dbits.size())
        return nullptr;
    return fs->dbits[idx];
}

bbt_status bbt_filter_upload(bbt_filterset *fs, uint32_t idx, const void *host_bits,
                             uint64_t nbytes) {
    if (!fs || idx >= fs->desc.size() || !host_bits)
        return fail(BBT_ERR_ARG, "bbt_filter_upload: bad argument");
    if (nbytes != fs->info[idx].m / 8)
        return fail(BBT_ERR_ARG, "bbt_filter_upload: size does not match the filter");
    HIP_OK(hipMemcpy(fs->dbits[idx], host_bits, nbytes, hipMemcpyHostToDevice));
    const uint8_t *b = static_cast<const uint8_t *>(host_bits);
    uint64_t pop = 0;
    for (uint64_t i = 0; i + 8 <= nbytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        pop += uint64_t(__builtin_popcountll(w));
    }
    set_fpr(fs, idx, pop);
    return BBT_OK;
}

bbt_status bbt_filter_download(const bbt_filterset *fs, uint32_t idx, void *host_bits,
                               uint64_t nbytes) {
    if (!fs || idx >= fs->desc.size() || !host_bits)
        return fail(BBT_ERR_ARG, "bbt_filter_download: bad argument");
    if (nbytes != fs->info[idx].m / 8)
        return fail(BBT_ERR_ARG, "bbt_filter_download: size does not match the filter");
    HIP_OK(hipMemcpy(host_bits, fs->dbits[idx], nbytes, hipMemcpyDeviceToHost));
    return BBT_OK;
}

bbt_status bbt_filter_refresh(bbt_filterset *fs, uint32_t idx, void *stream) {
    if (!fs || idx >= fs->desc.size())
        return fail(BBT_ERR_ARG, "bbt_filter_refresh: bad index");
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> g(fs->mu);
    unsigned long long *pop = fs->dscratch + 3;
    k_zero_u64<<<1, 64, 0, st>>>(pop, 1);
    const uint64_t nwords = fs->info[idx].m / 64;
    k_popcount<<<1024, kBlock, 0, st>>>(static_cast<const uint64_t *>(fs->dbits[idx]),
                                        nwords, pop);
    HIP_OK(hipGetLastError());
    unsigned long long hp = 0;
    HIP_OK(hipMemcpyAsync(&hp, pop, sizeof(hp), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    set_fpr(fs, idx, hp);
    return BBT_OK;
}

/* k_insert over dseq[0..nbytes), with the subtract filter when sub != null;
 * hc receives {new, already present, removed}.  Caller holds fs->mu. */
static bbt_status insert_impl(bbt_filterset *fs, uint32_t idx, const FilterDev *sub,
                              const char *dseq, uint64_t nbytes, unsigned long long hc[3],
                              hipStream_t st, bool ordered = false) {
    unsigned long long *cnt = fs->dscratch + 4;
    k_zero_u64<<<1, 64, 0, st>>>(cnt, 3);
    if (ordered && nbytes > 0xFFFFFFFFull)
        return fail(BBT_ERR_ARG, "bbt_filter_insert_ordered: sequence above 4 GB");
    /* ordered: one thread takes every window in order, so the new / present
     * split is the one a sequential insert gives */
    const uint32_t chunk = ordered ? uint32_t(std::max<uint64_t>(nbytes, 1)) : 256;
    const uint64_t threads = (nbytes + chunk - 1) / chunk;
    const uint64_t blocks = (threads + kBlock - 1) / kBlock;
    if (blocks > 0x7FFFFFFFull)
        return fail(BBT_ERR_ARG, "bbt_filter_insert: sequence too large for one call");
    const FilterDev &d = fs->desc[idx];
    const uint8_t *sq = reinterpret_cast<const uint8_t *>(dseq);
    if (blocks) {
        const dim3 g{static_cast<uint32_t>(blocks)};
        if (sub)
            d.k == 25 ? k_insert<25, true><<<g, kBlock, 0, st>>>(sq, nbytes, d, *sub, chunk, cnt)
                      : k_insert<0, true><<<g, kBlock, 0, st>>>(sq, nbytes, d, *sub, chunk, cnt);
        else
            d.k == 25 ? k_insert<25, false><<<g, kBlock, 0, st>>>(sq, nbytes, d, d, chunk, cnt)
                      : k_insert<0, false><<<g, kBlock, 0, st>>>(sq, nbytes, d, d, chunk, cnt);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipMemcpyAsync(hc, cnt, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return BBT_OK;
}

bbt_status bbt_filter_insert(bbt_filterset *fs, uint32_t idx, const char *dseq,
                             uint64_t nbytes, uint64_t *counts, void *stream) {
    if (!fs || idx >= fs->desc.size() || (!dseq && nbytes))
        return fail(BBT_ERR_ARG, "bbt_filter_insert: bad argument");
    std::lock_guard<std::mutex> g(fs->mu);
    unsigned long long hc[3] = {0, 0, 0};
    const bbt_status rc = insert_impl(fs, idx, nullptr, dseq, nbytes, hc,
                                      static_cast<hipStream_t>(stream));
    if (rc == BBT_OK && counts) {
        counts[0] = hc[0];
        counts[1] = hc[1];
    }
    return rc;
}

bbt_status bbt_filter_insert_ordered(bbt_filterset *fs, uint32_t idx, const char *dseq,
                                     uint64_t nbytes, uint64_t *counts, void *stream) {
    if (!fs || idx >= fs->desc.size() || (!dseq && nbytes))
        return fail(BBT_ERR_ARG, "bbt_filter_insert_ordered: bad argument");
    std::lock_guard<std::mutex> g(fs->mu);
    unsigned long long hc[3] = {0, 0, 0};
    const bbt_status rc = insert_impl(fs, idx, nullptr, dseq, nbytes, hc,
                                      static_cast<hipStream_t>(stream), true);
    if (rc == BBT_OK && counts) {
        counts[0] = hc[0];
        counts[1] = hc[1];
    }
    return rc;
}

bbt_status bbt_filter_insert_subtract(bbt_filterset *fs, uint32_t idx,
                                      const bbt_filterset *sub, uint32_t sub_idx,
                                      const char *dseq, uint64_t nbytes, uint64_t *counts,
                                      void *stream) {
    if (!fs || idx >= fs->desc.size() || !sub || sub_idx >= sub->desc.size() ||
        (!dseq && nbytes))
        return fail(BBT_ERR_ARG, "bbt_filter_insert_subtract: bad argument");
    const FilterDev &d = fs->desc[idx], &s = sub->desc[sub_idx];
    /* loadFilterSubtract's checks (BloomFilterGenerator.h:222-236) */
    if (d.h != s.h)
        return fail(BBT_ERR_ARG, "Error: Subtraction filter's hash number " +
                                     std::to_string(d.h) +
                                     " is a different size from than output filter's hash "
                                     "number " + std::to_string(s.h));
    if (d.k != s.k)
        return fail(BBT_ERR_ARG, "Error: Subtraction filter's k-mer size " +
                                     std::to_string(d.k) +
                                     " is a different size from than output filter's k-mer "
                                     "size " + std::to_string(s.k));
    if (s.bits == d.bits)
        return fail(BBT_ERR_ARG, "bbt_filter_insert_subtract: subtract filter is the output");
    std::lock_guard<std::mutex> g(fs->mu);
    unsigned long long hc[3] = {0, 0, 0};
    const FilterDev sd = s;
    const bbt_status rc = insert_impl(fs, idx, &sd, dseq, nbytes, hc,
                                      static_cast<hipStream_t>(stream));
    if (rc == BBT_OK && counts) {
        counts[0] = hc[0];
        counts[1] = hc[1];
        counts[2] = hc[2];
    }
    return rc;
}

bbt_status bbt_eval_subtract(const bbt_filterset *fs, uint32_t idx, const bbt_filterset *sub,
                             uint32_t sub_idx, const bbt_reads *in, uint32_t scorer,
                             double threshold, uint32_t streak, uint8_t *pass, void *stream) {
    if (!fs || idx >= fs->desc.size() || !in || (in->nreads && (!in->seq || !in->off || !pass)) ||
        (sub && sub_idx >= sub->desc.size()))
        return fail(BBT_ERR_ARG, "bbt_eval_subtract: bad argument");
    if (in->format != BBT_READS_ASCII)
        return fail(BBT_ERR_UNSUPPORTED, "bbt_eval_subtract: ASCII reads only");
    if (scorer != BBT_SIMPLE && scorer != BBT_LENGTH)
        return fail(BBT_ERR_UNSUPPORTED, "bbt_eval_subtract: SIMPLE or LENGTH scoring only");
    SubEvalArgs a;
    std::memset(&a, 0, sizeof(a));
    a.f = fs->desc[idx];
    if (sub) {
        a.sub = sub->desc[sub_idx];
        if (a.sub.h != a.f.h || a.sub.k != a.f.k)
            return fail(BBT_ERR_ARG, "bbt_eval_subtract: the subtract filter's hash number and "
                                     "k-mer size must equal the filter's");
        a.has_sub = 1;
    }
    if (!in->nreads)
        return BBT_OK;
    a.seq = reinterpret_cast<const uint8_t *>(in->seq);
    a.off = in->off;
    a.nreads = in->nreads;
    a.length_mode = scorer == BBT_LENGTH;
    a.threshold = threshold;
    a.streak = streak;
    a.min_len = uint32_t(std::lround(threshold));
    a.pass = pass;
    const uint64_t blocks = std::min<uint64_t>((in->nreads + kBlock - 1) / kBlock, 65536);
    k_eval_sub<<<dim3(uint32_t(blocks)), kBlock, 0, static_cast<hipStream_t>(stream)>>>(a);
    HIP_OK(hipGetLastError());
    return BBT_OK;
}

bbt_status bbt_filter_store(const bbt_filterset *fs, uint32_t idx, const char *prefix,
                            double desired_fpr, uint64_t expected_entries,
                            uint64_t num_entries, uint64_t redundant,
                            const char *seq_sources) {
    if (!fs || idx >= fs->desc.size() || !prefix)
        return fail(BBT_ERR_ARG, "bbt_filter_store: bad argument");
    const bbt_filter_info &inf = fs->info[idx];
    std::vector<uint8_t> bits(inf.m / 8);
    HIP_OK(hipMemcpy(bits.data(), fs->dbits[idx], bits.size(), hipMemcpyDeviceToHost));
    BfHeader hdr;
    std::memset(&hdr, 0, sizeof(hdr));
    hdr.size = inf.m;
    hdr.nhash = inf.h;
    hdr.kmer = inf.k;
    hdr.dFPR = desired_fpr;
    hdr.nEntry = num_entries;
    hdr.tEntry = expected_entries;
    std::string err;
    std::string p = prefix;
    if (!write_bf_file(p + ".bf", hdr, bits.data(), bits.size(), err))
        return fail(BBT_ERR_IO, err);
    InfoFile ini;
    ini.filter_id = fs->ids[idx];
    ini.kmer_size = inf.k;
    ini.desired_fpr = desired_fpr;
    ini.hash_num = inf.h;
    ini.expected_entries = expected_entries;
    ini.sequence_sources = seq_sources ? seq_sources : "";
    ini.size = inf.m;
    ini.num_entries = num_entries;
    ini.approx_fpr = calc_approx_fpr(inf.m, num_entries, inf.h);
    ini.redundant_sequences = redundant;
    ini.redundant_fpr = calc_redundancy_fpr(inf.m, num_entries, inf.h);
    if (!write_info_file(p + ".txt", ini, err))
        return fail(BBT_ERR_IO, err);
    return BBT_OK;
}

uint64_t bbt_calc_optimal_size(uint64_t entries, double fpr, uint32_t h) {
    return calc_optimal_size(entries, fpr, h);
}

uint32_t bbt_calc_optimal_hashnum(double fpr) { return calc_optimal_hashnum(fpr); }

uint64_t bbt_calc_min_count(uint64_t frame_len, double bf_fpr, double min_fpr) {
    return calc_min_count(frame_len, bf_fpr, min_fpr);
}

double bbt_binomial_score(uint64_t count, uint64_t frame_len, double bf_fpr) {
    /* evalScore BINOMIAL: log10(P(X > count)) * -10 (SeqEval.h:545,558) */
    return std::log10(binom_sf(int64_t(count), frame_len, bf_fpr)) * -10;
}

bbt_status bbt_summary_text(const bbt_filterset *fs, const uint64_t *summary,
                            uint64_t nrec, char *buf, uint64_t cap, uint64_t *needed) {
    if (!fs || !summary)
        return fail(BBT_ERR_ARG, "bbt_summary_text: bad argument");
    std::string s = summary_text(fs->ids, summary, nrec);
    if (needed)
        *needed = s.size();
    if (buf && cap) {
        uint64_t n = std::min<uint64_t>(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return BBT_OK;
}

} // extern "C"

/* ======================================================================= */
/* classification dispatch                                                  */

namespace {

struct Launch {
    EvalArgs a;
    hipStream_t st;
    uint32_t path;
    int cu;
    uint32_t blocks_per_cu; /* persistent grid: 0 = as many as fit */
};

/* resident blocks per CU of a kernel of kBlock threads (at least 1) */
template <typename KERNEL>
int blocks_per_cu_of(KERNEL kernel) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess)
        per_cu = 1;
    return std::max(1, per_cu);
}

/* resident blocks per CU of one persistent-kernel instantiation, queried once
 * (a function-local static: initialised once, also under concurrent calls) */
template <int SC, bool EXH, int K>
int persistent_occupancy() {
    static const int cached = blocks_per_cu_of(k_walk_persistent<SC, EXH, K>);
    return cached;
}

template <int SC, bool EXH, int K>
bbt_status launch_walk(const Launch &L) {
    if (L.a.nitems == 0)
        return BBT_OK;
    if (L.path == BBT_PATH_PLAIN) {
        const uint64_t blocks = (L.a.nitems + kBlock - 1) / kBlock;
        if (blocks > 0x7FFFFFFFull)
            return fail(BBT_ERR_ARG, "batch too large for one launch");
        k_walk_plain<SC, EXH, K><<<dim3(uint32_t(blocks)), kBlock, 0, L.st>>>(L.a);
    } else {
        int per_cu = persistent_occupancy<SC, EXH, K>();
        if (L.blocks_per_cu)
            per_cu = std::min<int>(per_cu, int(L.blocks_per_cu));
        uint64_t blocks = uint64_t(L.cu) * uint64_t(per_cu);
        /* lane walks take 64 items per wave; when length buckets exist, any
         * item may be a whole-wave walk, which takes a wave of its own, so a
         * small batch of long reads still gets one wave per item */
        const uint64_t waves_per_block = kBlock / 64;
        const uint64_t need = L.a.list_counts
                                  ? (L.a.nitems + waves_per_block - 1) / waves_per_block
                                  : (L.a.nitems + kBlock - 1) / kBlock;
        blocks = std::max<uint64_t>(1, std::min(blocks, need));
        k_walk_persistent<SC, EXH, K><<<dim3(uint32_t(blocks)), kBlock, 0, L.st>>>(L.a);
    }
    HIP_OK(hipGetLastError());
    return BBT_OK;
}

template <int K, unsigned SH>
bbt_status launch_rows_k(const Launch &L, const uint8_t *ilv) {
    static const int per_cu = blocks_per_cu_of(k_walk_rows<K, SH>);
    int bpc = per_cu;
    if (L.blocks_per_cu)
        bpc = std::min<int>(bpc, int(L.blocks_per_cu));
    const uint64_t nreads = L.a.nitems / L.a.nw;
    const uint64_t blocks = std::max<uint64_t>(
        1, std::min<uint64_t>(uint64_t(L.cu) * uint64_t(bpc), (nreads + kBlock - 1) / kBlock));
    k_walk_rows<K, SH><<<dim3(uint32_t(blocks)), kBlock, 0, L.st>>>(L.a, ilv);
    HIP_OK(hipGetLastError());
    return BBT_OK;
}

template <int K, unsigned SH>
bbt_status launch_group_long_k(const Launch &L, const uint8_t *ilv) {
    static const int per_cu = blocks_per_cu_of(k_walk_group_long<K, SH>);
    int bpc = per_cu;
    if (L.blocks_per_cu)
        bpc = std::min<int>(bpc, int(L.blocks_per_cu));
    /* the long-read count is on the device: a full persistent grid, whose
     * waves leave at once when the list is short */
    k_walk_group_long<K, SH><<<dim3(uint32_t(L.cu) * uint32_t(bpc)), kBlock, 0, L.st>>>(L.a, ilv);
    HIP_OK(hipGetLastError());
    return BBT_OK;
}

bbt_status launch_rows(const Launch &L, const uint8_t *ilv, uint32_t sh, uint32_t k) {
    if (L.a.nitems == 0)
        return BBT_OK;
    if (L.a.list_counts) {
        /* length buckets: groups of up to 8 filters (rows_group) */
        const bbt_status s = k == 25 ? (sh == 3 ? launch_group_long_k<25, 3>(L, ilv)
                                                : launch_group_long_k<25, 2>(L, ilv))
                                     : (sh == 3 ? launch_group_long_k<0, 3>(L, ilv)
                                                : launch_group_long_k<0, 2>(L, ilv));
        if (s != BBT_OK)
            return s;
    }
    if (k == 25)
        return sh == 4 ? launch_rows_k<25, 4>(L, ilv)
                       : (sh == 3 ? launch_rows_k<25, 3>(L, ilv) : launch_rows_k<25, 2>(L, ilv));
    return sh == 4 ? launch_rows_k<0, 4>(L, ilv)
                   : (sh == 3 ? launch_rows_k<0, 3>(L, ilv) : launch_rows_k<0, 2>(L, ilv));
}

template <int SC, bool EXH>
bbt_status launch_k(const Launch &L, uint32_t k) {
    return k == 25 ? launch_walk<SC, EXH, 25>(L) : launch_walk<SC, EXH, 0>(L);
}

bbt_status launch_scorer(const Launch &L, uint32_t scorer, bool exh, uint32_t k) {
    switch (scorer) {
    case BBT_SIMPLE:
        return exh ? launch_k<SIMPLE, true>(L, k) : launch_k<SIMPLE, false>(L, k);
    case BBT_HARMONIC:
        return exh ? launch_k<HARMONIC, true>(L, k) : launch_k<HARMONIC, false>(L, k);
    case BBT_LENGTH:
        return exh ? launch_k<LENGTH, true>(L, k) : launch_k<LENGTH, false>(L, k);
    case BBT_BINOMIAL:
        return exh ? launch_k<BINOMIAL, true>(L, k) : launch_k<BINOMIAL, false>(L, k);
    }
    return fail(BBT_ERR_ARG, "unknown scorer");
}

template <int SC, int K>
bbt_status launch_ordered_k(const Launch &L, uint64_t nrec, uint32_t paired,
                            uint32_t inclusive) {
    const uint64_t blocks = (nrec + kBlock - 1) / kBlock;
    if (blocks)
        k_ordered<SC, K><<<dim3(uint32_t(blocks)), kBlock, 0, L.st>>>(L.a, nrec, paired,
                                                                      inclusive);
    HIP_OK(hipGetLastError());
    return BBT_OK;
}

template <int SC>
bbt_status launch_ordered_sc(const Launch &L, uint64_t nrec, uint32_t paired,
                             uint32_t inclusive, uint32_t k) {
    return k == 25 ? launch_ordered_k<SC, 25>(L, nrec, paired, inclusive)
                   : launch_ordered_k<SC, 0>(L, nrec, paired, inclusive);
}

bbt_status launch_ordered(const Launch &L, uint32_t scorer, uint64_t nrec, uint32_t paired,
                          uint32_t inclusive, uint32_t k) {
    switch (scorer) {
    case BBT_SIMPLE:
        return launch_ordered_sc<SIMPLE>(L, nrec, paired, inclusive, k);
    case BBT_HARMONIC:
        return launch_ordered_sc<HARMONIC>(L, nrec, paired, inclusive, k);
    case BBT_LENGTH:
        return launch_ordered_sc<LENGTH>(L, nrec, paired, inclusive, k);
    case BBT_BINOMIAL:
        return launch_ordered_sc<BINOMIAL>(L, nrec, paired, inclusive, k);
    }
    return fail(BBT_ERR_ARG, "unknown scorer");
}

/* calcMinCount tables for every frame length up to max_frames */
/* (remade when min_fpr, the frame range or any filter's FPR changed: uploads
 * and refreshes change the FPR, (popcount / m)^h) */
/* true when the tables are made for these parameters and every filter's
 * current FPR (uploads and refreshes change it) */
bool binom_fresh(const bbt_filterset *fs, double min_fpr, uint32_t max_frames) {
    bool fresh = !fs->dbinom.empty() && fs->binom_min_fpr == min_fpr &&
                 fs->binom_frames >= max_frames && fs->binom_fpr.size() == fs->desc.size();
    for (size_t f = 0; fresh && f < fs->desc.size(); ++f)
        fresh = fs->binom_fpr[f] == fs->info[f].fpr;
    return fresh;
}

bool bscore_fresh(const bbt_filterset *fs, uint32_t max_frames) {
    bool fresh = !fs->dbscore.empty() && fs->bscore_frames >= max_frames &&
                 fs->bscore_fpr.size() == fs->desc.size();
    for (size_t f = 0; fresh && f < fs->desc.size(); ++f)
        fresh = fs->bscore_fpr[f] == fs->info[f].fpr;
    return fresh;
}

bbt_status ensure_binom(bbt_filterset *fs, double min_fpr, uint32_t max_frames) {
    if (binom_fresh(fs, min_fpr, max_frames))
        return BBT_OK;
    HIP_OK(hipDeviceSynchronize()); /* in-flight batches read the old tables */
    for (uint32_t *p : fs->dbinom)
        if (p)
            (void)hipFree(p);
    fs->dbinom.clear();
    fs->binom_fpr.clear();
    for (size_t f = 0; f < fs->desc.size(); ++f) {
        std::vector<uint32_t> t(size_t(max_frames) + 1, 1u);
        for (uint32_t fl = 1; fl <= max_frames; ++fl)
            t[fl] = uint32_t(calc_min_count(fl, fs->info[f].fpr, min_fpr));
        uint32_t *d = nullptr;
        HIP_OK(hipMalloc(&d, t.size() * sizeof(uint32_t)));
        HIP_OK(hipMemcpy(d, t.data(), t.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        fs->dbinom.push_back(d);
        fs->binom_fpr.push_back(fs->info[f].fpr);
    }
    fs->binom_min_fpr = min_fpr;
    fs->binom_frames = max_frames;
    fs->dirty = true;
    return BBT_OK;
}

/* BINOMIAL evalScore values for every frame length up to max_frames and
 * every hit count: log10(P(X > c)) * -10, computed by the same host
 * binomial tail as the thresholds; remade when a filter's FPR changed. */
constexpr uint32_t kBinomScoreMaxFrames = 1024;

bbt_status ensure_bscore(bbt_filterset *fs, uint32_t max_frames) {
    if (bscore_fresh(fs, max_frames))
        return BBT_OK;
    HIP_OK(hipDeviceSynchronize()); /* in-flight batches read the old tables */
    for (double *p : fs->dbscore)
        if (p)
            (void)hipFree(p);
    fs->dbscore.clear();
    fs->bscore_fpr.clear();
    const uint32_t F = std::max(max_frames, fs->bscore_frames);
    const size_t n = (size_t(F) + 1) * (size_t(F) + 2) / 2;
    for (size_t f = 0; f < fs->desc.size(); ++f) {
        const double fpr = fs->info[f].fpr;
        std::vector<double> t(n, 0.0);
        for (uint32_t fl = 1; fl <= F; ++fl)
            for (uint32_t c = 0; c <= fl; ++c)
                t[size_t(fl) * (fl + 1) / 2 + c] = std::log10(binom_sf(int64_t(c), fl, fpr)) * -10;
        double *d = nullptr;
        HIP_OK(hipMalloc(&d, n * sizeof(double)));
        HIP_OK(hipMemcpy(d, t.data(), n * sizeof(double), hipMemcpyHostToDevice));
        fs->dbscore.push_back(d);
        fs->bscore_fpr.push_back(fpr);
    }
    fs->bscore_frames = F;
    fs->dirty = true;
    return BBT_OK;
}

} // namespace

/* ---- the tuning table ---------------------------------------------------
 * Knobs bbt_autotune chose, by batch geometry: filter-set size (log2 of its
 * bytes), filter count, common k (0 if mixed), largest h, longest read
 * (log2), mode, scorer, pairing. */
static std::mutex g_tune_mu;
static std::map<std::string, bbt_tuning> g_tune_table;

static unsigned floor_log2(uint64_t v) {
    unsigned l = 0;
    while (v > 1) {
        v >>= 1;
        ++l;
    }
    return l;
}

static std::string tune_key(const bbt_filterset *fs, const bbt_reads *in,
                            const bbt_eval_params *p) {
    uint64_t bytes = 0;
    uint32_t hmax = 0, kc = fs->desc.empty() ? 0 : fs->desc[0].k;
    for (size_t i = 0; i < fs->desc.size(); ++i) {
        bytes += fs->info[i].m / 8;
        hmax = std::max(hmax, fs->desc[i].h);
        if (fs->desc[i].k != kc)
            kc = 0;
    }
    char buf[192];
    std::snprintf(buf, sizeof(buf),
                  "bytes_log2=%u nf=%zu k=%u h=%u len_log2=%u mode=%u scorer=%u paired=%u",
                  floor_log2(bytes), fs->desc.size(), kc, hmax,
                  floor_log2(std::max<uint64_t>(1, in->max_len)), p->mode, p->scorer,
                  p->paired ? 1u : 0u);
    return buf;
}

/* Everything of one bbt_eval_batch that is host work (checks, tables,
 * allocations) and the argument block of its kernels. */
struct Prepared {
    Launch L;
    uint32_t nf;
    uint64_t nrec;
    uint32_t kk;
    uint32_t wave_min;
    const uint8_t *codes; /* BBT_READS_2BIT: the packed batch, unpacked first */
    uint64_t code_bytes;
    bool fix_scores;      /* BINOMIAL scores past the table, finished on the host */
    bool dust_big;        /* -D: handed-on reads in a workgroup's LDS (k_dust_big) */
    bool group;           /* read-major lane walks over the interleaved filters */
    bool rows;            /* one read per lane against every filter (k_walk_rows) */
    uint32_t rows_n;      /* the filters it walks: the group's, fs->dmap[0 .. rows_n) */
};

/* Whether k_dust_big can have `bytes` of dynamic LDS per workgroup on the
 * current device (asked once per device). */
static bool dust_big_ok(uint64_t bytes) {
    static std::mutex mu;
    static std::map<int, int> limit; /* device -> bytes, 0 if unavailable */
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return false;
    std::lock_guard<std::mutex> g(mu);
    auto it = limit.find(dev);
    if (it == limit.end()) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess ||
            hipFuncSetAttribute(reinterpret_cast<const void *>(k_dust_big),
                                hipFuncAttributeMaxDynamicSharedMemorySize, v) != hipSuccess) {
            (void)hipGetLastError();
            v = 0;
        }
        it = limit.emplace(dev, v).first;
    }
    return bytes <= uint64_t(it->second);
}

/* noalloc: the call may not allocate (bbt_eval_batch_noalloc): W already
 * holds its control words and lists, and what would need an allocation (the
 * -D pre-pass, graphs, timing events, BINOMIAL tables not yet built) is
 * rejected instead. */
static bbt_status prepare(bbt_filterset *fs, bbt_workspace &W, const bbt_reads *in,
                          const bbt_eval_params *p, const bbt_eval_out *out, hipStream_t st,
                          Prepared &P, bool noalloc) {
    const uint32_t nf = uint32_t(fs->desc.size());
    if (noalloc) {
        if (p->dust)
            return fail(BBT_ERR_UNSUPPORTED,
                        "bbt_eval_batch_noalloc: -D (dust) needs the sdust pre-pass's "
                        "scratch; use bbt_eval_batch");
        if (p->flags & (BBT_FLAG_GRAPH | BBT_FLAG_TIME_WALK))
            return fail(BBT_ERR_UNSUPPORTED,
                        "bbt_eval_batch_noalloc: graph and timing flags are not available");
    }
    if (!in->seq || !in->off)
        return fail(BBT_ERR_ARG, "bbt_eval_batch: reads not given");
    if (reinterpret_cast<uintptr_t>(in->seq) % 8 != 0 || in->seq_bytes % 8 != 0)
        return fail(BBT_ERR_ARG,
                    "bbt_eval_batch: seq must be 8-byte aligned with a size multiple of 8");
    if (in->format > BBT_READS_2BIT)
        return fail(BBT_ERR_ARG, "bbt_eval_batch: unknown read format");
    if (in->nreads > 0xFFFFFFFFull)
        return fail(BBT_ERR_ARG, "bbt_eval_batch: more than 2^32 - 1 reads in one batch "
                                 "(read lists hold 32-bit indices); split the batch");
    if (in->format == BBT_READS_2BIT) {
        if (noalloc)
            return fail(BBT_ERR_UNSUPPORTED,
                        "bbt_eval_batch_noalloc: 2-bit reads need an ASCII staging buffer; "
                        "use bbt_eval_batch");
        /* the unpacked batch: 4 bases per code byte, and a word of padding */
        const uint64_t need = 4 * in->seq_bytes + 8;
        if (W.ascii_cap < need) {
            if (W.dascii) {
                HIP_OK(hipStreamSynchronize(st)); /* the previous batch may read it */
                HIP_OK(hipFree(W.dascii));
            }
            W.dascii = nullptr;
            W.ascii_cap = 0;
            HIP_OK(hipMalloc(&W.dascii, need));
            W.ascii_cap = need;
        }
    }
    if (p->paired && (in->nreads & 1))
        return fail(BBT_ERR_ARG, "bbt_eval_batch: paired batch needs an even read count");
    if (p->mode > BBT_SCORES || p->scorer > BBT_BINOMIAL || p->path > BBT_PATH_BATCHED)
        return fail(BBT_ERR_ARG, "bbt_eval_batch: bad mode, scorer or path");
    if (!out->hits)
        return fail(BBT_ERR_ARG, "bbt_eval_batch: hits output is required");
    const bool need_scores = p->mode == BBT_BESTHIT || p->mode == BBT_SCORES;
    bool fix_scores = false; /* BINOMIAL scores of reads past the table */
    if (need_scores && !out->scores)
        return fail(BBT_ERR_ARG, "bbt_eval_batch: scores output required for this mode");
    if (need_scores && p->scorer == BBT_BINOMIAL) {
        uint32_t kmin = 64;
        for (const FilterDev &d : fs->desc)
            kmin = std::min(kmin, d.k);
        const uint32_t frames = in->max_len >= kmin ? in->max_len - kmin + 1 : 0;
        if (frames > kBinomScoreMaxFrames) {
            /* reads past the table: the device counts, the host scores */
            fix_scores = true;
            if (noalloc || (p->flags & BBT_FLAG_GRAPH))
                return fail(BBT_ERR_UNSUPPORTED,
                            "binomial evalScore of reads over 1024 k-mers is finished on the "
                            "host: not available without allocation or in a graph");
            const uint64_t need = in->nreads * uint64_t(nf);
            if (W.fix_cap < need) {
                if (W.dfix) {
                    HIP_OK(hipStreamSynchronize(st));
                    HIP_OK(hipFree(W.dfix));
                }
                W.dfix = nullptr;
                W.fix_cap = 0;
                HIP_OK(hipMalloc(&W.dfix, need * sizeof(ScoreFix)));
                W.fix_cap = need;
            }
        }
        const uint32_t table = std::min(frames, kBinomScoreMaxFrames);
        if (noalloc && !bscore_fresh(fs, table))
            return fail(BBT_ERR_UNSUPPORTED,
                        "bbt_eval_batch_noalloc: the binomial score tables for these reads "
                        "are not built; run bbt_eval_batch once with the same parameters");
        bbt_status s = ensure_bscore(fs, table);
        if (s != BBT_OK)
            return s;
    }
    if (out->summary && (2 * size_t(nf) + 2) * sizeof(unsigned int) > 64 * 1024)
        return fail(BBT_ERR_UNSUPPORTED, "too many filters for device summary counters");
    uint32_t common_k = fs->desc[0].k;
    for (const FilterDev &d : fs->desc)
        if (d.k != common_k)
            common_k = 0;
    if (p->scorer == BBT_BINOMIAL) {
        uint32_t kmin = 64;
        for (const FilterDev &d : fs->desc)
            kmin = std::min(kmin, d.k);
        const uint32_t frames = in->max_len >= kmin ? in->max_len - kmin + 1 : 0;
        if (noalloc && !binom_fresh(fs, p->threshold, frames))
            return fail(BBT_ERR_UNSUPPORTED,
                        "bbt_eval_batch_noalloc: the binomial thresholds for these reads are "
                        "not built; run bbt_eval_batch once with the same parameters");
        bbt_status s = ensure_binom(fs, p->threshold, frames);
        if (s != BBT_OK)
            return s;
    }
    bbt_status s = sync_desc(fs);
    if (s != BBT_OK)
        return s;
    if (!W.dctl)
        HIP_OK(hipMalloc(&W.dctl, 128 * sizeof(unsigned long long)));
    bool dust_big = false; /* -D: handed-on reads in a workgroup's LDS */
    if (p->dust) {
        if (p->dust_w < 3 || p->dust_w > uint32_t(kDustMaxW) || p->dust_t > 1000000)
            return fail(BBT_ERR_UNSUPPORTED, "dust window (-W) must be 3.." +
                                                 std::to_string(kDustMaxW));
        /* one bit per base position of the batch (4 per code byte when packed) */
        const uint64_t positions =
            in->format == BBT_READS_2BIT ? 4 * in->seq_bytes : in->seq_bytes;
        const uint64_t words = positions / 32 + 2;
        if (W.dmask_words < words) {
            if (W.dmask) {
                HIP_OK(hipStreamSynchronize(st));
                HIP_OK(hipFree(W.dmask));
            }
            W.dmask = nullptr;
            W.dmask_words = 0;
            HIP_OK(hipMalloc(&W.dmask, words * sizeof(uint32_t)));
            W.dmask_words = words;
        }
        if (W.dscratch_w < p->dust_w) {
            /* the scratch grows with the window (W^2 for the interval list) */
            const DustLayout lay = dust_layout(int(p->dust_w));
            if (W.dscratch) {
                HIP_OK(hipStreamSynchronize(st));
                HIP_OK(hipFree(W.dscratch));
            }
            W.dscratch = nullptr;
            W.dscratch_w = 0;
            HIP_OK(hipMalloc(&W.dscratch, uint64_t(lay.threads) * lay.stride));
            W.dscratch_w = p->dust_w;
        }
        /* asked here, before any capture of the batch */
        dust_big = p->dust_w <= uint32_t(kDustFastW) &&
                   dust_big_ok(dust_layout(int(p->dust_w)).stride);
        if (p->dust_w <= uint32_t(kDustFastW) && W.dfall_cap < in->nreads) {
            if (W.dfall) {
                HIP_OK(hipStreamSynchronize(st));
                HIP_OK(hipFree(W.dfall));
            }
            W.dfall = nullptr;
            W.dfall_cap = 0;
            HIP_OK(hipMalloc(&W.dfall, in->nreads * sizeof(uint32_t)));
            W.dfall_cap = in->nreads;
        }
    }
    if (p->path == BBT_PATH_BATCHED && W.lists_cap < in->nreads) {
        /* the lists may still be read by this workspace's previous batch */
        if (W.dlists) {
            HIP_OK(hipStreamSynchronize(st));
            HIP_OK(hipFree(W.dlists));
        }
        W.dlists = nullptr;
        W.lists_cap = 0;
        HIP_OK(hipMalloc(&W.dlists, 2 * in->nreads * sizeof(uint32_t)));
        W.lists_cap = in->nreads;
    }
    std::memset(&P, 0, sizeof(P));
    P.dust_big = dust_big;
    Launch &L = P.L;
    L.st = st;
    L.path = p->path;
    L.cu = fs->cu_count;
    bbt_tuning tune = fs->tune;
    bool from_table = false;
    if (!fs->tuned && !noalloc) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        if (!g_tune_table.empty()) {
            const auto it = g_tune_table.find(tune_key(fs, in, p));
            if (it != g_tune_table.end()) {
                tune = it->second;
                from_table = true;
            }
        }
    }
    /* a set of one geometry without knobs of its own or in the table: the
     * interleaved read-major walks, four lookups at a time, with lookahead
     * (10 x 20 Mbp filters: 95.2 -> 52.4 ms per 4M reads), and one read per
     * lane where that applies to a large batch (24.2 ms); without allocation
     * the same load scheduling over the filters' own arrays.  A
     * set with only part of its filters in a group takes the read-per-lane
     * walk for the group and filter-major walks over their own arrays for
     * the others where the read-per-lane walk applies to a large batch
     * (the serving set, one big filter and eight of one geometry, 516k
     * reads: 57.8 -> 36.8 ms), else filter-major walks for all (57.9 ms
     * against 67.2 ms with the read-major walks) */
    P.wave_min = p->wave_min_len ? p->wave_min_len
                                 : (tune.wave_min_len ? tune.wave_min_len : kDefaultWaveMin);
    if (!fs->tuned && !from_table) {
        const bool rows_fit = !noalloc && in->nreads >= kRowsMinReads &&
                              rows_group(fs, ilv_plan(fs), p, in->max_len, P.wave_min) != 0;
        if (ilv_whole(fs))
            tune.probe_ahead =
                noalloc ? (kDefaultIlvAhead & 3u) : (rows_fit ? kDefaultRowsAhead : kDefaultIlvAhead);
        else if (rows_fit)
            tune.probe_ahead = kDefaultRowsAhead;
    }
    L.blocks_per_cu = tune.blocks_per_cu;
    EvalArgs &a = L.a;
    a.seq = reinterpret_cast<const uint8_t *>(in->seq);
    if (in->format == BBT_READS_2BIT) {
        P.codes = a.seq;
        P.code_bytes = in->seq_bytes;
        a.seq = W.dascii;
    }
    a.off = in->off;
    a.order = in->order;
    a.filters = fs->ddesc;
    if ((tune.probe_ahead & 4u) && p->path == BBT_PATH_BATCHED && p->mode != BBT_ORDERED &&
        !noalloc) {
        bool ok = false;
        bbt_status s2 = ensure_ilv(fs, ok);
        if (s2 != BBT_OK)
            return s2;
        if (ok) {
            P.group = true;
            a.filters = fs->ddesc_ilv;
            a.read_major = 1;
        }
    }
    a.hits = out->hits;
    a.scores = out->scores;
    a.counters = reinterpret_cast<unsigned long long *>(out->counters);
    a.nf = nf;
    a.nw = nf;
    a.fmap = nullptr;
    a.refill_min = tune.refill_min ? tune.refill_min : kDefaultRefillMin;
    a.probe_ahead = tune.probe_ahead & 3u;
    if (fix_scores) {
        P.fix_scores = true;
        a.fix_list = W.dfix;
        a.fix_count = W.dctl + 102;
    }
    a.wp.threshold = p->threshold;
    a.wp.streak = p->streak;
    a.wp.min_len = uint32_t(std::round(p->threshold));
    a.nitems = in->nreads * nf;
    P.nf = nf;
    P.nrec = p->paired ? in->nreads / 2 : in->nreads;
    P.kk = common_k == 25 ? 25u : 0u;
    /* probe_ahead bit 3: the group's filters by the read-per-lane walk (and
     * whole-wave walks of long reads against all of them); the set's other
     * filters, if any, by the lane and wave walks over their own arrays,
     * filter-major, in a second launch */
    P.rows_n = P.group && (tune.probe_ahead & 8u)
                   ? rows_group(fs, fs->ilv, p, in->max_len, P.wave_min)
                   : 0;
    P.rows = P.rows_n != 0;
    if (P.rows)
        a.filters = fs->ddesc;
    return BBT_OK;
}

/* The kernels of one batch, in stream order; capturable (no memset nodes:
 * counters are cleared by a fill kernel). */
static bbt_status enqueue_classify(const bbt_eval_params *p, const bbt_eval_out *out,
                                   const Prepared &P);

static bbt_status enqueue(bbt_filterset *fs, bbt_workspace &W, const bbt_reads *in,
                          const bbt_eval_params *p, const bbt_eval_out *out, Prepared &P,
                          bool timed) {
    Launch &L = P.L;
    EvalArgs &a = L.a;
    hipStream_t st = L.st;
    const uint32_t nf = P.nf;
    bbt_status s = BBT_OK;
    /* work queues of the (up to two) walks of this call, and for the batched
     * path the length buckets: long reads walked by whole waves */
    k_zero_u64<<<1, 64, 0, st>>>(W.dctl, 48);
    if (P.fix_scores)
        k_zero_u64<<<1, 64, 0, st>>>(W.dctl + 102, 1);
    if (P.code_bytes) {
        const uint64_t nw = P.code_bytes / 8;
        k_unpack_2bit<<<uint32_t(std::min<uint64_t>((nw + kBlock - 1) / kBlock, 65536)), kBlock,
                        0, st>>>(P.codes, P.code_bytes, const_cast<uint8_t *>(a.seq));
        HIP_OK(hipGetLastError());
    }
    a.queue = W.dctl + 0;
    a.lqueue = W.dctl + 2;
    a.dust = nullptr;
    if (p->dust) {
        /* -D pre-pass: clear the bitmap, then sdust every read */
        k_fill_u8<<<1024, kBlock, 0, st>>>(reinterpret_cast<uint8_t *>(W.dmask),
                                            W.dmask_words * sizeof(uint32_t), 0);
        /* [104] LDS-pass queue, [105] error sink, [106] scratch-pass queue,
         * [107] reads listed for the scratch pass */
        k_zero_u64<<<1, 64, 0, st>>>(W.dctl + 104, 4);
        unsigned long long *err = out->counters
                                      ? reinterpret_cast<unsigned long long *>(out->counters) + 3
                                      : W.dctl + 105;
        /* the layout the scratch was allocated with (its window >= this one's) */
        const DustLayout lay = dust_layout(int(W.dscratch_w));
        if (p->dust_w <= uint32_t(kDustFastW)) {
            /* LDS pass over every read, then the reads whose interval lists
             * outgrew it: one per workgroup with the whole list in LDS when
             * the device allows it, else on the scratch pass */
            const uint32_t grid = uint32_t(std::min<uint64_t>(
                (in->nreads + 63) / 64, uint64_t(fs->cu_count) * (kLdsPerCu / kDustFastLds)));
            k_dust_lds<<<grid, 64, 0, st>>>(a.seq, in->off, in->nreads, int(p->dust_t),
                                             int(p->dust_w), W.dmask, W.dctl + 104, W.dfall,
                                             W.dctl + 107);
            HIP_OK(hipGetLastError());
            const DustLayout big = dust_layout(int(p->dust_w));
            if (P.dust_big) {
                k_dust_big<<<uint32_t(fs->cu_count), 64, size_t(big.stride), st>>>(
                    a.seq, in->off, int(p->dust_t), int(p->dust_w), big, W.dmask, W.dctl + 106,
                    W.dfall, W.dctl + 107, err);
            } else {
                k_dust<<<(lay.threads + 63) / 64, 64, 0, st>>>(
                    a.seq, in->off, in->nreads, int(p->dust_t), int(p->dust_w), W.dscratch, lay,
                    W.dmask, W.dctl + 106, W.dfall, W.dctl + 107, err);
            }
        } else {
            k_dust<<<(lay.threads + 63) / 64, 64, 0, st>>>(a.seq, in->off, in->nreads,
                                                    int(p->dust_t), int(p->dust_w), W.dscratch,
                                                    lay, W.dmask, W.dctl + 106, nullptr,
                                                    nullptr, err);
        }
        HIP_OK(hipGetLastError());
        a.dust = W.dmask;
    }
    /* the length buckets are needed only when some read may reach the
     * whole-wave length; otherwise every read is a lane walk in input order */
    if (p->path == BBT_PATH_BATCHED && p->mode != BBT_ORDERED && in->max_len >= P.wave_min) {
        const uint32_t gb =
            uint32_t(std::min<uint64_t>(1024, (in->nreads + kBlock - 1) / kBlock));
        k_len_count<<<gb, kBlock, 0, st>>>(in->off, in->nreads, P.wave_min, W.dctl + 8);
        k_len_scan<<<1, 64, 0, st>>>(W.dctl + 8, W.dctl + 48, W.dctl + 4);
        k_len_scatter<<<gb, kBlock, 0, st>>>(in->off, in->nreads, P.wave_min,
                                              W.dctl + 48, W.dctl + 8, W.dlists,
                                              W.dlists + W.lists_cap);
        HIP_OK(hipGetLastError());
        a.long_list = W.dlists;
        a.short_list = W.dlists + W.lists_cap;
        a.list_counts = W.dctl + 4;
        a.bucket_base = W.dctl + 48;
        a.bucket_cnt = W.dctl + 8; /* the scatter cursors end at the counts */
    }
    if (p->mode == BBT_ORDERED)
        k_fill_u8<<<1024, kBlock, 0, st>>>(out->hits, in->nreads * nf, 0);
    if (P.group) {
        /* the interleaved copies, from the filters' own arrays as they are now */
        for (const auto &g : fs->ilv) {
            const uint64_t nbytes = fs->desc[g.mem.f[0]].m / 8;
            k_interleave<<<uint32_t(std::min<uint64_t>((nbytes + kBlock - 1) / kBlock, 65536)),
                           kBlock, 0, st>>>(fs->ddesc, g.mem, nbytes, g.shift, g.buf);
        }
        HIP_OK(hipGetLastError());
    }
    if (timed)
        HIP_OK(hipEventRecord(W.ev0, st));
    switch (p->mode) {
    case BBT_ORDERED:
        s = launch_ordered(L, p->scorer, P.nrec, p->paired, p->inclusive, P.kk);
        break;
    case BBT_BESTHIT:
        s = launch_scorer(L, p->scorer, true, P.kk);
        break;
    case BBT_SCORES:
        s = launch_scorer(L, p->scorer, false, P.kk);
        a.queue = W.dctl + 1;
        a.lqueue = W.dctl + 3;
        if (s == BBT_OK)
            s = launch_scorer(L, p->scorer, true, P.kk);
        break;
    default:
        if (P.rows) {
            Launch Lr = L;
            Lr.a.fmap = fs->dmap;
            Lr.a.nw = P.rows_n;
            Lr.a.nitems = in->nreads * P.rows_n;
            s = launch_rows(Lr, fs->ilv[0].buf, fs->ilv[0].shift, P.kk);
            if (s == BBT_OK && P.rows_n < nf) {
                /* the others: filter-major walks over their own arrays, with
                 * queues of their own */
                Launch Lo = L;
                Lo.a.fmap = fs->dmap + P.rows_n;
                Lo.a.nw = nf - P.rows_n;
                Lo.a.nitems = in->nreads * (nf - P.rows_n);
                Lo.a.read_major = 0;
                Lo.a.queue = W.dctl + 1;
                Lo.a.lqueue = W.dctl + 3;
                s = launch_scorer(Lo, p->scorer, false, P.kk);
            }
        } else {
            s = launch_scorer(L, p->scorer, false, P.kk);
        }
        break;
    }
    if (s != BBT_OK)
        return s;
    if (timed)
        HIP_OK(hipEventRecord(W.ev1, st));
    if (P.fix_scores)
        return BBT_OK; /* eval_impl finishes the scores, then classifies */
    return enqueue_classify(p, out, P);
}

/* hit rows (and best hits from the scores) -> classes and summary */
static bbt_status enqueue_classify(const bbt_eval_params *p, const bbt_eval_out *out,
                                   const Prepared &P) {
    const uint32_t nf = P.nf;
    hipStream_t st = P.L.st;
    if (out->cls || out->summary || p->mode == BBT_BESTHIT) {
        const uint64_t blocks = (P.nrec + kBlock - 1) / kBlock;
        const size_t shm = out->summary ? (2 * size_t(nf) + 2) * sizeof(unsigned int) : 0;
        k_classify<<<dim3(uint32_t(blocks)), kBlock, shm, st>>>(
            out->hits, out->scores, nf, P.nrec, p->paired, p->inclusive,
            p->mode == BBT_BESTHIT, out->cls,
            reinterpret_cast<unsigned long long *>(out->summary));
        HIP_OK(hipGetLastError());
    }
    return BBT_OK;
}

/* identity of a captured batch: every pointer and value the kernels see,
 * the filter set's descriptors included (a workspace may serve several
 * filter sets with the same buffers) */
static std::string graph_key(const bbt_filterset *fs, const bbt_reads *in,
                             const bbt_eval_params *p, const bbt_eval_out *out,
                             const Prepared &P, hipStream_t st) {
    std::string k;
    auto add = [&k](const void *v, size_t n) { k.append(static_cast<const char *>(v), n); };
    add(&fs, sizeof(fs));
    add(&P.L.a.filters, sizeof(P.L.a.filters));
    add(&P.nf, sizeof(P.nf));
    add(&P.kk, sizeof(P.kk));
    add(&P.L.a.seq, sizeof(P.L.a.seq)); /* the staging buffer of 2-bit reads */
    add(in, sizeof(*in));
    add(p, sizeof(*p));
    add(out, sizeof(*out));
    add(&P.wave_min, sizeof(P.wave_min));
    add(&P.L.a.refill_min, sizeof(P.L.a.refill_min));
    add(&P.L.a.probe_ahead, sizeof(P.L.a.probe_ahead));
    add(&P.L.a.read_major, sizeof(P.L.a.read_major));
    add(&P.rows, sizeof(P.rows));
    add(&P.rows_n, sizeof(P.rows_n));
    add(&P.L.blocks_per_cu, sizeof(P.L.blocks_per_cu));
    add(&st, sizeof(st));
    return k;
}

/* BINOMIAL evalScore of the listed items on the host: -10 log10 P(X > count),
 * X ~ Bin(frames, FPR of the item's filter), the function that fills the
 * device table (SeqEval.h:450-491, 558), then written into the rows. */
static bbt_status finish_scores(bbt_filterset *fs, bbt_workspace &W, const bbt_eval_out *out,
                                const Prepared &P, hipStream_t st) {
    unsigned long long n = 0;
    HIP_OK(hipMemcpyAsync(&n, W.dctl + 102, sizeof(n), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (n == 0)
        return BBT_OK;
    if (n > W.fix_cap)
        return fail(BBT_ERR_HIP, "binomial score list overflow");
    std::vector<ScoreFix> fx(n);
    HIP_OK(hipMemcpyAsync(fx.data(), W.dfix, n * sizeof(ScoreFix), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    const uint32_t nf = P.nf;
    const unsigned T = unsigned(std::max<uint64_t>(
        1, std::min<uint64_t>({16, std::max(1u, std::thread::hardware_concurrency()), n / 256})));
    const auto work = [&](unsigned t) {
        for (uint64_t i = t; i < n; i += T) {
            const double fpr = fs->info[fx[i].idx % nf].fpr;
            const double v =
                std::log10(binom_sf(int64_t(fx[i].count), fx[i].frames, fpr)) * -10;
            uint64_t bits;
            std::memcpy(&bits, &v, 8);
            fx[i].frames = uint32_t(bits);
            fx[i].count = uint32_t(bits >> 32);
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; ++t)
        th.emplace_back(work, t);
    work(0);
    for (std::thread &x : th)
        x.join();
    HIP_OK(hipMemcpyAsync(W.dfix, fx.data(), n * sizeof(ScoreFix), hipMemcpyHostToDevice, st));
    k_fix_scores<<<uint32_t(std::min<uint64_t>((n + kBlock - 1) / kBlock, 4096)), kBlock, 0,
                   st>>>(W.dfix, n, out->scores);
    HIP_OK(hipGetLastError());
    /* the host vector is read by the copy: wait before it goes away */
    HIP_OK(hipStreamSynchronize(st));
    return BBT_OK;
}

static bbt_status eval_impl(bbt_filterset *fs, bbt_workspace &W, const bbt_reads *in,
                            const bbt_eval_params *p, bbt_eval_out *out, hipStream_t st,
                            bool noalloc = false) {
    std::lock_guard<std::mutex> g(fs->mu);
    Prepared P;
    bbt_status s = prepare(fs, W, in, p, out, st, P, noalloc);
    if (s != BBT_OK)
        return s;
    if (P.rows) {
        static const char *const names[4] = {
            "k_walk_rows", "k_walk_group_long+k_walk_rows", "k_walk_rows+k_walk_persistent",
            "k_walk_group_long+k_walk_rows+k_walk_persistent"};
        W.last_walk_kernel =
            names[(in->max_len >= P.wave_min ? 1 : 0) + (P.rows_n < P.nf ? 2 : 0)];
    } else {
        W.last_walk_kernel = p->mode == BBT_ORDERED     ? "k_ordered"
                             : p->path == BBT_PATH_PLAIN ? "k_walk_plain"
                                                         : "k_walk_persistent";
    }
    const bool timed = (p->flags & BBT_FLAG_TIME_WALK) != 0;
    if (timed) {
        if (!W.ev0)
            HIP_OK(hipEventCreate(&W.ev0));
        if (!W.ev1)
            HIP_OK(hipEventCreate(&W.ev1));
    }
    const bool graph = (p->flags & BBT_FLAG_GRAPH) != 0 && !timed;
    if (!graph) {
        s = enqueue(fs, W, in, p, out, P, timed);
        if (s != BBT_OK)
            return s;
        if (P.fix_scores) {
            s = finish_scores(fs, W, out, P, st);
            if (s == BBT_OK)
                s = enqueue_classify(p, out, P);
            if (s != BBT_OK)
                return s;
        }
    } else {
        const std::string key = graph_key(fs, in, p, out, P, st);
        if (!W.graph_exec || key != W.graph_key) {
            if (W.graph_exec) {
                (void)hipGraphExecDestroy(W.graph_exec);
                W.graph_exec = nullptr;
            }
            if (!W.cap_stream)
                HIP_OK(hipStreamCreateWithFlags(&W.cap_stream, hipStreamNonBlocking));
            P.L.st = W.cap_stream;
            HIP_OK(hipStreamBeginCapture(W.cap_stream, hipStreamCaptureModeRelaxed));
            s = enqueue(fs, W, in, p, out, P, false);
            P.L.st = st;
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(W.cap_stream, &gr);
            if (s != BBT_OK) {
                if (gr)
                    (void)hipGraphDestroy(gr);
                return s;
            }
            if (e != hipSuccess)
                return fail(BBT_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&W.graph_exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) {
                W.graph_exec = nullptr;
                return fail(BBT_ERR_HIP, std::string("graph instantiate: ") +
                                             hipGetErrorString(e));
            }
            W.graph_key = key;
        }
        HIP_OK(hipGraphLaunch(W.graph_exec, st));
    }
    if (timed) {
        HIP_OK(hipEventSynchronize(W.ev1));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
        W.last_walk_ms = ms;
    }
    return BBT_OK;
}

extern "C" bbt_status bbt_eval_batch(bbt_filterset *fs, const bbt_reads *in,
                                     const bbt_eval_params *p, bbt_eval_out *out,
                                     void *stream) {
    if (!fs || !in || !p || !out)
        return fail(BBT_ERR_ARG, "bbt_eval_batch: null argument");
    if (in->nreads == 0)
        return BBT_OK;
    return eval_impl(fs, fs->ws, in, p, out, static_cast<hipStream_t>(stream));
}

/* ---- bbt_eval_batch_noalloc ---------------------------------------------
 * The control words and length-bucket lists of one batch live in this
 * array instead of a workspace: kNoAllocCtlWords 64-bit words, then the two
 * read lists of kNoAllocMaxReads 32-bit entries each (8 MiB + 1 KiB).  Each
 * GPU holds its own copy from the library's first use there until the
 * process ends.  In host code the name `scratch` is not a device address:
 * its address on a device comes from hipGetSymbolAddress, taken once per
 * device with that device current, and only that address reaches kernels. */
constexpr uint64_t kNoAllocMaxReads = uint64_t(1) << 20;
constexpr uint64_t kNoAllocCtlWords = 128;
static __device__ __align__(256) unsigned long long
    scratch[kNoAllocCtlWords + kNoAllocMaxReads]; /* lists: 2 x kNoAllocMaxReads x 4 B */

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned long long *g_scratch_ptr[kMaxDevices] = {};

/* scratch's address on the current device dev, taken on its first use there */
static bbt_status noalloc_scratch(int dev, unsigned long long **ptr) {
    if (dev < 0 || dev >= kMaxDevices)
        return fail(BBT_ERR_ARG, "bbt_eval_batch_noalloc: device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return fail(BBT_ERR_HIP, "bbt_eval_batch_noalloc: no device address for scratch");
        g_scratch_ptr[dev] = static_cast<unsigned long long *>(p);
    }
    *ptr = g_scratch_ptr[dev];
    return BBT_OK;
}

extern "C" bbt_status bbt_eval_batch_noalloc(bbt_filterset *fs, const bbt_reads *in,
                                             const bbt_eval_params *p, bbt_eval_out *out,
                                             void *stream) {
    if (!fs || !in || !p || !out)
        return fail(BBT_ERR_ARG, "bbt_eval_batch_noalloc: null argument");
    if (in->nreads > kNoAllocMaxReads)
        return fail(BBT_ERR_ARG, "bbt_eval_batch_noalloc: at most " +
                                     std::to_string(kNoAllocMaxReads) +
                                     " reads per call (the size of its scratch); split the "
                                     "batch or use bbt_eval_batch");
    if (in->nreads == 0)
        return BBT_OK;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev != fs->device)
        return fail(BBT_ERR_ARG,
                    "bbt_eval_batch_noalloc: the filter set belongs to another device");
    unsigned long long *base = nullptr;
    bbt_status s = noalloc_scratch(dev, &base);
    if (s != BBT_OK)
        return s;
    bbt_workspace W; /* views into scratch; nothing of it is freed */
    W.dctl = base;
    W.dlists = reinterpret_cast<uint32_t *>(base + kNoAllocCtlWords);
    W.lists_cap = kNoAllocMaxReads;
    return eval_impl(fs, W, in, p, out, static_cast<hipStream_t>(stream), true);
}

extern "C" bbt_status bbt_workspace_create(bbt_workspace **out) {
    if (!out)
        return fail(BBT_ERR_ARG, "bbt_workspace_create: null argument");
    *out = new bbt_workspace();
    return BBT_OK;
}

extern "C" bbt_status bbt_workspace_free(bbt_workspace *ws) {
    if (!ws)
        return BBT_OK;
    workspace_release(ws);
    delete ws;
    return BBT_OK;
}

extern "C" bbt_status bbt_eval_batch_ws(bbt_filterset *fs, bbt_workspace *ws,
                                        const bbt_reads *in, const bbt_eval_params *p,
                                        bbt_eval_out *out, void *stream) {
    if (!fs || !ws || !in || !p || !out)
        return fail(BBT_ERR_ARG, "bbt_eval_batch_ws: null argument");
    if (in->nreads == 0)
        return BBT_OK;
    return eval_impl(fs, *ws, in, p, out, static_cast<hipStream_t>(stream));
}

extern "C" bbt_status bbt_measure_gather(uint64_t nbytes, uint32_t rounds, void *stream,
                                         double *lines_per_s, double *ms_out) {
    if (nbytes < 4096 || (nbytes & (nbytes - 1)) || !rounds || !lines_per_s)
        return fail(BBT_ERR_ARG, "bbt_measure_gather: nbytes must be a power of two >= 4096");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = 0, cu = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    uint8_t *buf = nullptr;
    unsigned long long *sink = nullptr;
    if (hipMalloc(&buf, nbytes) != hipSuccess)
        return fail(BBT_ERR_NOMEM, "bbt_measure_gather: allocation failed");
    if (hipMalloc(&sink, 8) != hipSuccess) {
        (void)hipFree(buf);
        return fail(BBT_ERR_NOMEM, "bbt_measure_gather: allocation failed");
    }
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const uint32_t blocks = uint32_t(cu) * 8;
    k_fill_u8<<<4096, kBlock, 0, st>>>(buf, nbytes, 1);
    k_rand_gather<<<blocks, kBlock, 0, st>>>(buf, nbytes - 1, rounds / 8 + 1, sink); /* warm */
    HIP_OK(hipEventRecord(e0, st));
    k_rand_gather<<<blocks, kBlock, 0, st>>>(buf, nbytes - 1, rounds, sink);
    HIP_OK(hipEventRecord(e1, st));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    (void)hipFree(sink);
    const double loads = double(blocks) * kBlock * double(rounds) * 8;
    *lines_per_s = loads / (double(ms) * 1e-3);
    if (ms_out)
        *ms_out = ms;
    return BBT_OK;
}

extern "C" bbt_status bbt_set_tuning(bbt_filterset *fs, const bbt_tuning *t) {
    if (!fs || !t)
        return fail(BBT_ERR_ARG, "bbt_set_tuning: null argument");
    if (t->probe_ahead > 15)
        return fail(BBT_ERR_ARG, "bbt_set_tuning: probe_ahead takes four bits (0..15)");
    std::lock_guard<std::mutex> g(fs->mu);
    fs->tune = *t;
    fs->tuned = true;
    return BBT_OK;
}

extern "C" bbt_status bbt_get_tuning(const bbt_filterset *fs, bbt_tuning *t) {
    if (!fs || !t)
        return fail(BBT_ERR_ARG, "bbt_get_tuning: null argument");
    *t = fs->tune;
    return BBT_OK;
}

/* Coordinate search over the knobs, each candidate timed on the full batch
 * by the walk-kernel events. */
extern "C" bbt_status bbt_autotune(bbt_filterset *fs, const bbt_reads *in,
                                   const bbt_eval_params *p, bbt_eval_out *out,
                                   void *stream, bbt_tuning *best) {
    if (!fs || !in || !p || !out)
        return fail(BBT_ERR_ARG, "bbt_autotune: null argument");
    bbt_eval_params q = *p;
    q.path = BBT_PATH_BATCHED;
    q.flags = BBT_FLAG_TIME_WALK;
    q.wave_min_len = 0;
    bbt_tuning cur = fs->tune;
    auto trial = [&](const bbt_tuning &t, double &ms) -> bbt_status {
        bbt_status s = bbt_set_tuning(fs, &t);
        if (s == BBT_OK)
            s = bbt_eval_batch(fs, in, &q, out, stream);
        ms = fs->ws.last_walk_ms;
        return s;
    };
    double best_ms = 0;
    bbt_status s = trial(cur, best_ms);
    if (s != BBT_OK)
        return s;
    const uint32_t waves[] = {64, 128, 256, 512, 1024, 2048, 4096, 0x7FFFFFFFu};
    const uint32_t refills[] = {1, 4, 8, 16, 32};
    const uint32_t bpcs[] = {0, 1, 2, 3, 4};
    const uint32_t aheads[] = {0, 1, 2, 3, 4, 5, 6, 7, 12};
    /* the interleaved forms only for a set with a group of one geometry; the
     * read-per-lane walk (12) where it applies */
    const auto plan = ilv_plan(fs);
    const bool rows_ok = plan.size() == 1 && plan[0].mem.n >= 3 && plan[0].mem.n <= 16 &&
                         p->scorer == BBT_SIMPLE && p->mode == BBT_STD;
    const int nahead =
        ilv_any(fs) && p->mode != BBT_ORDERED ? (rows_ok ? 9 : 8) : 4;
    for (int dim = 0; dim < 4; ++dim) {
        const uint32_t *vals =
            dim == 0 ? waves : (dim == 1 ? refills : (dim == 2 ? bpcs : aheads));
        const int nv = dim == 0 ? 8 : (dim == 3 ? nahead : 5);
        for (int i = 0; i < nv; ++i) {
            bbt_tuning t = cur;
            if (dim == 0) {
                if (vals[i] != 0x7FFFFFFFu && vals[i] > in->max_len)
                    continue; /* no read would change path */
                t.wave_min_len = vals[i];
            } else if (dim == 1) {
                t.refill_min = vals[i];
            } else if (dim == 2) {
                t.blocks_per_cu = vals[i];
            } else {
                t.probe_ahead = vals[i];
            }
            double ms = 0;
            s = trial(t, ms);
            if (s != BBT_OK)
                return s;
            if (ms < best_ms) {
                best_ms = ms;
                cur = t;
            }
        }
    }
    s = bbt_set_tuning(fs, &cur);
    if (best)
        *best = cur;
    fs->ws.last_walk_ms = best_ms;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_table[tune_key(fs, in, p)] = cur;
    }
    return s;
}

extern "C" bbt_status bbt_tuning_lookup(const bbt_filterset *fs, const bbt_reads *in,
                                        const bbt_eval_params *p, bbt_tuning *t) {
    if (!fs || !in || !p || !t)
        return fail(BBT_ERR_ARG, "bbt_tuning_lookup: null argument");
    std::lock_guard<std::mutex> g(g_tune_mu);
    const auto it = g_tune_table.find(tune_key(fs, in, p));
    if (it == g_tune_table.end())
        return fail(BBT_ERR_ARG, "bbt_tuning_lookup: no entry for this geometry");
    *t = it->second;
    return BBT_OK;
}

/* one line per entry: the geometry, a tab, then the four knobs */
extern "C" bbt_status bbt_tuning_save(const char *path) {
    if (!path)
        return fail(BBT_ERR_ARG, "bbt_tuning_save: null path");
    std::string text;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        for (const auto &kv : g_tune_table) {
            char buf[96];
            std::snprintf(buf, sizeof(buf), "\t%u %u %u %u\n", kv.second.wave_min_len,
                          kv.second.refill_min, kv.second.blocks_per_cu,
                          kv.second.probe_ahead);
            text += kv.first + buf;
        }
    }
    FILE *f = std::fopen(path, "w");
    if (!f)
        return fail(BBT_ERR_IO, std::string("bbt_tuning_save: cannot write ") + path);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (std::fclose(f) != 0 || !ok)
        return fail(BBT_ERR_IO, std::string("bbt_tuning_save: cannot write ") + path);
    return BBT_OK;
}

extern "C" bbt_status bbt_tuning_load(const char *path) {
    if (!path)
        return fail(BBT_ERR_ARG, "bbt_tuning_load: null path");
    FILE *f = std::fopen(path, "r");
    if (!f)
        return fail(BBT_ERR_IO, std::string("bbt_tuning_load: cannot read ") + path);
    std::map<std::string, bbt_tuning> got;
    char line[512];
    unsigned lineno = 0;
    bbt_status st = BBT_OK;
    while (std::fgets(line, sizeof(line), f)) {
        ++lineno;
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r'))
            l.pop_back();
        if (l.empty())
            continue;
        const size_t tab = l.find('\t');
        bbt_tuning t = {0, 0, 0, 0};
        char extra = 0;
        if (tab == std::string::npos || tab == 0 ||
            std::sscanf(l.c_str() + tab + 1, "%u %u %u %u %c", &t.wave_min_len, &t.refill_min,
                        &t.blocks_per_cu, &t.probe_ahead, &extra) != 4 ||
            t.probe_ahead > 15) {
            st = fail(BBT_ERR_FORMAT, std::string("bbt_tuning_load: malformed line ") +
                                          std::to_string(lineno) + " of " + path);
            break;
        }
        got[l.substr(0, tab)] = t;
    }
    std::fclose(f);
    if (st != BBT_OK)
        return st;
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (const auto &kv : got)
        g_tune_table[kv.first] = kv.second;
    return BBT_OK;
}

extern "C" double bbt_last_walk_ms(const bbt_filterset *fs) {
    return fs ? fs->ws.last_walk_ms : -1.0;
}

extern "C" const char *bbt_last_walk_kernel(const bbt_filterset *fs) {
    return fs ? fs->ws.last_walk_kernel : "";
}
