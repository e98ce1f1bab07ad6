/*
 * bbt_mibf.hip -- multi-index Bloom filter (MIBF) on gfx950: build and query.
 *
 * Reference structure (BioBloom Tools 2.3.4; the MIBloomFilter classes are in
 * the absent btl_bloomfilter submodule):
 *   build  BioBloomMaker/MIBFGen.hpp:131-309 -- bit-vector pass, ID pass
 *          (random ownership), saturation pass ("if target frame does not
 *          have a single representative, mark frame as saturated")
 *   query  BioBloomCategorizer/MIBFClassifier.hpp:773-796 (classify), 79-94
 *          (allowed misses, frame probabilities), 930-939 (getMinCount)
 * The exact definitions restated here are listed in the oracle's header
 * (ownership by the largest (mix32(h ^ id*C), id), saturation per frame,
 * per-frame ID counting, ranking by nonSat/count/solid/totalNonSat/total/id).
 *
 * Device layout of the bit vector: 64-byte lines of one 64-bit rank sample
 * (set bits before the line) and seven 64-bit words, so a bit test and its
 * rank1 come from one line; the ID array is rank-ordered, 32-bit
 * (bit 31 = saturated).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bbt_hip.h"
#include "bbt_device.h"
#include "bbt_host.h"
#include "bbt_reads.h"

using namespace bbt;

namespace {

constexpr uint32_t kSatBit = 0x80000000u;
constexpr uint32_t kIdMask = 0x7FFFFFFFu;
constexpr uint32_t kMaxSeeds = 8;
constexpr uint32_t kWordsPerLine = 7;

bbt_status mfail(bbt_status c, const std::string &m) { return set_error(c, m); }

#define MHIP(expr)                                                                   \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return mfail(BBT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

/* Uniform description of a MIBF for the kernels. */
struct MibfDev {
    const uint64_t *lines; /* nlines * 8 words */
    const uint32_t *data;
    uint64_t m;
    uint64_t magic;
    uint32_t sh1, sh2;
    uint32_t S, span;
    uint32_t nids;
    uint32_t kmer;            /* 1: k-mer mode, S ntHash values per k-mer */
    uint32_t care[kMaxSeeds]; /* spaced seeds: bit j set = position j is used */
    uint32_t dc_union;        /* positions that are don't-care in some seed */
    uint32_t pad;
    /* direct-mapped table, or null: one 32-bit word per bit position, 0 for a
     * clear bit, else the slot's ID word (ID 0 stored as the all-ones ID) */
    const uint32_t *direct;
};

__device__ __forceinline__ uint32_t mix32(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return uint32_t(x >> 32);
}

/* bit p and rank1(p) from the interleaved lines */
__device__ __forceinline__ bool bv_rank(const uint64_t *lines, uint64_t p, uint64_t &rank) {
    const uint64_t w = p >> 6;
    const uint64_t line = w / kWordsPerLine;
    const uint32_t slot = uint32_t(w - line * kWordsPerLine);
    const uint64_t *L = lines + line * 8;
    uint64_t r = L[0];
    for (uint32_t s = 0; s < slot; ++s)
        r += uint64_t(__popcll(L[1 + s]));
    const uint64_t word = L[1 + slot];
    const uint32_t b = uint32_t(p & 63);
    rank = r + uint64_t(__popcll(b ? word & (~uint64_t(0) >> (64 - b)) : 0));
    return (word >> b) & 1;
}

/* The same from the whole 64-byte line, loaded as four independent 16-byte
 * loads with every word selected by constant index (no loop on the slot):
 * the loads of several lookups can all be in flight together. */
__device__ __forceinline__ bool bv_rank_line(const uint64_t *lines, uint64_t p, uint64_t &rank) {
    const uint64_t w = p >> 6;
    const uint64_t line = w / kWordsPerLine;
    const uint32_t slot = uint32_t(w - line * kWordsPerLine);
    const ulonglong2 *q = reinterpret_cast<const ulonglong2 *>(lines + line * 8);
    const ulonglong2 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const uint64_t W[8] = {q0.x, q0.y, q1.x, q1.y, q2.x, q2.y, q3.x, q3.y};
    uint64_t r = W[0], word = 0;
#pragma unroll
    for (uint32_t t = 0; t < kWordsPerLine; ++t) {
        r += t < slot ? uint64_t(__popcll(W[1 + t])) : 0;
        word = t == slot ? W[1 + t] : word;
    }
    const uint32_t b = uint32_t(p & 63);
    rank = r + uint64_t(__popcll(b ? word & (~uint64_t(0) >> (64 - b)) : 0));
    return (word >> b) & 1;
}

/* Hash sd of the window of span k starting at p, from the full window values
 * (fwd, rc).  Spaced seeds: the don't-care bases XORed back out, canonical
 * min.  k-mer mode: the ntHash multi-hash sd of the canonical value. */
__device__ __forceinline__ uint64_t frame_hash(const MibfDev &d, const uint8_t *s, uint64_t p,
                                               uint32_t sd, uint64_t fwd, uint64_t rc) {
    const uint32_t k = d.span;
    if (d.kmer)
        return multi_hash(rc < fwd ? rc : fwd, sd, uint64_t(k) * kMultiSeed);
    const uint32_t care = d.care[sd];
    for (uint32_t j = 0; j < k; ++j) {
        if ((care >> j) & 1)
            continue;
        const unsigned code = base_code(s[p + j]);
        fwd ^= rotl64(seed_fwd(code), k - 1 - j);
        rc ^= rotl64(seed_rc(code), j);
    }
    return rc < fwd ? rc : fwd;
}

/* ---------------------------------------------------------------- build */

/* hashes of every N-free window of records (separated by non-ACGT bytes):
 * the callback gets (window start, seed index, hash) */
struct FrameCursor {
    uint64_t fwd = 0, rc = 0, win = 0;
    uint32_t run = 0;
};

/* Thread t owns window starts [t*chunk, (t+1)*chunk) of the whole buffer. */
template <typename Fn>
__device__ void for_frames(const uint8_t *seq, uint64_t nbytes, const MibfDev &d,
                           uint32_t chunk, Fn fn) {
    const uint32_t k = d.span, S = d.S;
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t p0 = t * chunk;
    if (p0 >= nbytes)
        return;
    const uint64_t last = std::min<uint64_t>(p0 + chunk + k - 1, nbytes);
    FrameCursor c;
    for (uint64_t i = p0; i < last; ++i) {
        const unsigned ch = seq[i];
        if (!base_valid(ch)) {
            c.run = 0;
            c.fwd = c.rc = 0;
            continue;
        }
        const unsigned code = base_code(ch);
        if (c.run >= k) {
            const unsigned out = unsigned(c.win >> (2 * (k - 1))) & 3u;
            c.fwd = rotl1(c.fwd) ^ rotl64(seed_fwd(out), k) ^ seed_fwd(code);
            c.rc = rotr1(c.rc ^ seed_rc(out) ^ rotl64(seed_rc(code), k));
        } else {
            c.fwd = rotl1(c.fwd) ^ seed_fwd(code);
            c.rc = rotr1(c.rc ^ rotl64(seed_rc(code), k));
            ++c.run;
        }
        c.win = (c.win << 2) | code;
        if (c.run < k)
            continue;
        const uint64_t p = i + 1 - k;
        for (uint32_t sd = 0; sd < S; ++sd)
            fn(p, sd, frame_hash(d, seq, p, sd, c.fwd, c.rc));
    }
}

__global__ __launch_bounds__(256) void k_mibf_bits(const uint8_t *seq, uint64_t nbytes,
                                                   MibfDev d, uint64_t *bits, uint32_t chunk) {
    for_frames(seq, nbytes, d, chunk,
               [&](uint64_t, uint32_t, uint64_t h) {
                   const uint64_t p = fastmod(h, d.m, d.magic, d.sh1, d.sh2);
                   atomicOr(reinterpret_cast<unsigned long long *>(bits) + (p >> 6),
                            1ull << (p & 63));
               });
}

/* lines: copy the words, leave the per-line popcount in word 0 for the scan */
__global__ void k_mibf_lines(const uint64_t *bits, uint64_t nwords, uint64_t nlines,
                             uint64_t *lines) {
    for (uint64_t l = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; l < nlines;
         l += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t pc = 0;
        for (uint32_t s = 0; s < kWordsPerLine; ++s) {
            const uint64_t w = l * kWordsPerLine + s;
            const uint64_t v = w < nwords ? bits[w] : 0;
            lines[l * 8 + 1 + s] = v;
            pc += uint64_t(__popcll(v));
        }
        lines[l * 8] = pc;
    }
}

/* exclusive scan of the per-line counts in place (one block, chunked) */
__global__ __launch_bounds__(1024) void k_mibf_scan(uint64_t *lines, uint64_t nlines,
                                                    unsigned long long *total) {
    __shared__ unsigned long long part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (nlines + 1023) / 1024;
    const uint64_t a = std::min<uint64_t>(nlines, t * per), b = std::min<uint64_t>(nlines, a + per);
    unsigned long long s = 0;
    for (uint64_t l = a; l < b; ++l)
        s += lines[l * 8];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        unsigned long long run = 0;
        for (uint32_t i = 0; i < 1024; ++i) {
            const unsigned long long v = part[i];
            part[i] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    unsigned long long run = part[t];
    for (uint64_t l = a; l < b; ++l) {
        const unsigned long long v = lines[l * 8];
        lines[l * 8] = run;
        run += v;
    }
}

/* Records are separated by one non-ACGT byte; rec_of maps a window start to
 * its record (binary search in the record start offsets). */
__device__ __forceinline__ uint32_t record_of(const uint64_t *rec_off, uint64_t nrec,
                                              uint64_t p) {
    uint64_t lo = 0, hi = nrec;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (rec_off[mid] <= p)
            lo = mid;
        else
            hi = mid;
    }
    return uint32_t(lo);
}

__global__ __launch_bounds__(256) void k_mibf_own(const uint8_t *seq, uint64_t nbytes,
                                                  const uint64_t *rec_off, const uint32_t *rec_id,
                                                  uint64_t nrec, MibfDev d,
                                                  unsigned long long *best, uint32_t chunk) {
    for_frames(seq, nbytes, d, chunk,
               [&](uint64_t p, uint32_t, uint64_t h) {
                   const uint64_t id = rec_id[record_of(rec_off, nrec, p)];
                   uint64_t rank;
                   bv_rank(d.lines, fastmod(h, d.m, d.magic, d.sh1, d.sh2), rank);
                   const unsigned long long key =
                       (uint64_t(mix32(h ^ (id * 0x9E3779B97F4A7C15ULL))) << 32) | id;
                   atomicMax(best + rank, key);
               });
}

__global__ void k_mibf_settle(const unsigned long long *best, uint64_t n, uint32_t *data) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        data[i] = uint32_t(best[i]) & kIdMask;
}

/* saturation: a frame whose seed slots all hold other IDs marks them all */
__global__ __launch_bounds__(256) void k_mibf_sat(const uint8_t *seq, uint64_t nbytes,
                                                  const uint64_t *rec_off, const uint32_t *rec_id,
                                                  uint64_t nrec, MibfDev d, uint32_t *data,
                                                  uint32_t chunk) {
    uint64_t pos_cur = ~uint64_t(0);
    uint64_t ranks[kMaxSeeds];
    bool rep = false;
    uint32_t nseen = 0;
    for_frames(seq, nbytes, d, chunk,
               [&](uint64_t p, uint32_t sd, uint64_t h) {
                   if (p != pos_cur) {
                       pos_cur = p;
                       rep = false;
                       nseen = 0;
                   }
                   const uint32_t id = rec_id[record_of(rec_off, nrec, p)];
                   uint64_t rank;
                   bv_rank(d.lines, fastmod(h, d.m, d.magic, d.sh1, d.sh2), rank);
                   ranks[nseen++] = rank;
                   if ((data[rank] & kIdMask) == id)
                       rep = true;
                   if (sd + 1 == d.S && !rep)
                       for (uint32_t j = 0; j < nseen; ++j)
                           atomicOr(data + ranks[j], kSatBit);
               });
}

/* The direct-mapped table from the lines and the rank-ordered ID array:
 * thread w fills the 64 words of bit-vector word w (rank of its first set
 * bit from the line's sample and the words before it in the line). */
__global__ __launch_bounds__(256) void k_mibf_direct(const uint64_t *lines, uint64_t nwords,
                                                     const uint32_t *data, uint32_t *direct) {
    for (uint64_t w = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; w < nwords;
         w += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t line = w / kWordsPerLine;
        const uint32_t slot = uint32_t(w - line * kWordsPerLine);
        const uint64_t *L = lines + line * 8;
        uint64_t r = L[0];
        for (uint32_t t = 0; t < slot; ++t)
            r += uint64_t(__popcll(L[1 + t]));
        const uint64_t word = L[1 + slot];
        uint4 *out = reinterpret_cast<uint4 *>(direct + w * 64);
        for (uint32_t q = 0; q < 16; ++q) {
            uint32_t v[4];
            for (uint32_t e = 0; e < 4; ++e) {
                v[e] = 0;
                if ((word >> (4 * q + e)) & 1) {
                    uint32_t x = data[r++];
                    if ((x & kIdMask) == 0)
                        x |= kIdMask;
                    v[e] = x;
                }
            }
            out[q] = make_uint4(v[0], v[1], v[2], v[3]);
        }
    }
}

/* slots per ID (hist[id]) and saturated slots (hist[nids + 1]); with few
 * IDs each block counts in LDS first (thousands of slots share one ID) */
constexpr uint32_t kHistLds = 4096;

__global__ __launch_bounds__(256) void k_mibf_hist(const uint32_t *data, uint64_t n,
                                                   uint32_t nids, unsigned long long *hist) {
    __shared__ unsigned int local[kHistLds];
    const bool lds = nids + 2 <= kHistLds;
    if (lds) {
        for (uint32_t i = threadIdx.x; i < nids + 2; i += blockDim.x)
            local[i] = 0;
        __syncthreads();
    }
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t v = data[i];
        const uint32_t id = v & kIdMask;
        if (lds) {
            if (id <= nids)
                atomicAdd(local + id, 1u);
            if (v & kSatBit)
                atomicAdd(local + nids + 1, 1u);
        } else {
            if (id <= nids)
                atomicAdd(hist + id, 1ull);
            if (v & kSatBit)
                atomicAdd(hist + nids + 1, 1ull);
        }
    }
    if (lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nids + 2; i += blockDim.x)
            if (local[i])
                atomicAdd(hist + i, (unsigned long long)local[i]);
    }
}

/* every record must end with a byte that is not A/C/G/T (any case), so that
 * no window spans two records; err[0] counts the records that do not */
__global__ void k_mibf_check_sep(const uint8_t *seq, const uint64_t *rec_off, uint64_t nrec,
                                 unsigned long long *err) {
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrec;
         r += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t a = rec_off[r], b = rec_off[r + 1];
        if (b <= a || base_valid(seq[b - 1]))
            atomicAdd(err, 1ull);
    }
}

/* ---------------------------------------------------------------- query */

struct QueryArgs {
    MibfDev d;
    const uint8_t *seq;
    const uint64_t *off;
    uint64_t nrec;             /* reads, or pairs (reads 2j, 2j+1) */
    const uint32_t *min_count; /* [frameCount][nids + 1] */
    uint32_t max_frames;       /* rows of min_count - 1 */
    uint32_t allowed_miss;
    double multi_thresh;
    uint32_t min_nonsat;
    uint32_t max_hits;
    uint32_t paired, best_agree;
    uint32_t *nhits;  /* per record */
    uint32_t *hits;   /* per record max_hits x kHitWords */
    uint32_t *sat;    /* per record */
    unsigned long long *counters; /* bv lookups, id gathers, overflow, wide records */
    unsigned long long *queue;
    /* records with more than kMaxCand candidates: the first pass lists them
     * (wide[0] = how many, wide_list the record indices) and the wide pass
     * evaluates them again with the candidates past kMaxCand in spill
     * (spill_stride candidates per workgroup) */
    unsigned long long *wide;
    unsigned long long *wide_queue;
    uint32_t *wide_list;
    uint32_t *spill;
    uint32_t spill_stride;
    /* CAP = 0: each workgroup's frame records in device memory, grec_cap
     * (a power of two) of them per workgroup */
    uint64_t *grec;
    uint32_t grec_cap;
    uint32_t grec_blocks; /* workgroups the area has room for */
};

constexpr uint32_t kHitWords = 7; /* id count nonSat total totalNonSat solid nonSatFrame */
constexpr uint32_t kMaxCand = 64;
constexpr uint32_t kWideBlocks = 64; /* workgroups of the wide pass */
/* candidates a query may keep per record: a record has at most one per
 * (frame, seed) slot, and at most 8192 of those reach the sorted path */
constexpr uint32_t kMaxHitsOut = 8192;
/* up to this many IDs (with the repeat ID 0), per-ID totals are accumulated
 * directly in LDS, six 32-bit counters per ID, instead of sorting the frame
 * records; such a filter takes records of any length */
constexpr uint32_t kDirectIds = 1024;
constexpr uint32_t kDirectWords = 6; /* count nonSat total totalNonSat solid nonSatFrame */

/* one distinct ID of one frame, sortable by ID:
 * id << 32 | clean frame << 16 | slots holding it << 8 | unsaturated ones */
__device__ __forceinline__ uint64_t make_rec(uint32_t id, uint32_t clean, uint32_t occ,
                                             uint32_t ns) {
    return (uint64_t(id) << 32) | (clean << 16) | (occ << 8) | ns;
}

/* rank order of two candidates (MIBFQuerySupport ranking as restated in the
 * oracle): nonSat, count, solid, totalNonSat, total descending, then ID */
__device__ __forceinline__ bool ranks_before(const uint32_t *x, const uint32_t *y) {
    if (x[2] != y[2]) return x[2] > y[2];
    if (x[1] != y[1]) return x[1] > y[1];
    if (x[5] != y[5]) return x[5] > y[5];
    if (x[4] != y[4]) return x[4] > y[4];
    if (x[3] != y[3]) return x[3] > y[3];
    return x[0] < y[0];
}

/* Bases 64c .. 64c + 63 of a sequence, one per lane, in prefix form: H and G
 * (XOR prefixes of the position-rotated seeds through this base, as in
 * bbt_device.h), N (the last position at or before this base that is not
 * A/C/G/T, or past the sequence; -1 if none) and W (2-bit codes of this base
 * and the 31 before it, this base in bits 0-1). */
struct BaseChunk {
    uint64_t H, G, W;
    int N;
};

/* the same values for the last base scanned, carried into the next chunk */
struct ScanCarry {
    uint64_t H, G, W;
    int N;
};

__device__ __forceinline__ BaseChunk scan_chunk(unsigned c, uint32_t i, uint32_t len,
                                                uint32_t lane, bool spaced, ScanCarry &cy) {
    const bool ok = i < len && base_valid(c);
    const unsigned code = base_code(c);
    BaseChunk x;
    x.H = ok ? pfx_fwd(code, i) : 0;
    x.G = ok ? pfx_rc(code, i) : 0;
    x.N = ok ? -1 : int(i);
    for (int dd = 1; dd < 64; dd <<= 1) {
        const uint64_t th = __shfl_up(x.H, unsigned(dd));
        const uint64_t tg = __shfl_up(x.G, unsigned(dd));
        const int tl = __shfl_up(x.N, unsigned(dd));
        if (int(lane) >= dd) {
            x.H ^= th;
            x.G ^= tg;
            x.N = x.N > tl ? x.N : tl;
        }
    }
    x.H ^= cy.H;
    x.G ^= cy.G;
    x.N = x.N > cy.N ? x.N : cy.N;
    cy.H = __shfl(x.H, 63);
    cy.G = __shfl(x.G, 63);
    cy.N = __shfl(x.N, 63);
    x.W = 0;
    if (spaced) {
        /* doubling steps over the lanes, then the previous chunk's bases
         * above those of this chunk */
        uint64_t w = code;
#pragma unroll
        for (uint32_t dd = 1; dd < 32; dd <<= 1) {
            const uint64_t t = __shfl_up(w, dd);
            w |= (lane >= dd ? t : 0) << (2 * dd);
        }
        if (lane < 31)
            w |= cy.W << (2 * (lane + 1));
        cy.W = __shfl(w, 63);
        x.W = w;
    }
    return x;
}

/* Frame p = f0 + lane (the window [p, p + k - 1], k <= 64) from the base
 * chunk X starting at f0, the next chunk Y and the prefixes (hx, gx) through
 * base f0 - 1: the full window values, the packed codes of its bases (W at
 * its last base) and whether it lies inside the sequence's nfr frames
 * without an N. */
struct Frame {
    uint64_t fwd, rc, W;
    bool valid;
};

__device__ __forceinline__ Frame frame_at(const BaseChunk &X, const BaseChunk &Y, uint64_t hx,
                                          uint64_t gx, uint32_t f0, uint32_t lane, uint32_t k,
                                          uint32_t nfr, bool spaced) {
    uint64_t Hp = __shfl_up(X.H, 1u), Gp = __shfl_up(X.G, 1u);
    if (lane == 0) {
        Hp = hx;
        Gp = gx;
    }
    const uint32_t t = lane + k - 1;
    const int src = int(t & 63);
    const bool inX = t < 64;
    const uint64_t hX = __shfl(X.H, src), hY = __shfl(Y.H, src);
    const uint64_t gX = __shfl(X.G, src), gY = __shfl(Y.G, src);
    const int nX = __shfl(X.N, src), nY = __shfl(Y.N, src);
    const uint32_t p = f0 + lane;
    Frame fr;
    fr.valid = p < nfr && (inX ? nX : nY) < int(p);
    fr.fwd = pfx_win_fwd(inX ? hX : hY, Hp, p + k - 1);
    fr.rc = pfx_win_rc(inX ? gX : gY, Gp, p);
    fr.W = 0;
    if (spaced) {
        const uint64_t wX = __shfl(X.W, src), wY = __shfl(Y.W, src);
        fr.W = inX ? wX : wY;
    }
    return fr;
}

/* records claimed per queue atomic (first pass): their offsets come in one
 * load, and each sequence's first bases are loaded while the one before it
 * waits for its lookups */
constexpr uint32_t kRecBatch = 8;

/* One wavefront per record.  Each sequence is taken 64 frames per chunk
 * (lane = frame): the window values from prefix-XOR scans of 64-base chunks,
 * the S seed hashes per frame, then the slots -- with the direct-mapped
 * table (TABLE) one load per seed, two chunks' loads in flight together;
 * with the lines, bit + rank from one line, then the ID.  Each frame's
 * distinct IDs either add to per-ID totals in LDS (DIRECT: up to kDirectIds
 * IDs) or are appended to a record store (LDS, or device memory with
 * CAP = 0) that is sorted by ID and reduced per ID at the end of the record.
 * A record with more than kMaxCand candidates is listed for the WIDE pass,
 * which walks the listed records again with room for every ID as a
 * candidate and ranks them by selection; the WIDE pass adds nothing to the
 * lookup counters (the first pass counted the record's lookups). */
template <uint32_t CAP, uint32_t SMAX, bool DIRECT, bool WIDE, bool TABLE>
__global__ __launch_bounds__(64) void k_mibf_query(QueryArgs a) {
    /* frame chunks whose lookups are in flight together */
    constexpr uint32_t FG = TABLE ? 2 : 1;
    __shared__ uint64_t rec_lds[DIRECT || CAP == 0 ? 1 : CAP];
    /* the record store: LDS, or with CAP = 0 this workgroup's device area */
    uint64_t *const rec = CAP ? rec_lds : a.grec + uint64_t(blockIdx.x) * a.grec_cap;
    const uint32_t cap = CAP ? CAP : a.grec_cap;
    __shared__ uint32_t cand[kMaxCand * kHitWords];
    __shared__ uint32_t ncand_s;
    /* spaced seeds: for window offset j and base code c, the contributions
     * the don't-care base c at j made to the full window hashes, which the
     * seed hash XORs back out: {rotl(seed_fwd(c), k-1-j), rotl(seed_rc(c), j)} */
    __shared__ ulonglong2 dct[32 * 4];
    __shared__ uint8_t order_s[kMaxCand]; /* lane 0's candidate ranking */
    /* DIRECT: per ID, count, nonSat, total, totalNonSat, solid, nonSatFrame;
     * then the minimum counts of the current frame count, per ID */
    extern __shared__ uint32_t acc[];
    const uint32_t lane = threadIdx.x;
    const uint32_t S = a.d.S, k = a.d.span;
    const uint32_t nid1 = a.d.nids + 1;
    const bool spaced = !a.d.kmer;
    uint32_t *const mcs = acc + kDirectWords * nid1;
    uint32_t mc_row = 0xFFFFFFFFu; /* the frame count whose minimum counts mcs holds */
    unsigned long long bvl = 0, gat = 0;
    if (DIRECT)
        for (uint32_t x = lane; x < kDirectWords * nid1; x += 64)
            acc[x] = 0;
    for (uint32_t x = lane; x < 32 * 4; x += 64) {
        const uint32_t j = x >> 2, c = x & 3;
        dct[x] = j < k ? make_ulonglong2(rotl64(seed_fwd(c), k - 1 - j), rotl64(seed_rc(c), j))
                       : make_ulonglong2(0, 0);
    }
    __syncthreads();
    const uint32_t per = a.paired ? 2 : 1;
    while (true) {
        const uint32_t rb = WIDE ? 1 : kRecBatch;
        unsigned long long it0 = 0;
        if (lane == 0)
            it0 = atomicAdd(WIDE ? a.wide_queue : a.queue, (unsigned long long)rb);
        it0 = __shfl(it0, 0);
        const uint64_t lim = WIDE ? *a.wide : a.nrec;
        if (it0 >= lim)
            break;
        const uint32_t nb = uint32_t(lim - it0 < rb ? lim - it0 : rb);
        /* the batch's sequences q = 0 .. per * nb - 1 start at offs[q] (lane
         * q's value) and end at offs[q + 1] */
        const uint64_t r0 = WIDE ? uint64_t(a.wide_list[it0]) : it0;
        uint64_t offs = 0;
        if (lane <= per * nb)
            offs = a.off[per * r0 + lane];
        const uint32_t nseq = per * nb;
        /* the first three 64-base chunks of sequence pf_q, loaded ahead */
        uint32_t pf_q = 0xFFFFFFFFu;
        unsigned pf[3] = {0, 0, 0};
        const auto prefetch = [&](uint32_t q) {
            const uint64_t b = __shfl(offs, int(q));
            const uint32_t len = uint32_t(__shfl(offs, int(q + 1)) - b);
            const uint8_t *s = a.seq + b;
#pragma unroll
            for (uint32_t x = 0; x < 3; ++x) {
                const uint32_t i = 64 * x + lane;
                pf[x] = i < len ? unsigned(s[i]) : 0u;
            }
            pf_q = q;
        };
        for (uint32_t j = 0; j < nb; ++j) {
            const uint64_t r = r0 + j;
            /* frameCount: len - k + 1; for a pair, the nesting of
             * MIBFClassifier.hpp:804-805 (mate 2 only when mate 1 is too short) */
            const uint32_t q1 = per * j;
            const uint32_t len1 = uint32_t(__shfl(offs, int(q1 + 1)) - __shfl(offs, int(q1)));
            uint32_t fc = len1 >= k ? len1 - k + 1 : 0;
            if (a.paired && len1 < k) {
                const uint32_t len2 =
                    uint32_t(__shfl(offs, int(q1 + 2)) - __shfl(offs, int(q1 + 1)));
                fc = len2 >= k ? len2 - k + 1 : 0;
            }
            uint32_t satc = 0, nrec = 0;
            /* one frame chunk's slots into the per-ID totals or the record
             * store */
            const auto absorb = [&](bool valid, uint32_t present, uint32_t *raw) {
                const uint32_t nh = uint32_t(__popc(present));
                bvl += valid ? S : 0;
                gat += nh;
                const bool eval = valid && nh + a.allowed_miss >= S;
                /* the distinct IDs of the frame, in seed order of first
                 * occurrence; emit(x, record) for each */
                bool allsat = nh > 0, anysat = false;
#pragma unroll
                for (uint32_t x = 0; x < SMAX; ++x)
                    if ((present >> x) & 1) {
                        if (raw[x] & kSatBit)
                            anysat = true;
                        else
                            allsat = false;
                    }
                const auto each_id = [&](auto emit) {
                    uint32_t n = 0;
#pragma unroll
                    for (uint32_t x = 0; x < SMAX; ++x) {
                        if (!((present >> x) & 1))
                            continue;
                        const uint32_t id = raw[x] & kIdMask;
                        bool seen = false;
#pragma unroll
                        for (uint32_t y = 0; y < x; ++y)
                            if (((present >> y) & 1) && (raw[y] & kIdMask) == id)
                                seen = true;
                        if (seen)
                            continue;
                        uint32_t occ = 0, ns = 0;
#pragma unroll
                        for (uint32_t y = x; y < SMAX; ++y)
                            if (((present >> y) & 1) && (raw[y] & kIdMask) == id) {
                                ++occ;
                                if (!(raw[y] & kSatBit))
                                    ++ns;
                            }
                        emit(n++, make_rec(id, anysat ? 0u : 1u, occ, ns));
                    }
                    return n;
                };
                if (eval)
                    satc += allsat ? 1 : 0;
                if (DIRECT) {
                    /* The lanes whose frame's first distinct ID is the first
                     * evaluated lane's (most frames of a read: its
                     * reference) add their six statistics by ballots and
                     * one lane's update; every other distinct ID of a frame
                     * goes through LDS atomics.  The block is one wave, so
                     * its LDS operations happen in program order. */
                    uint32_t lid = 0xFFFFFFFFu, locc = 0, lns = 0;
                    if (eval)
                        each_id([&](uint32_t x, uint64_t v) {
                            if (x == 0) {
                                lid = uint32_t(v >> 32);
                                locc = (uint32_t(v) >> 8) & 0xFF;
                                lns = uint32_t(v) & 0xFF;
                            }
                        });
                    const uint64_t em = __ballot(lid != 0xFFFFFFFFu);
                    uint32_t leader = 0xFFFFFFFFu;
                    if (em) {
                        const int first = int(__ffsll((unsigned long long)em) - 1);
                        leader = uint32_t(__shfl(int(lid), first));
                        const bool mine = lid == leader;
                        const uint64_t bm = __ballot(mine);
                        uint32_t tot = 0, tns = 0;
                        for (uint32_t b = 0; (1u << b) <= S; ++b) {
                            tot += uint32_t(__popcll(__ballot(mine && ((locc >> b) & 1)))) << b;
                            tns += uint32_t(__popcll(__ballot(mine && ((lns >> b) & 1)))) << b;
                        }
                        const uint32_t cnt = uint32_t(__popcll(bm));
                        const uint32_t nsat = uint32_t(__popcll(__ballot(mine && lns != 0)));
                        const uint32_t solid = uint32_t(__popcll(__ballot(mine && locc == S)));
                        const uint32_t nsf = uint32_t(__popcll(__ballot(mine && !anysat)));
                        if (lane == uint32_t(first) && leader < nid1) {
                            uint32_t *A = acc + kDirectWords * leader;
                            A[0] += cnt;
                            A[1] += nsat;
                            A[2] += tot;
                            A[3] += tns;
                            A[4] += solid;
                            A[5] += nsf;
                        }
                    }
                    if (eval)
                        each_id([&](uint32_t x, uint64_t v) {
                            const uint32_t id = uint32_t(v >> 32);
                            if (x == 0 && id == leader)
                                return;
                            const uint32_t occ = (uint32_t(v) >> 8) & 0xFF, ns = uint32_t(v) & 0xFF;
                            if (id < nid1) {
                                uint32_t *A = acc + kDirectWords * id;
                                atomicAdd(A, 1u);
                                atomicAdd(A + 2, occ);
                                if (ns) {
                                    atomicAdd(A + 1, 1u);
                                    atomicAdd(A + 3, ns);
                                }
                                if (occ == S)
                                    atomicAdd(A + 4, 1u);
                                if ((uint32_t(v) >> 16) & 1)
                                    atomicAdd(A + 5, 1u);
                            }
                        });
                    return;
                }
                uint32_t myrec = 0;
                if (eval)
                    myrec = each_id([](uint32_t, uint64_t) {});
                /* append this chunk's records: exclusive scan of myrec */
                uint32_t pre = myrec;
                for (int dd = 1; dd < 64; dd <<= 1) {
                    const uint32_t t = __shfl_up(pre, unsigned(dd));
                    if (int(lane) >= dd)
                        pre += t;
                }
                const uint32_t total = __shfl(pre, 63);
                const uint32_t base = nrec + pre - myrec;
                if (eval)
                    each_id([&](uint32_t x, uint64_t v) {
                        if (base + x < cap)
                            rec[base + x] = v;
                    });
                nrec += total;
            };
            /* the S seed hashes of a frame: the don't-care bases XORed out of
             * the window values (spaced seeds), or ntHash's multi-hash of the
             * canonical value (k-mer mode) */
            const auto seed_hashes = [&](const Frame &fr, uint64_t *hs) {
                if (!spaced) {
                    const uint64_t h0 = fr.rc < fr.fwd ? fr.rc : fr.fwd;
#pragma unroll
                    for (uint32_t sd = 0; sd < SMAX; ++sd)
                        hs[sd] = multi_hash(h0, sd, uint64_t(k) * kMultiSeed);
                    return;
                }
                uint64_t fw[SMAX], rv[SMAX];
#pragma unroll
                for (uint32_t sd = 0; sd < SMAX; ++sd) {
                    fw[sd] = fr.fwd;
                    rv[sd] = fr.rc;
                }
                for (uint32_t dcm = a.d.dc_union; dcm; dcm &= dcm - 1) {
                    const uint32_t jj = uint32_t(__builtin_ctz(dcm));
                    const uint32_t cj = uint32_t(fr.W >> (2 * (k - 1 - jj))) & 3u;
                    const ulonglong2 t = dct[jj * 4 + cj];
#pragma unroll
                    for (uint32_t sd = 0; sd < SMAX; ++sd)
                        if (sd < S && !((a.d.care[sd] >> jj) & 1)) {
                            fw[sd] ^= t.x;
                            rv[sd] ^= t.y;
                        }
                }
#pragma unroll
                for (uint32_t sd = 0; sd < SMAX; ++sd)
                    hs[sd] = rv[sd] < fw[sd] ? rv[sd] : fw[sd];
            };
            for (uint32_t mate = 0; mate < per && fc; ++mate) {
                const uint32_t q = q1 + mate;
                const uint64_t b = __shfl(offs, int(q));
                const uint32_t len = uint32_t(__shfl(offs, int(q + 1)) - b);
                if (len < k)
                    continue;
                const uint8_t *s = a.seq + b;
                if (pf_q != q)
                    prefetch(q);
                unsigned c0 = pf[0], c1 = pf[1], c2 = pf[2];
                pf_q = 0xFFFFFFFFu;
                const uint32_t nfr = len - k + 1;
                ScanCarry cy;
                cy.H = cy.G = cy.W = 0;
                cy.N = -1;
                BaseChunk X = scan_chunk(c0, lane, len, lane, spaced, cy);
                uint64_t hx = 0, gx = 0; /* prefixes through the base before X */
                for (uint32_t c = 0; 64 * c < nfr; c += FG) {
                    /* base chunks c + 1 .. c + FG */
                    BaseChunk Y[FG];
                    unsigned cb[FG];
#pragma unroll
                    for (uint32_t f = 0; f < FG; ++f) {
                        const uint32_t i = 64 * (c + 1 + f) + lane;
                        if (c == 0)
                            cb[f] = f == 0 ? c1 : c2;
                        else
                            cb[f] = i < len ? unsigned(s[i]) : 0u;
                    }
#pragma unroll
                    for (uint32_t f = 0; f < FG; ++f)
                        Y[f] = scan_chunk(cb[f], 64 * (c + 1 + f) + lane, len, lane, spaced, cy);
                    /* frame chunks c .. c + FG - 1: hashes, then every lookup */
                    uint32_t raw[FG][SMAX];
                    uint32_t present[FG];
                    bool valid[FG];
#pragma unroll
                    for (uint32_t f = 0; f < FG; ++f) {
                        const BaseChunk &Xf = f == 0 ? X : Y[f > 0 ? f - 1 : 0];
                        /* prefixes through the last base of the chunk before Xf */
                        const BaseChunk &Wf = f == 1 ? X : Y[f > 1 ? f - 2 : 0];
                        const uint64_t hp = f == 0 ? hx : __shfl(Wf.H, 63);
                        const uint64_t gp = f == 0 ? gx : __shfl(Wf.G, 63);
                        const Frame fr = frame_at(Xf, Y[f], hp, gp, 64 * (c + f), lane, k, nfr,
                                                  spaced);
                        valid[f] = fr.valid;
                        uint64_t hs[SMAX];
                        seed_hashes(fr, hs);
                        present[f] = 0;
                        if (TABLE) {
                            /* one load per seed: the ID word, or 0 for a clear bit */
#pragma unroll
                            for (uint32_t sd = 0; sd < SMAX; ++sd)
                                raw[f][sd] = (sd < S && fr.valid)
                                                 ? a.d.direct[fastmod(hs[sd], a.d.m, a.d.magic,
                                                                      a.d.sh1, a.d.sh2)]
                                                 : 0u;
                        } else {
                            /* bit and rank from the line, then the ID slot */
                            uint64_t rk[SMAX];
#pragma unroll
                            for (uint32_t sd = 0; sd < SMAX; ++sd) {
                                rk[sd] = 0;
                                if (sd < S && fr.valid) {
                                    const bool bit = bv_rank_line(
                                        a.d.lines,
                                        fastmod(hs[sd], a.d.m, a.d.magic, a.d.sh1, a.d.sh2),
                                        rk[sd]);
                                    present[f] |= uint32_t(bit) << sd;
                                }
                            }
#pragma unroll
                            for (uint32_t sd = 0; sd < SMAX; ++sd)
                                raw[f][sd] = ((present[f] >> sd) & 1) ? a.d.data[rk[sd]] : 0u;
                        }
                    }
                    /* the next sequence's first bases load while these
                     * lookups are in flight */
                    if (c == 0 && q + 1 < nseq)
                        prefetch(q + 1);
#pragma unroll
                    for (uint32_t f = 0; f < FG; ++f) {
                        if (TABLE) {
#pragma unroll
                            for (uint32_t sd = 0; sd < SMAX; ++sd) {
                                present[f] |= uint32_t(raw[f][sd] != 0) << sd;
                                if ((raw[f][sd] & kIdMask) == kIdMask)
                                    raw[f][sd] &= kSatBit; /* ID 0 */
                            }
                        }
                        if (64 * (c + f) < nfr)
                            absorb(valid[f], present[f], raw[f]);
                    }
                    hx = __shfl(FG == 1 ? X.H : Y[FG - 2].H, 63);
                    gx = __shfl(FG == 1 ? X.G : Y[FG - 2].G, 63);
                    X = Y[FG - 1];
                }
            }
        for (int o = 32; o > 0; o >>= 1)
            satc += __shfl_xor(satc, o);
        bool over = false;
        const uint32_t row = fc <= a.max_frames ? fc : 0;
        if (DIRECT) {
            if (lane == 0)
                ncand_s = 0;
            if (row != mc_row) {
                for (uint32_t id = lane; id < nid1; id += 64)
                    mcs[id] = (id >= 1 && row) ? a.min_count[uint64_t(row) * nid1 + id]
                                               : 0xFFFFFFFFu;
                mc_row = row;
            }
            __syncthreads();
            /* candidates straight from the per-ID totals; each lane clears
             * the IDs it read for the next record */
            for (uint32_t id = lane; id < nid1; id += 64) {
                uint32_t *A = acc + kDirectWords * id;
                const uint32_t cnt = A[0];
                if (!cnt)
                    continue;
                const uint32_t nsat = A[1], tot = A[2], tns = A[3], solid = A[4], nsf = A[5];
#pragma unroll
                for (uint32_t x = 0; x < kDirectWords; ++x)
                    A[x] = 0;
                const uint32_t mc = mcs[id];
                if (cnt >= mc && nsat >= a.min_nonsat) {
                    const uint32_t slot = atomicAdd(&ncand_s, 1u);
                    if (slot < kMaxCand || (WIDE && slot - kMaxCand < a.spill_stride)) {
                        uint32_t *c7 = slot < kMaxCand
                                           ? cand + slot * kHitWords
                                           : a.spill + (uint64_t(blockIdx.x) * a.spill_stride +
                                                        (slot - kMaxCand)) * kHitWords;
                        c7[0] = id;
                        c7[1] = cnt;
                        c7[2] = nsat;
                        c7[3] = tot;
                        c7[4] = tns;
                        c7[5] = solid;
                        c7[6] = nsf;
                    }
                }
            }
            nrec = 0;
        } else if (nrec > cap) { /* cannot happen when the host sized cap by frames x S */
            over = true;
            nrec = 0;
        }
        __syncthreads();
        /* bitonic sort of the records (IDs in the high bits) */
        uint32_t n2 = 1;
        while (n2 < nrec)
            n2 <<= 1;
        for (uint32_t x = nrec + lane; x < n2; x += 64)
            rec[x] = ~uint64_t(0);
        __syncthreads();
        for (uint32_t sz = 2; sz <= n2; sz <<= 1) {
            for (uint32_t st = sz >> 1; st > 0; st >>= 1) {
                for (uint32_t x = lane; x < n2; x += 64) {
                    const uint32_t y = x ^ st;
                    if (y > x) {
                        const bool up = (x & sz) == 0;
                        const uint64_t vx = rec[x], vy = rec[y];
                        if ((vx > vy) == up) {
                            rec[x] = vy;
                            rec[y] = vx;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (!DIRECT) {
            if (lane == 0)
                ncand_s = 0;
            __syncthreads();
        }
        /* per-ID totals: the lane at the head of each run of one ID reduces it */
        for (uint32_t x = lane; x < nrec; x += 64) {
            const uint32_t id = uint32_t(rec[x] >> 32);
            if (x > 0 && uint32_t(rec[x - 1] >> 32) == id)
                continue;
            uint32_t cnt = 0, nsat = 0, tot = 0, tns = 0, solid = 0, nsf = 0;
            for (uint32_t y = x; y < nrec && uint32_t(rec[y] >> 32) == id; ++y) {
                const uint32_t v = uint32_t(rec[y]);
                const uint32_t occ = (v >> 8) & 0xFF, ns = v & 0xFF;
                cnt += 1;
                nsat += ns ? 1 : 0;
                tot += occ;
                tns += ns;
                solid += occ == S ? 1 : 0;
                nsf += (v >> 16) & 1;
            }
            const uint32_t mc = (id >= 1 && id <= a.d.nids && row)
                                    ? a.min_count[uint64_t(row) * (a.d.nids + 1) + id]
                                    : 0xFFFFFFFFu;
            if (cnt >= mc && nsat >= a.min_nonsat) {
                const uint32_t slot = atomicAdd(&ncand_s, 1u);
                if (slot < kMaxCand || (WIDE && slot - kMaxCand < a.spill_stride)) {
                    uint32_t *c7 = slot < kMaxCand
                                       ? cand + slot * kHitWords
                                       : a.spill + (uint64_t(blockIdx.x) * a.spill_stride +
                                                    (slot - kMaxCand)) * kHitWords;
                    c7[0] = id;
                    c7[1] = cnt;
                    c7[2] = nsat;
                    c7[3] = tot;
                    c7[4] = tns;
                    c7[5] = solid;
                    c7[6] = nsf;
                }
            }
        }
        __syncthreads();
        /* ranking and output by lane 0 */
        if (lane == 0) {
            const uint32_t ncand = ncand_s;
            uint32_t nout = 0;
            bool listed = false;
            if (over) {
            } else if (ncand > kMaxCand && !WIDE) {
                /* evaluated again by the wide pass, which writes the result */
                const unsigned long long w = atomicAdd(a.wide, 1ull);
                a.wide_list[w] = uint32_t(r);
                listed = true;
                atomicAdd(a.counters + 3, 1ull);
            } else if (ncand > kMaxCand) {
                /* wide pass: best first by selection over every candidate */
                const auto C = [&](uint32_t x) -> const uint32_t * {
                    return x < kMaxCand ? cand + x * kHitWords
                                        : a.spill + (uint64_t(blockIdx.x) * a.spill_stride +
                                                     (x - kMaxCand)) * kHitWords;
                };
                const uint32_t nc = ncand < kMaxCand + a.spill_stride ? ncand
                                                                      : kMaxCand + a.spill_stride;
                const uint32_t *top = C(0);
                for (uint32_t x = 1; x < nc; ++x)
                    if (ranks_before(C(x), top))
                        top = C(x);
                bool agree = true;
                if (a.best_agree)
                    for (uint32_t x = 0; x < nc; ++x)
                        for (uint32_t f = 1; f < kHitWords; ++f)
                            if (C(x)[f] > top[f])
                                agree = false;
                if (agree) {
                    const double bar = a.multi_thresh * double(top[2]);
                    uint32_t npass = 0;
                    for (uint32_t x = 0; x < nc; ++x)
                        npass += double(C(x)[2]) >= bar ? 1 : 0;
                    const uint32_t nkeep = npass < a.max_hits ? npass : a.max_hits;
                    const uint32_t *prev = nullptr;
                    for (uint32_t j = 0; j < nkeep; ++j) {
                        const uint32_t *sel = nullptr;
                        for (uint32_t x = 0; x < nc; ++x) {
                            const uint32_t *cv = C(x);
                            if (double(cv[2]) < bar || (prev && !ranks_before(prev, cv)))
                                continue;
                            if (!sel || ranks_before(cv, sel))
                                sel = cv;
                        }
                        for (uint32_t f = 0; f < kHitWords; ++f)
                            a.hits[(r * a.max_hits + j) * kHitWords + f] = sel[f];
                        prev = sel;
                    }
                    if (npass > nkeep)
                        atomicAdd(a.counters + 2, (unsigned long long)(npass - nkeep));
                    nout = npass;
                }
            } else {
                uint32_t nc = ncand;
                uint8_t *order = order_s;
                for (uint32_t x = 0; x < nc; ++x) {
                    uint32_t y = x;
                    while (y > 0 &&
                           ranks_before(cand + x * kHitWords, cand + order[y - 1] * kHitWords)) {
                        order[y] = order[y - 1];
                        --y;
                    }
                    order[y] = x;
                }
                if (nc && a.best_agree) {
                    const uint32_t *t = cand + order[0] * kHitWords;
                    for (uint32_t x = 0; x < nc; ++x)
                        for (uint32_t f = 1; f < kHitWords; ++f)
                            if (cand[x * kHitWords + f] > t[f])
                                nc = 0;
                }
                if (nc) {
                    const double bar = a.multi_thresh * double(cand[order[0] * kHitWords + 2]);
                    for (uint32_t x = 0; x < nc; ++x) {
                        const uint32_t *cv = cand + order[x] * kHitWords;
                        if (double(cv[2]) < bar)
                            break;
                        if (nout < a.max_hits)
                            for (uint32_t f = 0; f < kHitWords; ++f)
                                a.hits[(r * a.max_hits + nout) * kHitWords + f] = cv[f];
                        else
                            atomicAdd(a.counters + 2, 1ull);
                        ++nout;
                    }
                }
            }
            /* a listed record's 0 is a placeholder until the wide pass */
            a.nhits[r] = over ? 0xFFFFFFFFu : (listed ? 0u : nout);
            a.sat[r] = satc;
        }
        __syncthreads();
        }
    }
    if (WIDE)
        return;
    for (int o = 32; o > 0; o >>= 1) {
        bvl += __shfl_xor(bvl, o);
        gat += __shfl_xor(gat, o);
    }
    if (lane == 0) {
        atomicAdd(a.counters + 0, bvl);
        atomicAdd(a.counters + 1, gat);
    }
}

__global__ void k_mibf_zero(unsigned long long *p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        p[i] = 0;
}

/* the first pass over every record, then the wide pass over the records it
 * listed (with an empty list the wide workgroups end at once) */
template <uint32_t CAP, uint32_t SMAX, bool DIRECT, bool TABLE>
void query_pair(const QueryArgs &a, int cu, size_t shm, hipStream_t st) {
    /* as many one-wave workgroups as fit on every CU (registers and LDS),
     * each pulling records from the queue */
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_mibf_query<CAP, SMAX, DIRECT, false, TABLE>, 64, shm) != hipSuccess ||
        per_cu < 1)
        per_cu = 16;
    uint64_t nb = std::max<uint64_t>(1, std::min<uint64_t>(a.nrec, uint64_t(cu) * per_cu));
    if (CAP == 0)
        nb = std::min<uint64_t>(nb, a.grec_blocks); /* the device record area's room */
    const uint32_t blocks = uint32_t(nb);
    k_mibf_query<CAP, SMAX, DIRECT, false, TABLE><<<blocks, 64, shm, st>>>(a);
    if (a.spill_stride)
        k_mibf_query<CAP, SMAX, DIRECT, true, TABLE><<<kWideBlocks, 64, shm, st>>>(a);
}

/* the kernel for the filter's ID count (per-ID totals in LDS, or the record
 * store sized by the longest record), seed count and layout */
template <bool TABLE>
void query_launch(const QueryArgs &a, bool direct, uint64_t recs, bool few, int blocks, size_t shm,
                  hipStream_t st) {
    if (direct)
        few ? query_pair<1, 4, true, TABLE>(a, blocks, shm, st)
            : query_pair<1, 8, true, TABLE>(a, blocks, shm, st);
    else if (recs <= 1024)
        few ? query_pair<1024, 4, false, TABLE>(a, blocks, 0, st)
            : query_pair<1024, 8, false, TABLE>(a, blocks, 0, st);
    else if (recs <= 4096)
        few ? query_pair<4096, 4, false, TABLE>(a, blocks, 0, st)
            : query_pair<4096, 8, false, TABLE>(a, blocks, 0, st);
    else if (recs <= 8192)
        few ? query_pair<8192, 4, false, TABLE>(a, blocks, 0, st)
            : query_pair<8192, 8, false, TABLE>(a, blocks, 0, st);
    else
        few ? query_pair<0, 4, false, TABLE>(a, blocks, 0, st)
            : query_pair<0, 8, false, TABLE>(a, blocks, 0, st);
}

} // namespace

/* ======================================================================= */

struct bbt_mibf {
    MibfDev d;
    uint64_t *dlines = nullptr;
    uint32_t *ddata = nullptr;
    uint32_t *ddirect = nullptr; /* direct-mapped table (d.direct), or null */
    uint64_t nlines = 0, pop = 0, pop_sat = 0;
    std::vector<std::string> seeds;
    std::vector<std::string> names;  /* names[0] = "repeat", then IDs 1..n */
    std::vector<uint64_t> id_slots;  /* slots per ID (index = ID) */
    /* minCount table for (score, allowed_miss), rows 0..min_frames */
    uint32_t *dmin = nullptr;
    std::vector<uint32_t> hmin; /* host copy of the table */
    double min_score = -1;
    uint32_t min_miss = 0xFFFFFFFFu, min_frames = 0;
    /* [0] scan total, [1] separator errors, [8] queue, [9] records listed for
     * the wide pass, [10] wide-pass queue */
    unsigned long long *dctl = nullptr;
    /* the wide pass: listed record indices, and per workgroup room for the
     * candidates past kMaxCand (nids - kMaxCand of them at most) */
    uint32_t *dwide_list = nullptr;
    uint64_t wide_cap = 0;
    uint32_t *dspill = nullptr;
    int cu = 0;
    /* recorded after each query's kernels: the work queue, the wide-pass
     * lists and the minimum-count table belong to the filter, so the next
     * query (on any stream) waits for it, and the host waits for it before
     * freeing or refilling those buffers */
    hipEvent_t last = nullptr;
    bool has_last = false;
    /* BBT_READS_2BIT queries: the batch unpacked to ASCII */
    uint8_t *dascii = nullptr;
    uint64_t ascii_cap = 0;
    /* frame records of records too long for LDS (more than 1023 IDs) */
    uint64_t *dgrec = nullptr;
    uint32_t grec_cap = 0, grec_blocks = 0;
    std::mutex mu;
};

namespace {

bbt_status mibf_init(bbt_mibf *mf, const char *const *seeds, uint32_t nseeds,
                     uint32_t kmer_size) {
    if (nseeds == 0 || nseeds > kMaxSeeds)
        return mfail(BBT_ERR_ARG, "MIBF: 1..8 spaced seeds or hash functions required");
    const size_t span = seeds ? std::strlen(seeds[0]) : kmer_size;
    if (span == 0 || span > 32)
        return mfail(BBT_ERR_UNSUPPORTED, "MIBF: seed span / k-mer size must be 1..32");
    std::memset(&mf->d, 0, sizeof(mf->d));
    mf->d.kmer = seeds ? 0 : 1;
    for (uint32_t i = 0; seeds && i < nseeds; ++i) {
        const std::string sd = seeds[i];
        if (sd.size() != span)
            return mfail(BBT_ERR_ARG, "MIBF: all seeds must have the same span");
        uint32_t care = 0;
        for (size_t j = 0; j < span; ++j) {
            if (sd[j] != '0' && sd[j] != '1')
                return mfail(BBT_ERR_ARG, "MIBF: seeds are strings of 0 and 1");
            if (sd[j] == '1')
                care |= 1u << j;
        }
        mf->d.care[i] = care;
        mf->seeds.push_back(sd);
    }
    if (!seeds)
        for (uint32_t i = 0; i < nseeds; ++i)
            mf->d.care[i] = span == 32 ? 0xFFFFFFFFu : (1u << span) - 1;
    const uint32_t all = span == 32 ? 0xFFFFFFFFu : (1u << span) - 1;
    mf->d.dc_union = 0;
    for (uint32_t i = 0; i < nseeds; ++i)
        mf->d.dc_union |= all & ~mf->d.care[i];
    mf->d.S = nseeds;
    mf->d.span = uint32_t(span);
    int dev = 0;
    MHIP(hipGetDevice(&dev));
    MHIP(hipDeviceGetAttribute(&mf->cu, hipDeviceAttributeMultiprocessorCount, dev));
    MHIP(hipMalloc(&mf->dctl, 16 * sizeof(unsigned long long)));
    MHIP(hipEventCreateWithFlags(&mf->last, hipEventDisableTiming));
    return BBT_OK;
}

void set_m(bbt_mibf *mf, uint64_t m) {
    mf->d.m = m;
    const FastMod fm = fastmod_make(m);
    mf->d.magic = fm.magic;
    mf->d.sh1 = fm.sh1;
    mf->d.sh2 = fm.sh2;
    mf->nlines = (m / 64 + kWordsPerLine - 1) / kWordsPerLine;
}

/* slot statistics from the device ID array */
bbt_status mibf_stats(bbt_mibf *mf, hipStream_t st) {
    const uint32_t n = mf->d.nids;
    unsigned long long *hist = nullptr;
    MHIP(hipMalloc(&hist, (size_t(n) + 2) * 8));
    bool ok = hipMemsetAsync(hist, 0, (size_t(n) + 2) * 8, st) == hipSuccess;
    if (ok && mf->pop)
        k_mibf_hist<<<4096, 256, 0, st>>>(mf->ddata, mf->pop, n, hist);
    std::vector<uint64_t> h(size_t(n) + 2, 0);
    ok = ok && hipMemcpyAsync(h.data(), hist, h.size() * 8, hipMemcpyDeviceToHost, st) ==
                   hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    (void)hipFree(hist);
    if (!ok)
        return mfail(BBT_ERR_HIP, "MIBF: slot statistics failed");
    mf->id_slots.assign(h.begin(), h.begin() + n + 1);
    mf->pop_sat = h[n + 1];
    return BBT_OK;
}

/* The direct-mapped table (m 32-bit words) is built when it fits in half
 * of the free device memory, unless BBT_MIBF_DIRECT=0: on a GPU with
 * hundreds of GB of HBM it turns each query lookup (a line load for the bit
 * and its rank, then a dependent ID load) into one random load. */
bbt_status mibf_make_direct(bbt_mibf *mf, hipStream_t st) {
    const char *env = std::getenv("BBT_MIBF_DIRECT");
    if (env && std::strcmp(env, "0") == 0)
        return BBT_OK;
    size_t free_b = 0, total_b = 0;
    MHIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t bytes = mf->d.m * 4;
    if (bytes > free_b / 2)
        return BBT_OK;
    if (hipMalloc(&mf->ddirect, bytes) != hipSuccess) {
        (void)hipGetLastError();
        mf->ddirect = nullptr;
        return BBT_OK; /* the lines and ID array serve queries as before */
    }
    const uint64_t nwords = mf->d.m / 64;
    k_mibf_direct<<<uint32_t(std::min<uint64_t>((nwords + 255) / 256, 1u << 16)), 256, 0, st>>>(
        mf->dlines, nwords, mf->ddata, mf->ddirect);
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(st));
    mf->d.direct = mf->ddirect;
    return BBT_OK;
}

constexpr char kMagic[8] = {'B', 'B', 'T', 'M', 'I', 'B', 'F', '1'};

struct MibfHeader {
    char magic[8];
    uint32_t nseeds, span, nids, kmer; /* kmer = 1: no seed strings follow */
    uint64_t m, pop;
};
static_assert(sizeof(MibfHeader) == 40, "MibfHeader layout");

} // namespace

extern "C" {

uint64_t bbt_mibf_calc_size(uint64_t entries, uint32_t nseeds, double occupancy) {
    const uint64_t v = uint64_t(-double(entries) * double(nseeds) / std::log(1.0 - occupancy));
    return v + (64 - v % 64);
}

bbt_status bbt_mibf_free(bbt_mibf *mf) {
    if (!mf)
        return BBT_OK;
    (void)hipFree(mf->dlines);
    (void)hipFree(mf->ddata);
    (void)hipFree(mf->ddirect);
    (void)hipFree(mf->dmin);
    (void)hipFree(mf->dctl);
    (void)hipFree(mf->dwide_list);
    (void)hipFree(mf->dspill);
    (void)hipFree(mf->dascii);
    (void)hipFree(mf->dgrec);
    if (mf->last)
        (void)hipEventDestroy(mf->last);
    delete mf;
    return BBT_OK;
}

bbt_status bbt_mibf_build(const char *dseq, const uint64_t *drec_off, const uint32_t *drec_id,
                          const uint64_t *rec_off, uint64_t nrec, const char *const *names,
                          uint32_t nids, const char *const *seeds, uint32_t nseeds,
                          uint32_t kmer_size, double occupancy, uint64_t entries,
                          uint64_t m_override, bbt_mibf **out, void *stream) {
    if (!out || !dseq || !drec_off || !drec_id || !rec_off || nrec == 0 || nids == 0)
        return mfail(BBT_ERR_ARG, "bbt_mibf_build: bad argument");
    if (nids >= kIdMask)
        return mfail(BBT_ERR_ARG, "bbt_mibf_build: IDs must fit in 31 bits");
    if (!(occupancy > 0 && occupancy < 1) && !m_override)
        return mfail(BBT_ERR_ARG, "bbt_mibf_build: occupancy must be in (0, 1)");
    *out = nullptr;
    hipStream_t st = static_cast<hipStream_t>(stream);
    bbt_mibf *mf = new bbt_mibf();
    bbt_status s = mibf_init(mf, seeds, nseeds, kmer_size);
    if (s != BBT_OK) {
        bbt_mibf_free(mf);
        return s;
    }
    const uint32_t k = mf->d.span;
    const uint64_t nbytes = rec_off[nrec];
    if (!entries) /* windows per record, as MIBFGen counts them (len - k + 1) */
        for (uint64_t r = 0; r < nrec; ++r) {
            const uint64_t len = rec_off[r + 1] - rec_off[r] - 1;
            entries += len >= k ? len - k + 1 : 0;
        }
    const uint64_t m = m_override ? m_override : bbt_mibf_calc_size(entries, nseeds, occupancy);
    if (m == 0 || m % 64) {
        bbt_mibf_free(mf);
        return mfail(BBT_ERR_ARG, "bbt_mibf_build: m must be a positive multiple of 64");
    }
    set_m(mf, m);
    mf->d.nids = nids;
    mf->names.push_back("repeat");
    for (uint32_t i = 0; i < nids; ++i)
        mf->names.push_back(names && names[i] ? names[i] : std::to_string(i + 1));
    uint64_t *bits = nullptr;
    unsigned long long *best = nullptr;
    auto done = [&](bbt_status e) {
        (void)hipFree(bits);
        (void)hipFree(best);
        if (e != BBT_OK)
            bbt_mibf_free(mf);
        else
            *out = mf;
        return e;
    };
    const uint8_t *seq = reinterpret_cast<const uint8_t *>(dseq);
    const uint32_t chunk = 256;
    const uint64_t threads = (nbytes + chunk - 1) / chunk;
    const uint32_t blocks = uint32_t(std::max<uint64_t>(1, (threads + 255) / 256));
    unsigned long long ctl[2] = {0, 0};
    if (hipMemsetAsync(mf->dctl, 0, 16 * 8, st) != hipSuccess)
        return done(mfail(BBT_ERR_HIP, "bbt_mibf_build: memset failed"));
    k_mibf_check_sep<<<64, 256, 0, st>>>(seq, drec_off, nrec, mf->dctl + 1);
    if (hipMalloc(&bits, (m / 64) * 8) != hipSuccess ||
        hipMalloc(&mf->dlines, mf->nlines * 64) != hipSuccess)
        return done(mfail(BBT_ERR_NOMEM, "bbt_mibf_build: bit vector allocation failed"));
    mf->d.lines = mf->dlines;
    if (hipMemsetAsync(bits, 0, (m / 64) * 8, st) != hipSuccess)
        return done(mfail(BBT_ERR_HIP, "bbt_mibf_build: memset failed"));
    /* pass 1: bit vector (MIBFGen.hpp:144 generateBV) */
    k_mibf_bits<<<blocks, 256, 0, st>>>(seq, nbytes, mf->d, bits, chunk);
    k_mibf_lines<<<4096, 256, 0, st>>>(bits, m / 64, mf->nlines, mf->dlines);
    k_mibf_scan<<<1, 1024, 0, st>>>(mf->dlines, mf->nlines, mf->dctl);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(ctl, mf->dctl, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return done(mfail(BBT_ERR_HIP, "bbt_mibf_build: bit vector pass failed"));
    if (ctl[1])
        return done(mfail(BBT_ERR_ARG, "bbt_mibf_build: " + std::to_string(ctl[1]) +
                                           " record(s) do not end with a separator byte"));
    (void)hipFree(bits);
    bits = nullptr;
    mf->pop = ctl[0];
    const uint64_t slots = std::max<uint64_t>(1, mf->pop);
    if (hipMalloc(&mf->ddata, slots * 4) != hipSuccess ||
        hipMalloc(&best, slots * 8) != hipSuccess)
        return done(mfail(BBT_ERR_NOMEM, "bbt_mibf_build: ID array allocation failed"));
    mf->d.data = mf->ddata;
    if (hipMemsetAsync(best, 0, slots * 8, st) != hipSuccess)
        return done(mfail(BBT_ERR_HIP, "bbt_mibf_build: memset failed"));
    /* pass 2: IDs (MIBFGen.hpp:155-230 insertMIBF), order-independent ownership */
    k_mibf_own<<<blocks, 256, 0, st>>>(seq, nbytes, drec_off, drec_id, nrec, mf->d, best, chunk);
    k_mibf_settle<<<4096, 256, 0, st>>>(best, mf->pop, mf->ddata);
    /* pass 3: saturation (MIBFGen.hpp:200-262 insertSaturation) */
    k_mibf_sat<<<blocks, 256, 0, st>>>(seq, nbytes, drec_off, drec_id, nrec, mf->d, mf->ddata,
                                       chunk);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return done(mfail(BBT_ERR_HIP, "bbt_mibf_build: ID passes failed"));
    (void)hipFree(best);
    best = nullptr;
    bbt_status s2 = mibf_stats(mf, st);
    if (s2 == BBT_OK)
        s2 = mibf_make_direct(mf, st);
    return done(s2);
}

bbt_status bbt_mibf_info(const bbt_mibf *mf, bbt_mibf_info_t *info) {
    if (!mf || !info)
        return mfail(BBT_ERR_ARG, "bbt_mibf_info: null argument");
    info->m = mf->d.m;
    info->pop = mf->pop;
    info->pop_saturated = mf->pop_sat;
    info->nids = mf->d.nids;
    info->nseeds = mf->d.S;
    info->span = mf->d.span;
    info->kmer_mode = mf->d.kmer;
    info->direct = mf->d.direct ? 1 : 0;
    info->reserved = 0;
    return BBT_OK;
}

const char *bbt_mibf_name(const bbt_mibf *mf, uint32_t id) {
    if (!mf || id >= mf->names.size())
        return nullptr;
    return mf->names[id].c_str();
}

const char *bbt_mibf_seed(const bbt_mibf *mf, uint32_t i) {
    if (!mf || i >= mf->seeds.size())
        return nullptr;
    return mf->seeds[i].c_str();
}

bbt_status bbt_mibf_download(const bbt_mibf *mf, uint64_t *bv, uint32_t *data) {
    if (!mf)
        return mfail(BBT_ERR_ARG, "bbt_mibf_download: null");
    if (bv) {
        std::vector<uint64_t> lines(mf->nlines * 8);
        MHIP(hipMemcpy(lines.data(), mf->dlines, lines.size() * 8, hipMemcpyDeviceToHost));
        for (uint64_t w = 0; w < mf->d.m / 64; ++w)
            bv[w] = lines[(w / kWordsPerLine) * 8 + 1 + w % kWordsPerLine];
    }
    if (data && mf->pop)
        MHIP(hipMemcpy(data, mf->ddata, mf->pop * 4, hipMemcpyDeviceToHost));
    return BBT_OK;
}

bbt_status bbt_mibf_store(const bbt_mibf *mf, const char *prefix) {
    if (!mf || !prefix)
        return mfail(BBT_ERR_ARG, "bbt_mibf_store: null argument");
    std::vector<uint64_t> bv(mf->d.m / 64);
    std::vector<uint32_t> data(mf->pop);
    bbt_status s = bbt_mibf_download(mf, bv.data(), data.data());
    if (s != BBT_OK)
        return s;
    const std::string base(prefix);
    std::ofstream f(base + ".bf", std::ios::binary);
    MibfHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.nseeds = mf->d.S;
    h.span = mf->d.span;
    h.nids = mf->d.nids;
    h.kmer = mf->d.kmer;
    h.m = mf->d.m;
    h.pop = mf->pop;
    f.write(reinterpret_cast<const char *>(&h), sizeof(h));
    for (const std::string &sd : mf->seeds)
        f.write(sd.data(), std::streamsize(sd.size()));
    f.write(reinterpret_cast<const char *>(bv.data()), std::streamsize(bv.size() * 8));
    f.write(reinterpret_cast<const char *>(data.data()), std::streamsize(data.size() * 4));
    if (!f)
        return mfail(BBT_ERR_IO, "bbt_mibf_store: cannot write " + base + ".bf");
    /* ID file (MIBFGen.hpp:424-430): "<id>\t<name>" per line */
    std::ofstream ids(base + "_ids.txt");
    for (uint32_t i = 1; i < mf->names.size(); ++i)
        ids << i << "\t" << mf->names[i] << "\n";
    if (!ids)
        return mfail(BBT_ERR_IO, "bbt_mibf_store: cannot write " + base + "_ids.txt");
    return BBT_OK;
}

bbt_status bbt_mibf_load(const char *path, bbt_mibf **out, void *stream) {
    if (!path || !out)
        return mfail(BBT_ERR_ARG, "bbt_mibf_load: null argument");
    *out = nullptr;
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return mfail(BBT_ERR_IO, std::string("bbt_mibf_load: cannot open ") + path);
    MibfHeader h;
    f.read(reinterpret_cast<char *>(&h), sizeof(h));
    if (!f || std::memcmp(h.magic, kMagic, 8) != 0 || h.nseeds == 0 || h.nseeds > kMaxSeeds ||
        h.span == 0 || h.span > 32 || h.m == 0 || h.m % 64 || h.pop > h.m || h.nids == 0 ||
        h.kmer > 1)
        return mfail(BBT_ERR_FORMAT, std::string("bbt_mibf_load: not a multi-index filter: ") + path);
    std::vector<std::string> seeds(h.kmer ? 0 : h.nseeds, std::string(h.span, '0'));
    for (std::string &sd : seeds)
        f.read(&sd[0], h.span);
    std::vector<uint64_t> bv(h.m / 64);
    std::vector<uint32_t> data(h.pop);
    f.read(reinterpret_cast<char *>(bv.data()), std::streamsize(bv.size() * 8));
    f.read(reinterpret_cast<char *>(data.data()), std::streamsize(data.size() * 4));
    if (!f)
        return mfail(BBT_ERR_FORMAT, std::string("bbt_mibf_load: truncated file ") + path);
    uint64_t pop = 0;
    for (uint64_t w : bv)
        pop += uint64_t(__builtin_popcountll(w));
    if (pop != h.pop)
        return mfail(BBT_ERR_FORMAT, "bbt_mibf_load: popcount does not match the header");
    /* names from the ID file beside the filter: <prefix>_ids.txt for <prefix>.bf
     * (MIBFClassifier.hpp:52-76) */
    std::string p(path);
    const std::string stem = p.size() > 3 && p.compare(p.size() - 3, 3, ".bf") == 0
                                 ? p.substr(0, p.size() - 3)
                                 : p;
    std::ifstream idf(stem + "_ids.txt");
    if (!idf)
        return mfail(BBT_ERR_IO, "bbt_mibf_load: " + stem + "_ids.txt cannot be opened; "
                                 "a corresponding id file is needed");
    bbt_mibf *mf = new bbt_mibf();
    std::vector<const char *> sp;
    for (const std::string &sd : seeds)
        sp.push_back(sd.c_str());
    bbt_status s = mibf_init(mf, h.kmer ? nullptr : sp.data(), h.nseeds, h.span);
    if (s != BBT_OK) {
        bbt_mibf_free(mf);
        return s;
    }
    set_m(mf, h.m);
    mf->d.nids = h.nids;
    mf->pop = h.pop;
    mf->names.assign(h.nids + 1, std::string());
    mf->names[0] = "repeat";
    std::string line;
    while (std::getline(idf, line)) {
        const size_t tab = line.find_first_of(" \t");
        if (tab == std::string::npos)
            continue;
        const unsigned long id = std::strtoul(line.c_str(), nullptr, 10);
        size_t b = line.find_first_not_of(" \t", tab);
        if (id >= 1 && id <= h.nids && b != std::string::npos)
            mf->names[id] = line.substr(b, line.find_first_of(" \t", b) - b);
    }
    std::vector<uint64_t> lines(mf->nlines * 8, 0);
    uint64_t run = 0;
    for (uint64_t l = 0; l < mf->nlines; ++l) {
        lines[l * 8] = run;
        for (uint32_t t = 0; t < kWordsPerLine; ++t) {
            const uint64_t w = l * kWordsPerLine + t;
            const uint64_t v = w < bv.size() ? bv[w] : 0;
            lines[l * 8 + 1 + t] = v;
            run += uint64_t(__builtin_popcountll(v));
        }
    }
    if (hipMalloc(&mf->dlines, lines.size() * 8) != hipSuccess ||
        hipMalloc(&mf->ddata, std::max<uint64_t>(1, h.pop) * 4) != hipSuccess) {
        bbt_mibf_free(mf);
        return mfail(BBT_ERR_NOMEM, "bbt_mibf_load: device allocation failed");
    }
    mf->d.lines = mf->dlines;
    mf->d.data = mf->ddata;
    if (hipMemcpyAsync(mf->dlines, lines.data(), lines.size() * 8, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        (h.pop && hipMemcpyAsync(mf->ddata, data.data(), h.pop * 4, hipMemcpyHostToDevice, st) !=
                      hipSuccess)) {
        bbt_mibf_free(mf);
        return mfail(BBT_ERR_HIP, "bbt_mibf_load: upload failed");
    }
    s = mibf_stats(mf, st);
    if (s == BBT_OK)
        s = mibf_make_direct(mf, st);
    if (s != BBT_OK) {
        bbt_mibf_free(mf);
        return s;
    }
    *out = mf;
    return BBT_OK;
}

bbt_status bbt_mibf_summary_text(const bbt_mibf *mf, const uint64_t *summary, uint64_t nrec,
                                 char *buf, uint64_t cap, uint64_t *needed) {
    if (!mf || !summary)
        return mfail(BBT_ERR_ARG, "bbt_mibf_summary_text: bad argument");
    /* ResultsManager<ID>(m_fullIDs) rows: "repeat" and every ID name */
    const std::string s = summary_text(mf->names, summary, nrec);
    if (needed)
        *needed = s.size();
    if (buf && cap) {
        const size_t n = std::min<size_t>(s.size(), size_t(cap - 1));
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return BBT_OK;
}

bbt_status bbt_mibf_frame_probs(const bbt_mibf *mf, uint32_t allowed_miss, double *probs) {
    if (!mf || !probs)
        return mfail(BBT_ERR_ARG, "bbt_mibf_frame_probs: null argument");
    const uint32_t S = mf->d.S;
    const double occ = double(mf->pop) / double(mf->d.m);
    probs[0] = 0;
    for (uint32_t id = 1; id <= mf->d.nids; ++id) {
        const double f = mf->pop ? double(mf->id_slots[id]) / double(mf->pop) : 0.0;
        double P = 0;
        const uint32_t jmin = allowed_miss >= S ? 0 : S - allowed_miss;
        for (uint32_t j = jmin; j <= S; ++j) {
            double c = 1;
            for (uint32_t t = 0; t < j; ++t)
                c = c * double(S - t) / double(t + 1);
            P += c * std::pow(occ, double(j)) * std::pow(1 - occ, double(S - j)) *
                 (1 - std::pow(1 - f, double(j)));
        }
        probs[id] = P;
    }
    return BBT_OK;
}

uint32_t bbt_mibf_min_count(uint64_t frames, double prob, double score, uint32_t nids) {
    const uint64_t q = binom_isf_up(score / double(nids), frames, prob);
    return q > 1 ? uint32_t(q) : 1u;
}

bbt_status bbt_mibf_query_batch(bbt_mibf *mf, const bbt_reads *in, const bbt_mibf_params *p,
                                bbt_mibf_out *out, void *stream) {
    if (!mf || !in || !p || !out || !out->nhits || !out->hits || !out->sat || !out->counters)
        return mfail(BBT_ERR_ARG, "bbt_mibf_query_batch: bad argument");
    const bool paired = (p->flags & BBT_MIBF_PAIRED) != 0;
    if (paired && in->nreads % 2)
        return mfail(BBT_ERR_ARG, "bbt_mibf_query_batch: paired batch with an odd read count");
    const uint64_t nrec = paired ? in->nreads / 2 : in->nreads;
    if (nrec == 0)
        return BBT_OK;
    if (p->max_hits == 0 || p->max_hits > kMaxHitsOut)
        return mfail(BBT_ERR_ARG, "bbt_mibf_query_batch: max_hits must be 1..8192");
    if (nrec > 0xFFFFFFFFull)
        return mfail(BBT_ERR_ARG, "bbt_mibf_query_batch: more than 2^32 - 1 records in one batch");
    if (in->format > BBT_READS_2BIT || reinterpret_cast<uintptr_t>(in->seq) % 8 ||
        (in->format == BBT_READS_2BIT && in->seq_bytes % 8))
        return mfail(BBT_ERR_ARG, "bbt_mibf_query_batch: unknown read format, or 2-bit reads "
                                  "not 8-byte aligned");
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> g(mf->mu);
    const uint32_t k = mf->d.span;
    const uint32_t frames = in->max_len >= k ? in->max_len - k + 1 : 0;
    const uint64_t recs = uint64_t(frames) * mf->d.S * (paired ? 2 : 1);
    const bool direct = mf->d.nids + 1 <= kDirectIds;
    if (!direct && recs > (uint64_t(1) << 26))
        return mfail(BBT_ERR_UNSUPPORTED,
                     "bbt_mibf_query_batch: frames x hashes of the longest record above 2^26");
    if (uint64_t(frames + 1) * (mf->d.nids + 1) * 4 > (uint64_t(1) << 30))
        return mfail(BBT_ERR_UNSUPPORTED,
                     "bbt_mibf_query_batch: minimum-count table above 1 GiB (reads too long "
                     "for this many IDs)");
    const uint32_t miss = p->allowed_miss;
    /* the previous query (on any stream) may still read the buffers that are
     * refilled or replaced below */
    const auto wait_last = [&]() -> bbt_status {
        if (mf->has_last)
            MHIP(hipEventSynchronize(mf->last));
        return BBT_OK;
    };
    if (!mf->dmin || mf->min_score != p->score || mf->min_miss != miss ||
        mf->min_frames < frames) {
        if (wait_last() != BBT_OK)
            return BBT_ERR_HIP;
        /* getMinCount (MIBFClassifier.hpp:930-939) for every frame count up to
         * the longest read of the batch; rows already computed for the same
         * score and allowed misses are kept, the others are split over host
         * threads (one binomial quantile per frame count and ID) */
        const bool extend = mf->dmin && mf->min_score == p->score && mf->min_miss == miss;
        const uint32_t from = extend ? mf->min_frames + 1 : 1;
        std::vector<double> probs(mf->d.nids + 1);
        bbt_mibf_frame_probs(mf, miss, probs.data());
        const uint32_t cols = mf->d.nids + 1;
        std::vector<uint32_t> &t = mf->hmin;
        t.resize(size_t(frames + 1) * cols, 1);
        if (!extend)
            std::fill(t.begin(), t.begin() + cols, 1u);
        const uint32_t rows = frames + 1 - from;
        const unsigned nt = std::max(1u, std::min({16u, std::thread::hardware_concurrency(),
                                                   (rows * cols + 2047) / 2048}));
        const auto work = [&](unsigned w) {
            for (uint32_t fc = from + w; fc <= frames; fc += nt)
                for (uint32_t id = 1; id <= mf->d.nids; ++id)
                    t[size_t(fc) * cols + id] =
                        bbt_mibf_min_count(fc, probs[id], p->score, mf->d.nids);
        };
        std::vector<std::thread> th;
        for (unsigned w = 1; w < nt; ++w)
            th.emplace_back(work, w);
        work(0);
        for (std::thread &x : th)
            x.join();
        (void)hipFree(mf->dmin);
        mf->dmin = nullptr;
        MHIP(hipMalloc(&mf->dmin, t.size() * 4));
        MHIP(hipMemcpy(mf->dmin, t.data(), t.size() * 4, hipMemcpyHostToDevice));
        mf->min_score = p->score;
        mf->min_miss = miss;
        mf->min_frames = frames;
    }
    QueryArgs a;
    std::memset(&a, 0, sizeof(a));
    a.d = mf->d;
    if (p->flags & BBT_MIBF_LINES)
        a.d.direct = nullptr;
    a.seq = reinterpret_cast<const uint8_t *>(in->seq);
    a.off = in->off;
    a.nrec = nrec;
    a.min_count = mf->dmin;
    a.max_frames = mf->min_frames;
    a.allowed_miss = miss;
    a.multi_thresh = p->multi_thresh;
    a.min_nonsat = p->min_nonsat;
    a.max_hits = p->max_hits;
    a.paired = paired ? 1 : 0;
    a.best_agree = (p->flags & BBT_MIBF_BEST_AGREE) ? 1 : 0;
    a.nhits = out->nhits;
    a.hits = out->hits;
    a.sat = out->sat;
    a.counters = reinterpret_cast<unsigned long long *>(out->counters);
    a.queue = mf->dctl + 8;
    a.wide = mf->dctl + 9;
    a.wide_queue = mf->dctl + 10;
    /* with more IDs than kMaxCand a record may have more candidates than the
     * first pass keeps: those records go through the wide pass */
    a.spill_stride = mf->d.nids > kMaxCand ? mf->d.nids - kMaxCand : 0;
    if (a.spill_stride) {
        if (mf->wide_cap < nrec) {
            if (wait_last() != BBT_OK)
                return BBT_ERR_HIP;
            (void)hipFree(mf->dwide_list);
            mf->dwide_list = nullptr;
            mf->wide_cap = 0;
            MHIP(hipMalloc(&mf->dwide_list, nrec * sizeof(uint32_t)));
            mf->wide_cap = nrec;
        }
        if (!mf->dspill)
            MHIP(hipMalloc(&mf->dspill,
                           size_t(kWideBlocks) * a.spill_stride * kHitWords * sizeof(uint32_t)));
    }
    a.wide_list = mf->dwide_list;
    a.spill = mf->dspill;
    if (!direct && recs > 8192) {
        /* records too long for the LDS store: a device area per workgroup,
         * for as many workgroups as fit in about 1 GiB (at least the wide
         * pass's) */
        uint32_t cap = 1;
        while (cap < recs)
            cap <<= 1;
        const uint64_t per = uint64_t(cap) * 8;
        const uint64_t nblk = std::max<uint64_t>(kWideBlocks, (uint64_t(1) << 30) / per);
        if (mf->grec_cap != cap || mf->grec_blocks != nblk) {
            if (wait_last() != BBT_OK)
                return BBT_ERR_HIP;
            (void)hipFree(mf->dgrec);
            mf->dgrec = nullptr;
            mf->grec_cap = 0;
            mf->grec_blocks = 0;
            MHIP(hipMalloc(&mf->dgrec, nblk * per));
            mf->grec_cap = cap;
            mf->grec_blocks = uint32_t(nblk);
        }
        a.grec = mf->dgrec;
        a.grec_cap = cap;
        a.grec_blocks = mf->grec_blocks;
    }
    if (in->format == BBT_READS_2BIT) {
        const uint64_t need = 4 * in->seq_bytes + 8;
        if (mf->ascii_cap < need) {
            if (wait_last() != BBT_OK)
                return BBT_ERR_HIP;
            (void)hipFree(mf->dascii);
            mf->dascii = nullptr;
            mf->ascii_cap = 0;
            MHIP(hipMalloc(&mf->dascii, need));
            mf->ascii_cap = need;
        }
        a.seq = mf->dascii;
    }
    /* the queue counters are shared by every query of this filter: this
     * query's kernels start after the previous query's have finished */
    if (mf->has_last)
        MHIP(hipStreamWaitEvent(st, mf->last, 0));
    k_mibf_zero<<<1, 64, 0, st>>>(mf->dctl + 8, 3);
    if (in->format == BBT_READS_2BIT) {
        const uint64_t nw = in->seq_bytes / 8;
        k_unpack_2bit<<<uint32_t(std::min<uint64_t>((nw + 255) / 256, 65536)), 256, 0, st>>>(
            reinterpret_cast<const uint8_t *>(in->seq), in->seq_bytes, mf->dascii);
        MHIP(hipGetLastError());
    }
    const int blocks = mf->cu;
    const bool few = mf->d.S <= 4;
    /* per-ID totals and one row of minimum counts */
    const size_t shm = direct ? size_t(kDirectWords + 1) * (mf->d.nids + 1) * sizeof(uint32_t) : 0;
    if (a.d.direct)
        query_launch<true>(a, direct, recs, few, blocks, shm, st);
    else
        query_launch<false>(a, direct, recs, few, blocks, shm, st);
    MHIP(hipGetLastError());
    MHIP(hipEventRecord(mf->last, st));
    mf->has_last = true;
    return BBT_OK;
}

} // extern "C"
