/*
 * bbt_device.h -- the per-(read, filter) SeqEval walk as a lane state machine.
 *
 * Shared by the gfx950 kernels and by the host-side self test (the same code
 * compiled for the CPU), so the walk can be checked without a device.
 *
 * Reference behaviour restated here:
 *   ntHashIterator window order, N handling, canonical hash and multi-hash
 *       (btl_bloomfilter, absent; call site Common/SeqEval.h:50)
 *   BloomFilter::contains -- h bit tests, stop at the first clear bit
 *       (btl_bloomfilter, absent; call sites Common/SeqEval.h:54,79)
 *   evalSimple      Common/SeqEval.h:37-114
 *   evalHarmonic    Common/SeqEval.h:116-193
 *   evalBinomial    Common/SeqEval.h:218-290
 *   evalMinMatchLen Common/SeqEval.h:295-332
 *   eval*Score      Common/SeqEval.h:334-491
 *
 * The walk consumes one base per step (rolling the hash through N-runs from a
 * zeroed state), and stops whenever it needs a bit of the filter: the caller
 * issues that load and hands the byte back through consume().  Every lane
 * therefore has at most one filter load in flight, and performs exactly the
 * bit lookups the reference's contains() calls would, in the same order.
 */
#ifndef BBT_DEVICE_H
#define BBT_DEVICE_H

#include <math.h>
#include <stdint.h>

#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BBT_HD __host__ __device__ __forceinline__
#else
#define BBT_HD inline
#endif

namespace bbt {

/* ntHash v1 seeds (Mohamadi et al. 2016) and multi-hash constants. */
constexpr uint64_t kSeedA = 0x3c8bfbb395c60474ULL;
constexpr uint64_t kSeedC = 0x3193c18562a02b4cULL;
constexpr uint64_t kSeedG = 0x20323ed082572324ULL;
constexpr uint64_t kSeedT = 0x295549f54be24456ULL;
constexpr uint64_t kMultiSeed = 0x90b45d39fb6da1faULL;
constexpr unsigned kMultiShift = 27;
constexpr unsigned kMaxHash = 16;

enum Scorer : uint32_t { SIMPLE = 0, LENGTH = 1, HARMONIC = 2, BINOMIAL = 3 };

/* A load from device (global) memory.  Pointers that reach the walk through
 * descriptors or integer arithmetic are generic to the compiler, which then
 * emits flat loads; those also count against the LDS wait counter, so an LDS
 * read behind them would wait for the memory load. */
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ T ld_global(const T *p) {
    return *(const __attribute__((address_space(1))) T *)(p);
}
#else
template <class T>
BBT_HD T ld_global(const T *p) {
    return *p;
}
#endif

BBT_HD uint64_t rotl64(uint64_t v, unsigned s) {
    s &= 63;
    return s ? (v << s) | (v >> (64 - s)) : v;
}
BBT_HD uint64_t rotr64(uint64_t v, unsigned s) {
    s &= 63;
    return s ? (v >> s) | (v << (64 - s)) : v;
}
BBT_HD uint64_t rotl1(uint64_t v) { return (v << 1) | (v >> 63); }
BBT_HD uint64_t rotr1(uint64_t v) { return (v >> 1) | (v << 63); }

/* ASCII -> 2-bit code with (c >> 1) & 3: A 0, C 1, T 2, G 3 in either case;
 * the complement of a code is code ^ 2. */
BBT_HD bool base_valid(unsigned c) {
    /* (c & 0xDF) - 'A' is 0, 2, 6 or 19 for A, C, G, T in either case: one
     * bit test of 0x80045, branch-free (a chain of compares becomes branches) */
    const unsigned d = (c & 0xDFu) - unsigned('A');
    return (d < 20u) & ((0x80045u >> (d & 31u)) & 1u);
}
BBT_HD unsigned base_code(unsigned c) { return (c >> 1) & 3u; }

/* seed of a code, forward strand */
BBT_HD uint64_t seed_fwd(unsigned code) {
    return (code & 2) ? ((code & 1) ? kSeedG : kSeedT) : ((code & 1) ? kSeedC : kSeedA);
}
/* seed of the complement base */
BBT_HD uint64_t seed_rc(unsigned code) { return seed_fwd(code ^ 2u); }

/*
 * Prefix form of the window hashes, used where one wave hashes 64
 * consecutive windows at once.  With position-rotated contributions
 *   H_i = XOR_{j<=i} rotr(seed_fwd(c_j), j),  G_i = XOR_{j<=i} rotl(seed_rc(c_j), j)
 * (rotations mod 64), the window [p, i] with i = p + k - 1 hashes to
 *   fwd = rotl(H_i ^ H_{p-1}, i),  rc = rotr(G_i ^ G_{p-1}, p),
 * equal to XOR_j rotl(seed_fwd(c_j), i - j) and XOR_j rotl(seed_rc(c_j), j - p)
 * of the rolling form; both prefixes are plain XOR scans across lanes.
 */
BBT_HD uint64_t pfx_fwd(unsigned code, uint32_t i) { return rotr64(seed_fwd(code), i & 63); }
BBT_HD uint64_t pfx_rc(unsigned code, uint32_t i) { return rotl64(seed_rc(code), i & 63); }
BBT_HD uint64_t pfx_win_fwd(uint64_t Hi, uint64_t Hb, uint32_t i) { return rotl64(Hi ^ Hb, i & 63); }
BBT_HD uint64_t pfx_win_rc(uint64_t Gi, uint64_t Gb, uint32_t p) { return rotr64(Gi ^ Gb, p & 63); }

/*
 * Rolling ntHash terms of one base, precomputed for a fixed k.  Entry
 * e = (o << 2) | c, o and c codes, is the pair of terms when base c enters the
 * window and base o (k places back) leaves it:
 *   fwd' = rotl1(fwd) ^ rotl(seed_fwd(o), k) ^ seed_fwd(c)
 *   rc'  = rotr1(rc ^ seed_rc(o) ^ rotl(seed_rc(c), k));
 * entries 16 + c are the terms while the window fills (no base leaves).
 * Stored as t[2e] (forward term) and t[2e + 1] (reverse-complement term), so a
 * step is one 16-byte table read, two one-bit rotations and four XORs.
 */
constexpr unsigned kRollEntries = 20;

BBT_HD void roll_fill(uint64_t *t, unsigned k) {
    for (unsigned e = 0; e < kRollEntries; ++e) {
        const unsigned c = e & 3u;
        uint64_t tf = seed_fwd(c), tr = rotl64(seed_rc(c), k);
        if (e < 16) {
            const unsigned o = e >> 2;
            tf ^= rotl64(seed_fwd(o), k);
            tr ^= seed_rc(o);
        }
        t[2 * e] = tf;
        t[2 * e + 1] = tr;
    }
}

/* j-th value of ntHash's multi-hash expansion of the canonical hash h0:
 * t = h0 * (j ^ k*multiSeed); t ^ (t >> 27); j = 0 is h0 itself.  kseed is
 * k * multiSeed, a constant per filter. */
BBT_HD uint64_t multi_hash(uint64_t h0, uint32_t j, uint64_t kseed) {
    if (j == 0)
        return h0;
    const uint64_t t = h0 * (uint64_t(j) ^ kseed);
    return t ^ (t >> kMultiShift);
}

BBT_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return uint64_t((unsigned __int128)a * b >> 64);
#endif
}

/* Exact n % d for a fixed 64-bit divisor: Granlund & Montgomery, "Division
 * by invariant integers using multiplication" (PLDI 1994), Fig. 4.1. */
struct FastMod {
    uint64_t d;
    uint64_t magic;
    uint32_t sh1;
    uint32_t sh2;
};

inline FastMod fastmod_make(uint64_t d) {
    FastMod f;
    f.d = d;
    unsigned l = 0;
    while (l < 64 && (uint64_t(1) << l) < d)
        ++l; /* l = ceil(log2 d) */
    unsigned __int128 two_l = (unsigned __int128)1 << l;
    f.magic = uint64_t((((two_l - d) << 64) / d) + 1);
    f.sh1 = l < 1 ? l : 1;
    f.sh2 = l > 0 ? l - 1 : 0;
    return f;
}

BBT_HD uint64_t fastmod(uint64_t n, uint64_t d, uint64_t magic, uint32_t sh1,
                        uint32_t sh2) {
    uint64_t t1 = mulhi64(magic, n);
    uint64_t q = (t1 + ((n - t1) >> sh1)) >> sh2;
    return n - q * d;
}

/* Position of the n-th set bit (n >= 1) of x, or 64 when x has fewer. */
BBT_HD uint32_t nth_bit(uint64_t x, uint32_t n) {
    if (n == 0 || uint32_t(__builtin_popcountll(x)) < n)
        return 64;
    uint32_t pos = 0;
    for (uint32_t w = 32; w > 0; w >>= 1) {
        const uint32_t c = uint32_t(__builtin_popcountll(x & ((uint64_t(1) << w) - 1)));
        if (c < n) {
            n -= c;
            x >>= w;
            pos += w;
        }
    }
    return pos;
}

/* Device-resident description of one Bloom filter. */
struct FilterDev {
    const uint8_t *bits;     /* m/8 bytes, bit p = byte p>>3, mask 1<<(p&7) */
    uint64_t m;              /* bits */
    /* byte p>>3 is at bits[(p >> 3) << bshift]: 0 for a filter's own array;
     * the filters of one geometry interleaved byte by byte, filter f's bytes
     * at f, f + 2^bshift, ... (bits points at its first) */
    uint32_t bshift;
    uint64_t magic;          /* FastMod of m */
    uint32_t sh1, sh2;
    uint32_t h, k;
    double fpr;              /* (popcount/m)^h, for the binomial scorer */
    const uint32_t *binom;   /* calcMinCount per frame length, or null */
    uint32_t binom_len;      /* entries in binom */
    uint32_t bscore_frames;  /* frame lengths covered by bscore */
    /* BINOMIAL evalScore: -10 log10 P(X > c), X ~ Bin(fl, fpr), at
     * [fl (fl + 1) / 2 + c] for 1 <= fl <= bscore_frames, 0 <= c <= fl */
    const double *bscore;
    uint64_t mult[kMaxHash]; /* mult[j] = j ^ (k * kMultiSeed), j >= 1 */
    uint64_t roll[2 * kRollEntries]; /* roll_fill(roll, k) */
};

/* Per-walk parameters that are uniform over a launch. */
struct WalkParams {
    double threshold;  /* opt::score after CLI interpretation */
    uint32_t streak;   /* opt::streakThreshold */
    uint32_t min_len;  /* LENGTH: (unsigned) round(threshold) */
};

/* Walk outcome codes. */
enum : uint32_t { W_RUNNING = 0, W_ACCEPT = 1, W_REJECT = 2, W_DONE = 3 };

/*
 * Lane state machine for one (read, filter) walk.
 *   SC   scorer (Scorer)
 *   EXH  false: evalRead (early accept/reject); true: evalScore (exhaustive)
 *   K    compile-time k, or 0 for a runtime k
 */
template <int SC, bool EXH, int K>
struct Walker {
    /* read */
    const uint8_t *s;
    uint32_t len;
    uint32_t i; /* next base */
    const uint8_t *sw; /* s rounded down to 8 bytes */
    uint32_t mis;      /* s - sw */
    uint32_t wi;       /* index of word from sw */
    uint32_t lastw;    /* index of the word holding the read's last byte */
    uint64_t word;     /* 8 aligned bytes holding base i */
    uint64_t next;     /* the following 8 bytes, loaded one word ahead */
    /* rolling hash */
    uint64_t fwd, rc, win;
    uint32_t run;
    uint32_t k;
    const uint64_t *rt; /* the filter's roll terms for k */
    /* filter */
    const uint8_t *bits;
    uint32_t bshift; /* byte p>>3 at bits[(p >> 3) << bshift] */
    uint64_t m, magic;
    uint32_t sh1, sh2, h;
    /* probe in progress */
    uint64_t h0;
    uint32_t j;
    uint32_t pending;
    const uint8_t *addr;
    uint32_t mask;
    /* automaton */
    uint32_t p, prev;
    uint32_t state;
    uint32_t first;
    uint32_t skipping, skipEnd;
    uint32_t streak, anti, anti_thr;
    uint32_t score2;  /* SIMPLE: 2*score; BINOMIAL: count; LENGTH: matchLen */
    uint32_t thr2;    /* SIMPLE: ceil(2*thres); BINOMIAL: thres; LENGTH: minLen */
    double dscore;    /* HARMONIC score */
    double dthr;      /* HARMONIC thres */
    uint32_t streak_thr;
    /* instrumentation: contains() calls and bit lookups */
    uint32_t probes, lookups;
    /* set when a BINOMIAL threshold table does not cover the read */
    uint32_t bad;
    /* -D: low-complexity window starts, bit (dbase + p) of dm (null: no dust) */
    const uint32_t *dm;
    uint64_t dbase;

    BBT_HD unsigned kk() const { return K ? unsigned(K) : k; }

    BBT_HD void init(const uint8_t *seq, uint32_t length, const FilterDev *f,
                     const WalkParams &wp) {
        s = seq;
        len = length;
        i = 0;
        mis = uint32_t(reinterpret_cast<uintptr_t>(seq) & 7u);
        sw = seq - mis;
        wi = 0;
        lastw = length ? (length + mis - 1) >> 3 : 0;
        word = next = 0;
        if (length) {
            word = ld_global(reinterpret_cast<const uint64_t *>(sw));
            next = ld_global(reinterpret_cast<const uint64_t *>(sw + 8u * (lastw ? 1u : 0u)));
        }
        fwd = rc = win = 0;
        run = 0;
        rt = f->roll;
        bits = f->bits;
        bshift = f->bshift;
        m = f->m;
        magic = f->magic;
        sh1 = f->sh1;
        sh2 = f->sh2;
        h = f->h;
        k = f->k;
        pending = 0;
        prev = 0;
        first = 1;
        skipping = 0;
        skipEnd = 0;
        streak = 0;
        anti = 0;
        score2 = 0;
        dscore = 0.0;
        streak_thr = wp.streak;
        state = W_RUNNING;
        bad = 0;
        dm = nullptr;
        dbase = 0;
        const unsigned kv = kk();
        /* frame count as the size_t expression seqLen - k + 1 (SeqEval.h:28-35) */
        const uint64_t nframes = uint64_t(length) - uint64_t(kv) + 1;
        if (SC == SIMPLE || SC == HARMONIC) {
            const double n = double(nframes);
            const double thres = wp.threshold * n;
            const double at = floor((1.0 - wp.threshold) * n);
            /* a negative antiThres rejects at the first miss, as 0 does */
            anti_thr = at >= 4294967295.0 ? 0xFFFFFFFFu : (at <= 0 ? 0u : uint32_t(at));
            if (SC == SIMPLE) {
                const double t2 = ceil(2.0 * thres);
                thr2 = t2 >= 4294967295.0 ? 0xFFFFFFFFu : (t2 <= 0 ? 0u : uint32_t(t2));
            }
            dthr = thres;
        } else if (SC == BINOMIAL) {
            if (length < kv) {
                state = EXH ? W_DONE : W_REJECT;
                thr2 = 0;
                anti_thr = 0;
            } else {
                const uint32_t fl = uint32_t(nframes);
                if (EXH) {
                    thr2 = 1u;
                } else if (f->binom && fl < f->binom_len) {
                    thr2 = f->binom[fl];
                } else {
                    /* no threshold for this length: report, do not guess */
                    bad = 1;
                    state = W_REJECT;
                    thr2 = 1u;
                }
                anti_thr = fl - thr2; /* unsigned, as SeqEval.h:227 */
            }
        } else {
            thr2 = wp.min_len;
        }
    }

    /* byte idx of the read through an 8-byte aligned register window; idx
     * advances by one per call, so the word after the current one is loaded
     * a word ahead of its use.  Only words that hold bytes of the read are
     * loaded: past the last one the last is loaded again, so the load has no
     * alternative value and stays in flight until the word is used. */
    BBT_HD unsigned base_at(uint32_t idx) {
        const uint32_t a = idx + mis;
        const uint32_t w = a >> 3;
        if (w != wi) {
            wi = w;
            word = next;
            const uint32_t nw = w + 1 < lastw ? w + 1 : lastw;
            next = ld_global(reinterpret_cast<const uint64_t *>(sw + 8u * nw));
        }
        return unsigned(word >> ((a & 7u) * 8u)) & 0xFFu;
    }

    BBT_HD bool low_complexity(uint32_t pw) const {
        if (!dm)
            return false;
        const uint64_t g = dbase + pw;
        return (dm[g >> 5] >> (g & 31)) & 1u;
    }

    BBT_HD uint64_t kseed() const { return uint64_t(kk()) * kMultiSeed; }
    BBT_HD uint64_t hash_j(uint32_t jj) const { return multi_hash(h0, jj, kseed()); }

    BBT_HD void set_addr() {
        uint64_t pos = fastmod(hash_j(j), m, magic, sh1, sh2);
        addr = bits + ((pos >> 3) << bshift);
        mask = 1u << unsigned(pos & 7);
    }

    BBT_HD void finish_end() {
        if (EXH)
            state = W_DONE;
        else
            state = W_REJECT;
    }

    /* Rolls the base with 2-bit code into the window hash; true once the
     * window holds k bases. */
    BBT_HD bool roll_in(unsigned code) {
        const unsigned kv = kk();
        const bool full = run >= kv;
        const unsigned e = (full ? (unsigned(win >> (2 * (kv - 1))) & 3u) << 2 : 16u) | code;
        fwd = rotl1(fwd) ^ rt[2 * e];
        rc = rotr1(rc ^ rt[2 * e + 1]);
        run += full ? 0u : 1u;
        win = (win << 2) | code;
        return run >= kv;
    }

    /* One base of the walk, the flat form of advance(): true when the window
     * ending at this base needs a contains() call (h0 set, pending); the
     * caller then hands the call's bits back with take(), in hash order,
     * until pending clears.  Otherwise the base only rolled the hash, or
     * consumed a window without a probe, or the walk ended (state). */
    BBT_HD bool step() {
        if (state != W_RUNNING)
            return false;
        if (i >= len) {
            finish_end();
            return false;
        }
        uint32_t pw;
        uint64_t hv;
        if (!roll_step(pw, hv))
            return false;
        return decide(pw, hv);
    }

    /* The two halves of step(), for a caller that rolls the next base (and
     * issues its window's loads) before the pending call's bits are taken.
     * roll_step (i < len): one base in; true when it completes an N-free
     * window, whose start and canonical hash it returns.  It touches only the
     * rolling hash, never the automaton. */
    BBT_HD bool roll_step(uint32_t &pw, uint64_t &hv) {
        const unsigned c = base_at(i);
        ++i;
        if (!base_valid(c)) {
            run = 0;
            fwd = rc = 0;
            return false;
        }
        if (!roll_in(base_code(c)))
            return false;
        pw = i - kk();
        hv = rc < fwd ? rc : fwd;
        return true;
    }

    /* decide: window pw (hash hv) through the automaton, after every earlier
     * window's result; true when it needs a contains() call (pending). */
    BBT_HD bool decide(uint32_t pw, uint64_t hv) {
        if (window_pre(pw) != PRE_PROBE)
            return false;
        if (low_complexity(pw)) {
            on_result(false); /* -D: a miss without a contains() call */
            return false;
        }
        h0 = hv;
        j = 0;
        ++probes;
        pending = 1;
        return true;
    }

    /* The bit of hash j of the pending contains() call (step()). */
    BBT_HD void take(bool bit) {
        ++lookups;
        if (bit && j + 1 < h) {
            ++j;
            return;
        }
        pending = 0;
        on_result(bit);
    }

    /* Advance until a filter byte is needed (returns true, addr/mask set) or
     * the walk has ended (returns false). */
    BBT_HD bool advance() {
        if (pending)
            return true;
        if (state != W_RUNNING)
            return false;
        const unsigned kv = kk();
        while (true) {
            if (i >= len) {
                finish_end();
                return false;
            }
            const unsigned c = base_at(i);
            ++i;
            if (!base_valid(c)) {
                run = 0;
                fwd = rc = 0;
                continue;
            }
            if (!roll_in(base_code(c)))
                continue;
            p = i - kv; /* start of the N-free window ending at base i-1 */
            const int pre = window_pre(p);
            if (pre == PRE_END)
                return false;
            if (pre == PRE_NOPROBE)
                continue;
            if (low_complexity(p)) {
                /* -D: a miss without a contains() call (SeqEval.h:53) */
                on_result(false);
                if (state != W_RUNNING)
                    return false;
                continue;
            }
            /* contains(*itr) */
            h0 = rc < fwd ? rc : fwd;
            j = 0;
            ++probes;
            pending = 1;
            set_addr();
            return true;
        }
    }

    /* Window-level part of the walk, before contains(): returns PRE_PROBE
     * when window p must be probed, PRE_NOPROBE when it is consumed without
     * a probe (inside a jump-skip), PRE_END when the walk has ended. */
    enum { PRE_PROBE = 0, PRE_NOPROBE = 1, PRE_END = 2 };
    BBT_HD int window_pre(uint32_t pw) {
        p = pw;
        if (SC == LENGTH) {
            if (!EXH) {
                /* SeqEval.h:306: quit when the rest cannot reach minLen */
                if (uint64_t(len) - pw + score2 < uint64_t(thr2)) {
                    state = W_REJECT;
                    return PRE_END;
                }
            }
            if (pw != prev + 1)
                score2 = 0;
            return PRE_PROBE;
        }
        if (!EXH) {
            if (skipping) {
                if (pw < skipEnd) { /* SeqEval.h:103-108 */
                    if (anti_thr <= ++anti) {
                        state = W_REJECT;
                        return PRE_END;
                    }
                    prev = pw;
                    return PRE_NOPROBE;
                }
                skipping = 0;
            }
            if (!first && pw != prev + 1) { /* SeqEval.h:71-77 */
                anti += pw - prev - 1;
                if (anti_thr <= anti) {
                    state = W_REJECT;
                    return PRE_END;
                }
                streak = 0;
            }
            return PRE_PROBE;
        }
        if (skipping) {
            if (pw < skipEnd) {
                prev = pw;
                return PRE_NOPROBE;
            }
            skipping = 0;
        }
        if (pw != prev + 1)
            streak = 0;
        return PRE_PROBE;
    }

    /* Hand back the byte loaded from addr. */
    BBT_HD void consume(unsigned byte) {
        ++lookups;
        const bool bit = (byte & mask) != 0;
        if (bit && j + 1 < h) {
            ++j;
            set_addr();
            return;
        }
        pending = 0;
        on_result(bit);
    }

    BBT_HD void on_result(bool hit) {
        const unsigned kv = kk();
        if (SC == LENGTH) {
            score2 = hit ? (score2 == 0 ? kv : score2 + 1) : 0;
            if (!EXH && score2 >= thr2)
                state = W_ACCEPT;
            prev = p;
            return;
        }
        if (!EXH && first) { /* SeqEval.h:52-66 */
            first = 0;
            if (hit) {
                if (add_hit())
                    return;
                streak = 1;
            } else if (anti_thr <= ++anti) {
                state = W_REJECT;
                return;
            }
            prev = p;
            return;
        }
        if (hit) {
            if (add_hit())
                return;
            prev = p;
            ++streak;
            return;
        }
        if (streak < streak_thr) {
            if (!EXH && anti_thr <= ++anti) {
                state = W_REJECT;
                return;
            }
        } else {
            /* jump-skip: this window and the following ones below p + k
             * count as misses without probes (SeqEval.h:99-109) */
            skipping = 1;
            skipEnd = p + kv;
            if (!EXH && anti_thr <= ++anti) {
                state = W_REJECT;
                return;
            }
        }
        prev = p;
        streak = 0;
    }

    /* adds a hit to the score; true if the walk accepted */
    BBT_HD bool add_hit() {
        if (SC == HARMONIC) {
            if (streak == 0)
                dscore += 0.5;
            else
                dscore += 1.0 - 1.0 / (1.0 + double(streak));
            if (!EXH && dthr <= dscore) {
                state = W_ACCEPT;
                return true;
            }
            return false;
        }
        if (SC == BINOMIAL)
            score2 += 1;
        else
            score2 += streak == 0 ? 1u : 2u;
        if (!EXH && thr2 <= score2) {
            state = W_ACCEPT;
            return true;
        }
        return false;
    }

    /* The walk over one chunk of up to 64 consecutive windows whose contains()
     * answers are already known: lane L of the chunk is window
     * b0 + L - k + 1, vmask marks the N-free windows, hmask the windows whose
     * bits were all set, and lk(a, b) returns the bit lookups issued for lanes
     * [a, b) (the lookups a reference contains() makes: 1..h, stopping at the
     * first clear bit).  Three kinds of runs are consumed in one step each,
     * with the same effect as window_pre / on_result applied window by window:
     *   - windows inside a jump-skip (SeqEval.h:99-109): anti + 1 each, no probe;
     *   - contiguous misses while streak < streakThreshold (SeqEval.h:94-98):
     *     anti + 1 each, streak 0;
     *   - contiguous hits, SIMPLE and BINOMIAL (SeqEval.h:78-92, 236-262):
     *     +1 then +2 per hit (2 x score), or +1 per hit, accepting at the
     *     first window that reaches the threshold.
     * Everything else (the first window, gaps after N, the miss that starts a
     * jump-skip, LENGTH, HARMONIC hits) goes through the single-window
     * steps.  dmask marks the windows -D treats as misses without a
     * contains() call (their hmask bits are clear and lk counts nothing). */
    template <typename LK>
    BBT_HD void replay(uint32_t b0, uint64_t vmask, uint64_t hmask, uint64_t dmask,
                       const LK &lk) {
        const unsigned kv = kk();
        uint64_t vm = vmask;
        while (vm && state == W_RUNNING) {
            const uint32_t L = uint32_t(__builtin_ctzll(vm));
            const uint32_t pw = b0 + L - kv + 1;
            if (SC != LENGTH && (EXH || !first)) {
                if (skipping && pw < skipEnd) {
                    const uint32_t span = skipEnd - pw;
                    const uint64_t sm =
                        vm & (span >= 64 - L ? ~uint64_t(0) << L
                                             : ((uint64_t(1) << span) - 1) << L);
                    const uint32_t c = uint32_t(__builtin_popcountll(sm));
                    if (!EXH) {
                        if (anti_thr <= anti + c) {
                            state = W_REJECT;
                            return;
                        }
                        anti += c;
                    }
                    prev = b0 + uint32_t(63 - __builtin_clzll(sm)) - kv + 1;
                    vm &= ~sm;
                    continue;
                }
                skipping = 0;
                if (pw == prev + 1) {
                    const bool hit = (hmask >> L) & 1;
                    const uint64_t same = (hit ? hmask : ~hmask) & vm;
                    const uint64_t x = ~(same >> L);
                    const uint32_t r = x ? uint32_t(__builtin_ctzll(x)) : 64 - L;
                    const uint64_t run =
                        r >= 64 ? ~uint64_t(0) : ((uint64_t(1) << r) - 1) << L;
                    if ((SC == SIMPLE || SC == BINOMIAL) && !hit && streak < streak_thr &&
                        streak_thr >= 2) {
                        /* a miss, then isolated hits among misses: each hit
                         * follows a miss, so it scores as the first of a
                         * streak (SIMPLE +0.5, BINOMIAL +1), and the miss after
                         * it finds streak 1 < streakThreshold (SeqEval.h:78-98).
                         * Consumed in one step up to the first of two adjacent
                         * hits (included) or the end of the contiguous valid
                         * windows, stopping at the hit that accepts or the
                         * miss that rejects. */
                        const uint64_t cv = ~(vm >> L);
                        const uint32_t cl = cv ? uint32_t(__builtin_ctzll(cv)) : 64 - L;
                        const uint64_t cm = cl >= 64 ? ~uint64_t(0) : (uint64_t(1) << cl) - 1;
                        const uint64_t hm = (hmask >> L) & cm;
                        const uint64_t adj = hm & (hm >> 1);
                        const uint32_t R = adj ? uint32_t(__builtin_ctzll(adj)) + 1 : cl;
                        const uint64_t rm = R >= 64 ? ~uint64_t(0) : (uint64_t(1) << R) - 1;
                        const uint64_t hits_r = hm & rm, miss_r = ~hm & rm;
                        const uint64_t dust_r = (dmask >> L) & rm;
                        if (!EXH) {
                            const uint32_t nh = thr2 > score2 ? thr2 - score2 : 1u;
                            const uint32_t nm = anti_thr > anti ? anti_thr - anti : 1u;
                            const uint32_t ph = nth_bit(hits_r, nh), pm = nth_bit(miss_r, nm);
                            if (ph < 64 || pm < 64) {
                                const uint32_t e = ph < pm ? ph : pm;
                                const uint64_t em =
                                    e >= 63 ? ~uint64_t(0) : (uint64_t(1) << (e + 1)) - 1;
                                probes += e + 1 - uint32_t(__builtin_popcountll(dust_r & em));
                                lookups += lk(L, L + e + 1);
                                score2 += uint32_t(__builtin_popcountll(hits_r & em));
                                anti += uint32_t(__builtin_popcountll(miss_r & em));
                                state = ph < pm ? W_ACCEPT : W_REJECT;
                                return;
                            }
                            anti += uint32_t(__builtin_popcountll(miss_r));
                        }
                        score2 += uint32_t(__builtin_popcountll(hits_r));
                        probes += R - uint32_t(__builtin_popcountll(dust_r));
                        lookups += lk(L, L + R);
                        streak = uint32_t((hits_r >> (R - 1)) & 1);
                        prev = pw + R - 1;
                        vm &= ~(rm << L);
                        continue;
                    }
                    if (!hit && streak < streak_thr) {
                        if (!EXH && anti_thr <= anti + r) {
                            const uint32_t n = anti_thr > anti ? anti_thr - anti : 1u;
                            const uint64_t first_n =
                                n >= 64 ? run : run & (((uint64_t(1) << n) - 1) << L);
                            probes += n - uint32_t(__builtin_popcountll(first_n & dmask));
                            lookups += lk(L, L + n);
                            state = W_REJECT;
                            return;
                        }
                        if (!EXH)
                            anti += r;
                        probes += r - uint32_t(__builtin_popcountll(run & dmask));
                        lookups += lk(L, L + r);
                        streak = 0;
                        prev = pw + r - 1;
                        vm &= ~run;
                        continue;
                    }
                    if (hit && (SC == SIMPLE || SC == BINOMIAL)) {
                        /* score after n hits: SIMPLE 2n - (streak == 0), BINOMIAL n */
                        const uint32_t z = (SC == SIMPLE && streak == 0) ? 1u : 0u;
                        const uint32_t per = SC == SIMPLE ? 2u : 1u;
                        if (!EXH && uint64_t(score2) + uint64_t(per) * r - z >= thr2) {
                            const uint32_t d = thr2 - score2; /* >= 1: not accepted yet */
                            const uint32_t n = SC == SIMPLE ? (d + z + 1) / 2 : d;
                            score2 += per * n - z;
                            probes += n;
                            lookups += lk(L, L + n);
                            state = W_ACCEPT;
                            return;
                        }
                        score2 += per * r - z;
                        streak += r;
                        probes += r;
                        lookups += lk(L, L + r);
                        prev = pw + r - 1;
                        vm &= ~run;
                        continue;
                    }
                }
            }
            vm &= vm - 1;
            const int pre = window_pre(pw);
            if (pre == PRE_END)
                return;
            if (pre == PRE_NOPROBE)
                continue;
            if (!((dmask >> L) & 1)) { /* a dust window is a miss without a probe */
                ++probes;
                lookups += lk(L, L + 1);
            }
            on_result(((hmask >> L) & 1) != 0);
        }
    }

    /* Result of a finished walk: evalRead's bool, or evalScore's value (for
     * BINOMIAL, the hit count; the caller applies the binomial tail). */
    BBT_HD bool accepted() const { return state == W_ACCEPT; }
    BBT_HD double score_value() const {
        const double n = double(uint64_t(len) - uint64_t(kk()) + 1);
        if (SC == SIMPLE)
            return (double(score2) * 0.5) / n;
        if (SC == HARMONIC)
            return dscore / n;
        return double(score2);
    }
};

/* The contains() bits of one position for every filter of an interleaved
 * group: byte q of a 2^SH-byte group holds filter q's byte, bit b of it the
 * filter's bit; returns bit q = filter q's bit. */
template <unsigned SH>
BBT_HD uint32_t group_bits(const uint32_t *w, unsigned b) {
    uint32_t out = 0;
#pragma unroll
    for (unsigned i = 0; i < ((1u << SH) + 3) / 4; ++i) {
        const uint32_t x = (w[i] >> b) & 0x01010101u;
        out |= (((x * 0x01020408u) >> 24) & 0xFu) << (4 * i);
    }
    return (1u << SH) < 4 ? out & ((1u << (1u << SH)) - 1u) : out;
}

/*
 * One read against every filter of an interleaved group at once, SIMPLE
 * evalRead (SeqEval.h:37-114): the read's windows are hashed once (the
 * Walker's rolling hash, its own automaton unused), each of the h positions
 * of a window is one load of the group's 2^SH bytes, and each filter keeps
 * its own automaton -- the Walker's window_pre / on_result for SIMPLE,
 * restated over per-filter fields.  Reads shorter than 65536 bases (the
 * fields are 16-bit).  fetch(hv, j) returns the group bits of hash j.
 */
template <int K, unsigned MAXF>
struct RowsWalk {
    Walker<SIMPLE, false, K> w; /* hashing, thresholds, -D bitmap */
    uint32_t nf;
    uint32_t alive;             /* filters whose walk runs */
    uint32_t acc;               /* filters that accepted */
    uint32_t probes, lookups;   /* contains() calls and bit lookups, all filters */
    uint32_t issued;            /* group loads */
    /* per filter: prev | anti << 16; score2 | skipEnd << 16; streak | first << 30 | skipping << 31 */
    uint32_t pa[MAXF], ss[MAXF], st[MAXF];

    BBT_HD void init(const uint8_t *seq, uint32_t length, const FilterDev *f0, uint32_t n,
                     const WalkParams &wp) {
        w.init(seq, length, f0, wp);
        nf = n;
        alive = w.state == W_RUNNING ? (n >= 32 ? ~0u : (1u << n) - 1u) : 0u;
        acc = 0;
        probes = lookups = issued = 0;
#pragma unroll
        for (unsigned f = 0; f < MAXF; ++f) {
            pa[f] = 0;
            ss[f] = 0;
            st[f] = 1u << 30; /* first */
        }
    }

    /* window pw through filter f's automaton before contains():
     * 0 probe, 1 consumed without a probe, 2 the walk ended */
    template <unsigned F>
    BBT_HD int pre(uint32_t pw) {
        uint32_t prev = pa[F] & 0xFFFFu, anti = pa[F] >> 16;
        if (st[F] >> 31) { /* skipping */
            if (pw < (ss[F] >> 16)) {
                ++anti;
                pa[F] = pw | (anti << 16);
                if (w.anti_thr <= anti) {
                    alive &= ~(1u << F);
                    return 2;
                }
                return 1;
            }
            st[F] &= ~(1u << 31);
        }
        if (!((st[F] >> 30) & 1u) && pw != prev + 1) {
            anti += pw - prev - 1;
            pa[F] = prev | (anti << 16);
            if (w.anti_thr <= anti) {
                alive &= ~(1u << F);
                return 2;
            }
            st[F] &= 0xC0000000u; /* streak 0 */
        }
        return 0;
    }

    /* the contains() answer of window pw for filter f */
    template <unsigned F>
    BBT_HD void result(uint32_t pw, bool hit) {
        uint32_t anti = pa[F] >> 16, score2 = ss[F] & 0xFFFFu;
        uint32_t streak = st[F] & 0x3FFFFFFFu;
        const bool first = (st[F] >> 30) & 1u;
        if (hit) {
            score2 += streak == 0 ? 1u : 2u;
            ss[F] = (ss[F] & 0xFFFF0000u) | score2;
            if (w.thr2 <= score2) {
                alive &= ~(1u << F);
                acc |= 1u << F;
                return;
            }
            st[F] = (st[F] & 0x80000000u) | (first ? 1u : streak + 1u);
            pa[F] = pw | (anti << 16);
            return;
        }
        ++anti;
        if (!first && streak >= w.streak_thr) {
            /* jump-skip: windows below pw + k count as misses (SeqEval.h:99-109) */
            ss[F] = score2 | ((pw + w.kk()) << 16);
            st[F] = 0x80000000u;
        } else {
            st[F] = st[F] & 0x80000000u;
        }
        pa[F] = pw | (anti << 16);
        if (w.anti_thr <= anti)
            alive &= ~(1u << F);
    }

    template <unsigned F, class Fn>
    BBT_HD void each_alive(const Fn &fn) {
        if constexpr (F < MAXF) {
            if (F < nf && ((alive >> F) & 1u))
                fn(std::integral_constant<unsigned, F>());
            each_alive<F + 1, Fn>(fn);
        }
    }

    /* one base; false once every filter's walk has ended */
    template <class Fetch>
    BBT_HD bool step(const Fetch &fetch) {
        if (!alive)
            return false;
        if (w.i >= w.len) {
            alive = 0; /* the end of the read rejects the running walks */
            return false;
        }
        uint32_t pw;
        uint64_t hv;
        if (!w.roll_step(pw, hv))
            return true;
        const bool low = w.low_complexity(pw);
        uint32_t probe = 0;
        each_alive<0>([&](auto fc) {
            constexpr unsigned F = decltype(fc)::value;
            if (this->template pre<F>(pw) == 0) {
                if (low)
                    this->template result<F>(pw, false); /* -D: a miss, no contains() */
                else
                    probe |= 1u << F;
            }
        });
        if (!probe)
            return alive != 0;
        probes += uint32_t(__builtin_popcount(probe));
        uint32_t ones = probe; /* filters whose bits so far were all set */
        for (uint32_t j = 0; j < w.h && ones; ++j) {
            lookups += uint32_t(__builtin_popcount(ones));
            ++issued;
            ones &= fetch(hv, j);
        }
        each_alive<0>([&](auto fc) {
            constexpr unsigned F = decltype(fc)::value;
            if ((probe >> F) & 1u)
                this->template result<F>(pw, (ones >> F) & 1u);
        });
        return alive != 0;
    }
};

} // namespace bbt

#endif
