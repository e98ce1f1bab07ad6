/*
 * bbt_oracle.cpp -- CPU restatement of the BioBloom Tools read-classification
 * path.  TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); no
 * product code path uses it.  See bbt_oracle.h for the list of reference
 * files restated.
 *
 * Parity status per part (how each is pinned is in tests/test_oracle_*.py):
 *   ntHash constants, rotation variant and multi-hash: the btl_bloomfilter
 *     submodule is absent, so these follow the published ntHash v1 algorithm
 *     (plain 64-bit rotate, canonical = min(fwd, rc)).  Pinned by the closed
 *     form, reverse-complement symmetry, iterator positions vs brute force and
 *     the reference's own membership tests; the constant values themselves are
 *     "parity unpinned" against upstream.
 *   Spaced-seed hash: pinned by Tests/ntHashTests.cpp:32-40 symmetry and brute
 *     force masked windows; constants as above.
 *   Bloom contains/insert: pinned by brute force and Tests/BloomFilterTests.cpp
 *     membership assertions; bit order (LSB first) is "parity unpinned".
 *   SeqEval scorers and modes: restated from visible source; pinned by
 *     Tests/SeqEvalTests.cpp, closed forms and an independent brute-force
 *     automaton in the tests.
 *   Binomial: Loader's saddle-point pmf; pinned against scipy.stats.binom
 *     (which wraps Boost.Math); Boost's edge rounding of the quantile is
 *     "parity unpinned".
 *   Subtract filters (evalRead's subtract argument, loadFilterSubtract) and
 *     the progressive builders' one-thread loop: pinned by the independent
 *     SeqEval model fed both filters' answers, per-window composition of the
 *     pinned primitives, closed forms and the recruitment-order checks.
 */
#include "bbt_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ------------------------------------------------------------------ */
/* ntHash (published algorithm; upstream btl_bloomfilter/vendor/nthash.hpp) */

const uint64_t kSeedA = 0x3c8bfbb395c60474ULL;
const uint64_t kSeedC = 0x3193c18562a02b4cULL;
const uint64_t kSeedG = 0x20323ed082572324ULL;
const uint64_t kSeedT = 0x295549f54be24456ULL;
const uint64_t kMultiSeed = 0x90b45d39fb6da1faULL;
const unsigned kMultiShift = 27;

inline uint64_t rotl(uint64_t v, unsigned s) {
    s &= 63;
    return s ? (v << s) | (v >> (64 - s)) : v;
}
inline uint64_t rotr(uint64_t v, unsigned s) {
    s &= 63;
    return s ? (v >> s) | (v << (64 - s)) : v;
}

/* 0..3 for ACGT (either case), 4 for anything else ("N"). */
inline int base_code(unsigned char c) {
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}
const uint64_t kFwdSeed[4] = {kSeedA, kSeedC, kSeedG, kSeedT};
const uint64_t kRcSeed[4] = {kSeedT, kSeedG, kSeedC, kSeedA}; /* complement */

void multihash(uint64_t h0, unsigned k, unsigned h, uint64_t *out) {
    out[0] = h0;
    for (unsigned i = 1; i < h; ++i) {
        uint64_t t = h0 * (uint64_t(i) ^ (uint64_t(k) * kMultiSeed));
        t ^= t >> kMultiShift;
        out[i] = t;
    }
}

/* Iterator over N-free windows, visiting starts in increasing order; the
 * first window and every window after an N are hashed from scratch, the
 * others are rolled by one base.  Follows the contract the categorizer
 * relies on at Common/SeqEval.h:50-112 (pos(), *itr, ++itr, end()). */
class NtIter {
  public:
    NtIter(const char *s, uint64_t len, unsigned h, unsigned k)
        : m_s(reinterpret_cast<const unsigned char *>(s)), m_len(len),
          m_h(h ? h : 1), m_k(k), m_hv(m_h), m_pos(0), m_end(false) {
        seek();
    }
    bool end() const { return m_end; }
    uint64_t pos() const { return m_pos; }
    const uint64_t *hashes() const { return m_hv.data(); }
    uint64_t fwd() const { return m_fwd; }
    uint64_t rc() const { return m_rc; }
    void next() {
        ++m_pos;
        if (m_len < m_k || m_pos > m_len - m_k) {
            m_end = true;
            return;
        }
        int in = base_code(m_s[m_pos + m_k - 1]);
        if (in == 4) {
            m_pos += m_k; /* every window holding this base is skipped */
            seek();
            return;
        }
        int out = base_code(m_s[m_pos - 1]);
        m_fwd = rotl(m_fwd, 1) ^ rotl(kFwdSeed[out], m_k) ^ kFwdSeed[in];
        m_rc = rotr(m_rc ^ kRcSeed[out] ^ rotl(kRcSeed[in], m_k), 1);
        finish();
    }

  private:
    /* find the first N-free window starting at or after m_pos */
    void seek() {
        while (true) {
            if (m_len < m_k || m_pos > m_len - m_k) {
                m_end = true;
                return;
            }
            /* scan right to left so a hit jumps past the last N */
            uint64_t f = 0, r = 0;
            bool ok = true;
            for (int64_t i = int64_t(m_k) - 1; i >= 0; --i) {
                int c = base_code(m_s[m_pos + uint64_t(i)]);
                if (c == 4) {
                    m_pos += uint64_t(i) + 1;
                    ok = false;
                    break;
                }
                f ^= rotl(kFwdSeed[c], m_k - 1 - unsigned(i));
                r ^= rotl(kRcSeed[c], unsigned(i));
            }
            if (ok) {
                m_fwd = f;
                m_rc = r;
                finish();
                return;
            }
        }
    }
    void finish() {
        uint64_t canon = m_rc < m_fwd ? m_rc : m_fwd;
        multihash(canon, m_k, m_h, m_hv.data());
    }

    const unsigned char *m_s;
    uint64_t m_len;
    unsigned m_h, m_k;
    std::vector<uint64_t> m_hv;
    uint64_t m_pos;
    bool m_end;
    uint64_t m_fwd = 0, m_rc = 0;
};

/* ------------------------------------------------------------------ */
/* Bloom filter: bit p of the array is byte p>>3, mask 1<<(p&7). */

inline bool bf_contains(const orc_filter &f, const uint64_t *hv, uint64_t &lookups) {
    for (unsigned i = 0; i < f.h; ++i) {
        uint64_t p = hv[i] % f.m;
        ++lookups;
        if (!(f.bits[p >> 3] & (uint8_t(1) << (p & 7))))
            return false;
    }
    return true;
}

struct Counters {
    uint64_t probes = 0;
    uint64_t lookups = 0;
    /* the subtract filter of evalRead (SeqEval.h:38): a hit it also contains
     * scores nothing */
    const orc_filter *sub = nullptr;
    /* -D: the read's low-complexity intervals and the in-order query cursor
     * of SDust::isLowComp (Common/SDust.hpp:38-48) */
    const std::vector<uint32_t> *dust_start = nullptr, *dust_finish = nullptr;
    size_t dust_cur = 0;

    bool low_complexity(uint64_t pos) {
        const std::vector<uint32_t> &st = *dust_start, &fi = *dust_finish;
        while (dust_cur < st.size() && pos >= fi[dust_cur])
            ++dust_cur;
        return dust_cur < st.size() && pos >= st[dust_cur] && pos < fi[dust_cur];
    }
};

/* `!(sduster && sduster->isLowComp(pos)) && filter.contains(*itr)`
 * (SeqEval.h:53 and the same test in every scorer): a window starting in a
 * low-complexity interval is a miss without a contains() call. */
inline bool probe(const orc_filter &f, const NtIter &itr, Counters &c) {
    if (c.dust_start && c.low_complexity(itr.pos()))
        return false;
    ++c.probes;
    return bf_contains(f, itr.hashes(), c.lookups);
}

/* `subtract == NULL || !subtract->contains(*itr)` after a hit (SeqEval.h:55,
 * 81, 84, 316). */
inline bool scores(const NtIter &itr, const Counters &c) {
    if (!c.sub)
        return true;
    uint64_t l = 0;
    return !bf_contains(*c.sub, itr.hashes(), l);
}

/* With -D, evalRead / evalScore build an SDust of the read per call
 * (SeqEval.h:496-498, 537-539): the intervals of sdust(read, T, W). */
struct DustScope {
    Counters &c;
    std::vector<uint32_t> st, fi;
    DustScope(Counters &cc, const orc_eval_params &p, const char *s, uint64_t len) : c(cc) {
        if (!p.dust)
            return;
        st.resize(len + 1);
        fi.resize(len + 1);
        const int64_t n = orc_sdust(s, len, int(p.dust_t), int(p.dust_w), st.data(), fi.data(),
                                    len + 1);
        st.resize(size_t(n));
        fi.resize(size_t(n));
        c.dust_start = &st;
        c.dust_finish = &fi;
        c.dust_cur = 0;
    }
    ~DustScope() {
        c.dust_start = nullptr;
        c.dust_finish = nullptr;
    }
};

/* ------------------------------------------------------------------ */
/* SeqEval scorers (Common/SeqEval.h).  `n` below is the size_t expression
 * seqLen - kmerSize + 1 of denormalizeScore/normalizeScore (:28-35), which
 * wraps for reads shorter than k exactly as in the reference. */

inline double frames_as_double(uint64_t len, unsigned k) {
    return double(uint64_t(len - uint64_t(k) + 1));
}

/* evalSimple (:37-114) and evalHarmonic (:116-193): the same tiling walk,
 * differing only in the streak increment. */
bool eval_tiling(const orc_filter &f, const char *s, uint64_t len,
                 double threshold, unsigned streakThr, bool harmonic,
                 Counters &c) {
    const double thres = threshold * frames_as_double(len, f.k);
    const double antiThres = std::floor((1.0 - threshold) * frames_as_double(len, f.k));
    double score = 0;
    unsigned anti = 0, streak = 0;
    NtIter itr(s, len, std::max(f.k, f.h), f.k);
    uint64_t prev = 0;
    if (!itr.end()) {
        if (probe(f, itr, c)) {
            if (scores(itr, c))
                score += 0.5;
            if (thres <= score)
                return true;
            ++streak;
        } else if (antiThres <= ++anti) {
            return false;
        }
        prev = itr.pos();
        itr.next();
    }
    while (!itr.end()) {
        if (itr.pos() != prev + 1) {
            anti += unsigned(itr.pos() - prev - 1);
            if (antiThres <= anti)
                return false;
            streak = 0;
        }
        if (probe(f, itr, c)) {
            if (!scores(itr, c))
                ;
            else if (streak == 0)
                score += 0.5;
            else if (harmonic)
                score += 1.0 - 1.0 / (1.0 + double(streak));
            else
                score += 1;
            if (thres <= score)
                return true;
            prev = itr.pos();
            itr.next();
            ++streak;
        } else {
            if (streak < streakThr) {
                if (antiThres <= ++anti)
                    return false;
                prev = itr.pos();
                itr.next();
            } else {
                uint64_t skipEnd = itr.pos() + f.k;
                while (!itr.end() && itr.pos() < skipEnd) {
                    if (antiThres <= ++anti)
                        return false;
                    prev = itr.pos();
                    itr.next();
                }
            }
            streak = 0;
        }
    }
    return false;
}

/* evalBinomial (:218-290): integer hit count against a binomial quantile. */
bool eval_binomial(const orc_filter &f, const char *s, uint64_t len,
                   double minFPR, unsigned streakThr, Counters &c) {
    if (len < f.k)
        return false;
    const unsigned frameLen = unsigned(len - f.k + 1);
    const unsigned thres = unsigned(orc_calc_min_count(frameLen, f.fpr, minFPR));
    const unsigned antiThres = frameLen - thres;
    unsigned score = 0, anti = 0, streak = 0;
    NtIter itr(s, len, std::max(f.k, f.h), f.k);
    uint64_t prev = 0;
    if (!itr.end()) {
        if (probe(f, itr, c)) {
            ++score;
            if (thres <= score)
                return true;
            ++streak;
        } else if (antiThres <= ++anti) {
            return false;
        }
        prev = itr.pos();
        itr.next();
    }
    while (!itr.end()) {
        if (itr.pos() != prev + 1) {
            anti += unsigned(itr.pos() - prev - 1);
            if (antiThres <= anti)
                return false;
            streak = 0;
        }
        if (probe(f, itr, c)) {
            ++score;
            if (thres <= score)
                return true;
            prev = itr.pos();
            itr.next();
            ++streak;
        } else {
            if (streak < streakThr) {
                if (antiThres <= ++anti)
                    return false;
                prev = itr.pos();
                itr.next();
            } else {
                uint64_t skipEnd = itr.pos() + f.k;
                while (!itr.end() && itr.pos() < skipEnd) {
                    if (antiThres <= ++anti)
                        return false;
                    prev = itr.pos();
                    itr.next();
                }
            }
            streak = 0;
        }
    }
    return false;
}

/* evalMinMatchLen (:295-332): contiguous matched length. */
bool eval_min_match_len(const orc_filter &f, const char *s, uint64_t len,
                        unsigned minMatchLen, Counters &c) {
    unsigned matchLen = 0;
    NtIter itr(s, len, f.h, f.k);
    uint64_t prev = 0;
    while (!itr.end()) {
        if (len - itr.pos() + matchLen < minMatchLen)
            return false;
        if (itr.pos() != prev + 1)
            matchLen = 0;
        if (probe(f, itr, c)) {
            if (scores(itr, c))
                matchLen = matchLen == 0 ? f.k : matchLen + 1;
        } else {
            matchLen = 0;
        }
        if (matchLen >= minMatchLen)
            return true;
        prev = itr.pos();
        itr.next();
    }
    return false;
}

/* Exhaustive walks for evalScore (:334-491).  kind 0 simple, 2 harmonic,
 * 3 binomial count; LENGTH handled separately. */
double eval_tiling_score(const orc_filter &f, const char *s, uint64_t len,
                         unsigned streakThr, int kind, Counters &c,
                         unsigned *countOut) {
    double score = 0;
    unsigned count = 0, streak = 0;
    NtIter itr(s, len, std::max(f.k, f.h), f.k);
    uint64_t prev = 0;
    while (!itr.end()) {
        if (itr.pos() != prev + 1)
            streak = 0;
        if (probe(f, itr, c)) {
            ++count;
            if (streak == 0)
                score += 0.5;
            else if (kind == ORC_HARMONIC)
                score += 1.0 - 1.0 / (1.0 + double(streak));
            else
                score += 1;
            prev = itr.pos();
            itr.next();
            ++streak;
        } else {
            if (streak < streakThr) {
                prev = itr.pos();
                itr.next();
            } else {
                uint64_t skipEnd = itr.pos() + f.k;
                while (!itr.end() && itr.pos() < skipEnd) {
                    prev = itr.pos();
                    itr.next();
                }
            }
            streak = 0;
        }
    }
    if (countOut)
        *countOut = count;
    return score / frames_as_double(len, f.k);
}

unsigned eval_min_match_len_score(const orc_filter &f, const char *s,
                                  uint64_t len, Counters &c) {
    unsigned matchLen = 0;
    NtIter itr(s, len, f.h, f.k);
    uint64_t prev = 0;
    while (!itr.end()) {
        if (itr.pos() != prev + 1)
            matchLen = 0;
        if (probe(f, itr, c))
            matchLen = matchLen == 0 ? f.k : matchLen + 1;
        else
            matchLen = 0;
        prev = itr.pos();
        itr.next();
    }
    return matchLen;
}

/* evalRead dispatcher (:493-529); -D (dust) is not part of this path. */
bool eval_read(const orc_filter &f, const char *s, uint64_t len,
               const orc_eval_params &p, Counters &c) {
    DustScope dust(c, p, s, len);
    switch (p.scorer) {
    case ORC_LENGTH:
        return eval_min_match_len(f, s, len, unsigned(std::round(p.threshold)), c);
    case ORC_HARMONIC:
        return eval_tiling(f, s, len, p.threshold, p.streak, true, c);
    case ORC_BINOMIAL:
        return eval_binomial(f, s, len, p.threshold, p.streak, c);
    default:
        return eval_tiling(f, s, len, p.threshold, p.streak, false, c);
    }
}

/* evalScore dispatcher (:534-564). */
double eval_score(const orc_filter &f, const char *s, uint64_t len,
                  const orc_eval_params &p, Counters &c) {
    DustScope dust(c, p, s, len);
    switch (p.scorer) {
    case ORC_LENGTH:
        return double(eval_min_match_len_score(f, s, len, c));
    case ORC_HARMONIC:
        return eval_tiling_score(f, s, len, p.streak, ORC_HARMONIC, c, nullptr);
    case ORC_BINOMIAL: {
        /* evalBinomialScore (:450-491) then log10(.) * -10 */
        double tail;
        if (len < f.k) {
            tail = 1.0;
        } else {
            unsigned count = 0;
            eval_tiling_score(f, s, len, p.streak, ORC_BINOMIAL, c, &count);
            tail = orc_binom_sf(int64_t(count), len - f.k + 1, f.fpr);
        }
        return std::log10(tail) * -10;
    }
    default:
        return eval_tiling_score(f, s, len, p.streak, ORC_SIMPLE, c, nullptr);
    }
}

/* ------------------------------------------------------------------ */
/* Loader's saddle-point binomial pmf (C. Loader, "Fast and accurate
 * computation of binomial probabilities", 2000). */

double stirlerr(double n) {
    const double S0 = 1.0 / 12, S1 = 1.0 / 360, S2 = 1.0 / 1260,
                 S3 = 1.0 / 1680, S4 = 1.0 / 1188;
    if (n <= 15.0) {
        long double ln = (long double)n;
        long double v = lgammal(ln + 1.0L) - (ln + 0.5L) * logl(ln) + ln -
                        0.5L * logl(2.0L * 3.14159265358979323846264338327950288L);
        return double(v);
    }
    double nn = n * n;
    if (n > 500)
        return (S0 - S1 / nn) / n;
    if (n > 80)
        return (S0 - (S1 - S2 / nn) / nn) / n;
    if (n > 35)
        return (S0 - (S1 - (S2 - S3 / nn) / nn) / nn) / n;
    return (S0 - (S1 - (S2 - (S3 - S4 / nn) / nn) / nn) / nn) / n;
}

double bd0(double x, double np) {
    if (std::fabs(x - np) < 0.1 * (x + np)) {
        double v = (x - np) / (x + np);
        double s = (x - np) * v;
        double ej = 2 * x * v;
        v = v * v;
        for (int j = 1; j < 1000; ++j) {
            ej *= v;
            double s1 = s + ej / (2 * j + 1);
            if (s1 == s)
                return s1;
            s = s1;
        }
    }
    return x * std::log(x / np) + np - x;
}

double dbinom(double x, double n, double p) {
    double q = 1.0 - p;
    if (p == 0)
        return x == 0 ? 1 : 0;
    if (q == 0)
        return x == n ? 1 : 0;
    if (x == 0) {
        if (n == 0)
            return 1;
        double lc = p < 0.1 ? -bd0(n, n * q) - n * p : n * std::log(q);
        return std::exp(lc);
    }
    if (x == n) {
        double lc = q < 0.1 ? -bd0(n, n * p) - n * q : n * std::log(p);
        return std::exp(lc);
    }
    if (x < 0 || x > n)
        return 0;
    double lc = stirlerr(n) - stirlerr(x) - stirlerr(n - x) - bd0(x, n * p) -
                bd0(n - x, n * q);
    double lf = std::log(2 * 3.14159265358979323846) + std::log(x) + std::log1p(-x / n);
    return std::exp(lc - 0.5 * lf);
}

/* ------------------------------------------------------------------ */
/* ResultsManager (ResultsManager.hpp:66-189) on per-read hit lists. */

struct Summary {
    explicit Summary(unsigned nf) : nf(nf), above(nf + 2, 0), unique(nf + 2, 0) {}
    unsigned nf;
    std::vector<uint64_t> above, unique;
    uint64_t multi = 0, none = 0;

    unsigned note(unsigned idx, unsigned cur) {
        ++above[idx];
        return cur == nf ? idx : nf + 1;
    }
    unsigned finish(unsigned cls) {
        if (cls == nf)
            ++none;
        else if (cls == nf + 1)
            ++multi;
        else
            ++unique[cls];
        return cls;
    }
    unsigned single(const std::vector<unsigned> &hits) {
        unsigned cls = nf;
        for (unsigned id : hits)
            cls = note(id, cls);
        return finish(cls);
    }
    unsigned pair(const std::vector<unsigned> &h1, const std::vector<unsigned> &h2,
                  bool inclusive) {
        unsigned cls = nf;
        size_t i = 0, j = 0;
        if (inclusive) {
            while (i < h1.size() && j < h2.size()) {
                if (h1[i] == h2[j]) {
                    cls = note(h1[i], cls);
                    ++i;
                    ++j;
                } else if (h1[i] < h2[j]) {
                    cls = note(h1[i++], cls);
                } else {
                    cls = note(h2[j++], cls);
                }
            }
            while (i < h1.size())
                cls = note(h1[i++], cls);
            while (j < h2.size())
                cls = note(h2[j++], cls);
        } else {
            while (i < h1.size() && j < h2.size()) {
                if (h1[i] == h2[j]) {
                    cls = note(h1[i], cls);
                    ++i;
                    ++j;
                } else if (h1[i] < h2[j]) {
                    ++i;
                } else {
                    ++j;
                }
            }
        }
        return finish(cls);
    }
};

std::vector<unsigned> hit_list(const uint8_t *row, unsigned nf) {
    std::vector<unsigned> v;
    for (unsigned f = 0; f < nf; ++f)
        if (row[f])
            v.push_back(f);
    return v;
}

/* Per-read hits/scores for one read under a mode (BioBloomClassifier.cpp:
 * 1145-1237); ORDERED single reads only (pairs handled by the caller). */
void evaluate_read(const orc_filter *fs, unsigned nf, const char *s,
                   uint64_t len, uint32_t mode, const orc_eval_params &p,
                   uint8_t *hrow, double *srow, Counters &c) {
    switch (mode) {
    case ORC_ORDERED:
        for (unsigned i = 0; i < nf; ++i)
            if (eval_read(fs[i], s, len, p, c)) {
                hrow[i] = 1;
                break;
            }
        break;
    case ORC_BESTHIT: {
        double maxScore = 0;
        std::vector<unsigned> best;
        for (unsigned i = 0; i < nf; ++i) {
            double sc = eval_score(fs[i], s, len, p, c);
            if (srow)
                srow[i] = sc;
            if (maxScore < sc) {
                maxScore = sc;
                best.assign(1, i);
            } else if (maxScore == sc) {
                best.push_back(i);
            }
        }
        if (maxScore > 0)
            for (unsigned i : best)
                hrow[i] = 1;
        break;
    }
    case ORC_SCORES:
        for (unsigned i = 0; i < nf; ++i)
            hrow[i] = eval_read(fs[i], s, len, p, c) ? 1 : 0;
        for (unsigned i = 0; i < nf; ++i) {
            double sc = eval_score(fs[i], s, len, p, c);
            if (srow)
                srow[i] = sc;
        }
        break;
    default:
        for (unsigned i = 0; i < nf; ++i)
            hrow[i] = eval_read(fs[i], s, len, p, c) ? 1 : 0;
        break;
    }
}

} // namespace

/* ==================================================================== */
extern "C" {

int64_t orc_nthash(const char *seq, uint64_t len, uint32_t k, uint32_t h,
                   uint64_t *pos_out, uint64_t *hash_out, uint64_t cap) {
    int64_t n = 0;
    for (NtIter it(seq, len, h, k); !it.end(); it.next()) {
        if (uint64_t(n) >= cap)
            return -1;
        pos_out[n] = it.pos();
        std::memcpy(hash_out + uint64_t(n) * h, it.hashes(), sizeof(uint64_t) * h);
        ++n;
    }
    return n;
}

void orc_nthash_window(const char *win, uint32_t k, uint64_t *fwd, uint64_t *rc) {
    /* closed form: fwd = XOR_j rol(seed(s_j), k-1-j), rc = XOR_j rol(seed(comp s_j), j) */
    uint64_t f = 0, r = 0;
    for (uint32_t j = 0; j < k; ++j) {
        int c = base_code((unsigned char)win[j]);
        if (c == 4)
            c = 0; /* caller passes N-free windows */
        f ^= rotl(kFwdSeed[c], k - 1 - j);
        r ^= rotl(kRcSeed[c], j);
    }
    *fwd = f;
    *rc = r;
}

void orc_multihash(uint64_t h0, uint32_t k, uint32_t h, uint64_t *out) {
    multihash(h0, k, h, out);
}

int64_t orc_sthash(const char *seq, uint64_t len, const char *seeds,
                   uint32_t nseeds, uint32_t k, uint64_t *pos_out,
                   uint64_t *hash_out, uint64_t cap) {
    /* Spaced-seed ntHash (upstream stHashIterator, absent): per window, the
     * full fwd/rc values with the "don't care" positions' contributions
     * XORed back out, then canonical min per seed. */
    std::vector<std::vector<unsigned>> dontCare(nseeds);
    for (uint32_t j = 0; j < nseeds; ++j)
        for (uint32_t i = 0; i < k; ++i)
            if (seeds[j * k + i] != '1')
                dontCare[j].push_back(i);
    int64_t n = 0;
    const unsigned char *s = reinterpret_cast<const unsigned char *>(seq);
    for (NtIter it(seq, len, 1, k); !it.end(); it.next()) {
        if (uint64_t(n) >= cap)
            return -1;
        pos_out[n] = it.pos();
        for (uint32_t j = 0; j < nseeds; ++j) {
            uint64_t f = it.fwd(), r = it.rc();
            for (unsigned i : dontCare[j]) {
                int c = base_code(s[it.pos() + i]);
                f ^= rotl(kFwdSeed[c], k - 1 - i);
                r ^= rotl(kRcSeed[c], i);
            }
            hash_out[uint64_t(n) * nseeds + j] = r < f ? r : f;
        }
        ++n;
    }
    return n;
}

int orc_bf_contains(const uint8_t *bits, uint64_t m, uint32_t h,
                    const uint64_t *hashes, uint32_t *lookups) {
    orc_filter f{bits, m, h, 0, 0};
    uint64_t l = 0;
    bool r = bf_contains(f, hashes, l);
    if (lookups)
        *lookups = uint32_t(l);
    return r ? 1 : 0;
}

int orc_bf_insert(uint8_t *bits, uint64_t m, uint32_t h, const uint64_t *hashes) {
    bool all = true;
    for (uint32_t i = 0; i < h; ++i) {
        uint64_t p = hashes[i] % m;
        uint8_t mask = uint8_t(1) << (p & 7);
        if (!(bits[p >> 3] & mask)) {
            all = false;
            bits[p >> 3] |= mask;
        }
    }
    return all ? 1 : 0;
}

void orc_bf_insert_seq(uint8_t *bits, uint64_t m, uint32_t h, uint32_t k,
                       const char *seq, uint64_t len, uint64_t *n_new,
                       uint64_t *n_redundant) {
    /* BloomFilterGenerator::loadFilter (BloomFilterGenerator.h:143-188) */
    for (NtIter it(seq, len, h, k); !it.end(); it.next()) {
        if (orc_bf_insert(bits, m, h, it.hashes()))
            ++*n_redundant;
        else
            ++*n_new;
    }
}

void orc_bf_insert_seq_subtract(uint8_t *bits, uint64_t m, uint32_t h, uint32_t k,
                                const uint8_t *sub_bits, uint64_t sub_m, const char *seq,
                                uint64_t len, uint64_t *n_new, uint64_t *n_redundant,
                                uint64_t *n_removed) {
    /* BloomFilterGenerator::loadFilterSubtract (BloomFilterGenerator.h:238-270):
     * a k-mer the subtract filter contains is counted and not inserted */
    orc_filter sub{sub_bits, sub_m, h, k, 0};
    for (NtIter it(seq, len, h, k); !it.end(); it.next()) {
        uint64_t l = 0;
        if (bf_contains(sub, it.hashes(), l)) {
            ++*n_removed;
            continue;
        }
        if (orc_bf_insert(bits, m, h, it.hashes()))
            ++*n_redundant;
        else
            ++*n_new;
    }
}

uint64_t orc_popcount(const uint8_t *bits, uint64_t nbytes) {
    uint64_t c = 0;
    for (uint64_t i = 0; i < nbytes; ++i)
        c += uint64_t(__builtin_popcount(bits[i]));
    return c;
}

uint64_t orc_calc_optimal_size(uint64_t entries, double fpr, uint32_t h) {
    /* BloomFilterInfo::calcOptimalSize (BloomFilterInfo.h:57-65) */
    uint64_t v = uint64_t(-double(entries) * double(h) /
                          std::log(1.0 - std::pow(fpr, double(1 / double(h)))));
    return v + (64 - v % 64);
}

uint32_t orc_calc_optimal_hashnum(double fpr) {
    /* BloomFilterInfo::calcOptimalHashNum (BloomFilterInfo.h:73-76) */
    return unsigned(-std::log(fpr) / std::log(2));
}

double orc_binom_pmf(uint64_t x, uint64_t n, double p) {
    return dbinom(double(x), double(n), p);
}

double orc_binom_sf(int64_t x, uint64_t n, double p) {
    /* P(X > x) for X ~ Bin(n, p); summed over whichever side is the tail */
    if (x < 0)
        return 1.0;
    if (uint64_t(x) >= n)
        return 0.0;
    double mean = double(n) * p;
    double sum = 0;
    if (double(x) + 1 >= mean) {
        for (uint64_t i = uint64_t(x) + 1; i <= n; ++i) {
            double t = dbinom(double(i), double(n), p);
            sum += t;
            if (t < sum * 1e-18 && double(i) > mean)
                break;
        }
        return sum;
    }
    for (int64_t i = x; i >= 0; --i) {
        double t = dbinom(double(i), double(n), p);
        sum += t;
        if (t < sum * 1e-18 && double(i) < mean)
            break;
    }
    return 1.0 - sum;
}

uint64_t orc_binom_isf_up(double q, uint64_t n, double p) {
    /* smallest integer x with P(X > x) <= q (Boost discrete quantile of the
     * complement with integer_round_up, SeqEval.h:199-207) */
    if (q >= 1.0)
        return 0;
    double mean = double(n) * p;
    double sd = std::sqrt(mean * (1 - p));
    double top = std::ceil(mean + 40 * sd + 40);
    uint64_t x = top >= double(n) ? n : uint64_t(top);
    double s = orc_binom_sf(int64_t(x), n, p); /* sf(x) */
    while (x > 0) {
        double prev = s + dbinom(double(x), double(n), p); /* sf(x-1) */
        if (prev > q)
            return x;
        s = prev;
        --x;
    }
    return 0;
}

uint64_t orc_calc_min_count(uint64_t frame_len, double bf_fpr, double min_fpr) {
    /* SeqEval::calcMinCount (SeqEval.h:199-207): quantile, at least 1 */
    unsigned i = unsigned(orc_binom_isf_up(min_fpr, frame_len, bf_fpr));
    return i > 1 ? i : 1;
}

int orc_eval_read(const orc_filter *f, const char *seq, uint64_t len,
                  const orc_eval_params *p, uint64_t *probes, uint64_t *lookups) {
    Counters c;
    bool r = eval_read(*f, seq, len, *p, c);
    if (probes)
        *probes += c.probes;
    if (lookups)
        *lookups += c.lookups;
    return r ? 1 : 0;
}

int orc_eval_read_sub(const orc_filter *f, const orc_filter *sub, const char *seq,
                      uint64_t len, const orc_eval_params *p) {
    Counters c;
    c.sub = sub;
    return eval_read(*f, seq, len, *p, c) ? 1 : 0;
}

int orc_progressive(uint8_t *bits, uint64_t m, uint32_t h, uint32_t k, const orc_filter *bait,
                    const orc_filter *sub, const char *seq, const uint64_t *off,
                    uint64_t npairs, const orc_prog_params *p, uint64_t *total_entries,
                    uint64_t *total_reads, uint64_t *tagged) {
    /* BloomFilterGenerator::generateProgressive (BloomFilterGenerator.cpp:
     * 117-274) and generateProgressiveBait (:343-509) over one pair of read
     * files, as one thread runs them: per pair, count it, stop once the
     * filter holds the expected k-mers, else tag the pair when its reads pass
     * (PROG_STD: both; PROG_INC: either, if it has more k-mers than the score)
     * against the growing filter or the fixed bait filter, and insert both
     * reads of a tagged pair. */
    orc_filter f{bits, m, h, k, 0};
    orc_eval_params ep;
    std::memset(&ep, 0, sizeof(ep));
    ep.scorer = p->scorer;
    ep.threshold = p->score;
    ep.streak = p->streak;
    orc_eval_params bp = ep;
    bp.threshold = p->bait;
    const auto passes = [&](const char *s, uint64_t len) {
        Counters c;
        c.sub = sub;
        if (eval_read(f, s, len, ep, c))
            return true;
        if (!bait)
            return false;
        Counters cb;
        cb.sub = sub;
        return eval_read(*bait, s, len, bp, cb);
    };
    if (p->single) {
        /* generateProgressive over single-end files (BloomFilterGenerator.cpp:
         * 1038-1170): reads are counted only while the filter is short of the
         * expected k-mers, and a read is tagged alone */
        if (*total_entries >= p->expected)
            return 0;
        for (uint32_t it = 0; it < p->iterations; ++it) {
            for (uint64_t i = 0; i < npairs; ++i) {
                if (*total_entries >= p->expected)
                    break;
                ++*total_reads;
                const char *s1 = seq + off[i];
                const uint64_t l1 = off[i + 1] - off[i];
                const uint64_t nk1 = l1 > k ? l1 - k + 1 : 0;
                Counters c;
                c.sub = sub;
                if (!(double(nk1) > p->score && eval_read(f, s1, l1, ep, c)))
                    continue;
                ++*tagged;
                uint64_t n = 0, r = 0;
                orc_bf_insert_seq(bits, m, h, k, s1, l1, &n, &r);
                *total_entries += n;
            }
        }
        return 0;
    }
    for (uint32_t it = 0; it < p->iterations; ++it) {
        for (uint64_t i = 0; i < npairs; ++i) {
            ++*total_reads;
            if (*total_entries >= p->expected)
                break;
            const char *s1 = seq + off[2 * i], *s2 = seq + off[2 * i + 1];
            const uint64_t l1 = off[2 * i + 1] - off[2 * i], l2 = off[2 * i + 2] - off[2 * i + 1];
            const uint64_t nk1 = l1 > k ? l1 - k + 1 : 0, nk2 = l2 > k ? l2 - k + 1 : 0;
            bool tag;
            if (p->inclusive)
                tag = (double(nk1) > p->score && passes(s1, l1)) ||
                      (double(nk2) > p->score && passes(s2, l2));
            else
                tag = passes(s1, l1) && passes(s2, l2);
            if (!tag)
                continue;
            ++*tagged;
            uint64_t n = 0, r = 0;
            orc_bf_insert_seq(bits, m, h, k, s1, l1, &n, &r);
            orc_bf_insert_seq(bits, m, h, k, s2, l2, &n, &r);
            *total_entries += n;
        }
    }
    return 0;
}

double orc_eval_score(const orc_filter *f, const char *seq, uint64_t len,
                      const orc_eval_params *p, uint64_t *probes,
                      uint64_t *lookups) {
    Counters c;
    double r = eval_score(*f, seq, len, *p, c);
    if (probes)
        *probes += c.probes;
    if (lookups)
        *lookups += c.lookups;
    return r;
}

int orc_classify(const orc_filter *fs, uint32_t nf, const char *seq,
                 const uint64_t *off, uint64_t nreads, int paired, int inclusive,
                 uint32_t mode, const orc_eval_params *p, int nthreads,
                 uint8_t *hits, double *scores, uint32_t *cls,
                 uint64_t *counters) {
    if (paired && (nreads & 1))
        return 1;
    const uint64_t nrec = paired ? nreads / 2 : nreads;
    std::vector<uint8_t> hbuf;
    if (!hits) {
        hbuf.assign(nreads * nf, 0);
        hits = hbuf.data();
    } else {
        std::memset(hits, 0, nreads * nf);
    }
    uint64_t probes = 0, lookups = 0;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : probes, lookups)
    {
        Counters c;
        Summary local(nf);
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < int64_t(nrec); ++r) {
            if (!paired) {
                uint8_t *hrow = hits + uint64_t(r) * nf;
                double *srow = scores ? scores + uint64_t(r) * nf : nullptr;
                evaluate_read(fs, nf, seq + off[r], off[r + 1] - off[r], mode, *p,
                              hrow, srow, c);
                if (cls)
                    cls[r] = local.single(hit_list(hrow, nf));
                continue;
            }
            uint64_t a = 2 * uint64_t(r), b = a + 1;
            uint8_t *h1 = hits + a * nf, *h2 = hits + b * nf;
            const char *s1 = seq + off[a], *s2 = seq + off[b];
            uint64_t l1 = off[a + 1] - off[a], l2 = off[b + 1] - off[b];
            if (mode == ORC_ORDERED) {
                /* evaluateReadOrderedPair (BioBloomClassifier.cpp:1159-1180) */
                for (unsigned i = 0; i < nf; ++i) {
                    bool pass;
                    if (inclusive)
                        pass = eval_read(fs[i], s1, l1, *p, c) ||
                               eval_read(fs[i], s2, l2, *p, c);
                    else
                        pass = eval_read(fs[i], s1, l1, *p, c) &&
                               eval_read(fs[i], s2, l2, *p, c);
                    if (pass) {
                        h1[i] = 1;
                        h2[i] = 1;
                        break;
                    }
                }
            } else {
                evaluate_read(fs, nf, s1, l1, mode, *p, h1,
                              scores ? scores + a * nf : nullptr, c);
                evaluate_read(fs, nf, s2, l2, mode, *p, h2,
                              scores ? scores + b * nf : nullptr, c);
            }
            if (cls)
                cls[r] = local.pair(hit_list(h1, nf), hit_list(h2, nf), inclusive != 0);
        }
        probes += c.probes;
        lookups += c.lookups;
    }
    if (counters) {
        counters[0] = probes;
        counters[1] = lookups;
    }
    return 0;
}

uint64_t orc_summary(const char *const *filter_ids, uint32_t nf,
                     const uint8_t *hits, uint64_t nreads, int paired,
                     int inclusive, char *out, uint64_t cap) {
    Summary s(nf);
    uint64_t nrec = paired ? nreads / 2 : nreads;
    for (uint64_t r = 0; r < nrec; ++r) {
        if (paired)
            s.pair(hit_list(hits + 2 * r * nf, nf), hit_list(hits + (2 * r + 1) * nf, nf),
                   inclusive != 0);
        else
            s.single(hit_list(hits + r * nf, nf));
    }
    /* ResultsManager::getResultsSummary (ResultsManager.hpp:191-237) */
    std::stringstream o;
    o << "filter_id\thits\tmisses\tshared\trate_hit\trate_miss\trate_shared\n";
    for (unsigned i = 0; i < nf; ++i) {
        o << filter_ids[i];
        o << "\t" << s.above[i];
        o << "\t" << nrec - s.above[i];
        o << "\t" << (s.above[i] - s.unique[i]);
        o << "\t" << double(s.above[i]) / double(nrec);
        o << "\t" << double(nrec - s.above[i]) / double(nrec);
        o << "\t" << double(s.above[i] - s.unique[i]) / double(nrec);
        o << "\n";
    }
    o << "multiMatch\t" << s.multi << "\t" << nrec - s.multi << "\t" << 0 << "\t"
      << double(s.multi) / double(nrec) << "\t" << double(nrec - s.multi) / double(nrec)
      << "\t" << 0.0 << "\n";
    o << "noMatch\t" << s.none << "\t" << nrec - s.none << "\t" << 0 << "\t"
      << double(s.none) / double(nrec) << "\t" << double(nrec - s.none) / double(nrec)
      << "\t" << 0.0 << "\n";
    std::string str = o.str();
    if (out && cap) {
        uint64_t n = std::min<uint64_t>(cap - 1, str.size());
        std::memcpy(out, str.data(), n);
        out[n] = 0;
    }
    return str.size();
}

} // extern "C"
