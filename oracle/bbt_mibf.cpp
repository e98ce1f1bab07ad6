/*
 * bbt_mibf.cpp -- CPU restatement of the multi-index Bloom filter (MIBF)
 * build and query.  TEST INFRASTRUCTURE ONLY: the checker of the MIBF parity
 * tests; no product code path uses it.
 *
 * The reference's MIBloomFilter / MIBFConstructSupport / MIBFQuerySupport
 * live in the absent btl_bloomfilter submodule; what the mount shows is the
 * call structure (BioBloomMaker/MIBFGen.hpp:131-309: bit-vector pass,
 * ID pass with random ownership, saturation pass "if target frame does not
 * have a single representative, mark frame as saturated"; BioBloomCategorizer/
 * MIBFClassifier.hpp:79-94, 773-796, 930-939: allowed misses, per-frame
 * probabilities, binomial minimum counts, query, best hit first).  The
 * definitions below are this build's restatement of that structure and are
 * the specification the device path is checked against ("parity unpinned"
 * against upstream):
 *
 *  hashing     spaced-seed mode: one stHash value per seed per N-free window;
 *              k-mer mode (no seeds): the S ntHash values of each k-mer
 *              (sntHashIterator is ntHashIterator, Common/sntHashIterator.hpp).
 *  bit vector  m bits, m = calcOptimalSize(entries, #seeds, occupancy) with
 *              the occupancy form m = -n*S/ln(1-occ) rounded up as
 *              BloomFilterInfo::calcOptimalSize rounds; bit p of seed hash h
 *              is p = h mod m.
 *  IDs         data[rank1(p)] for every set bit p, 32-bit values: low 31 bits
 *              the ID (1-based, in record order), bit 31 saturation.
 *  ownership   a slot claimed by several IDs keeps the ID with the largest
 *              (mix32(h ^ id * 0x9E3779B97F4A7C15), id): random-looking but
 *              independent of insertion order, so a build is reproducible.
 *  saturation  for every frame of every record: when no seed slot of the
 *              frame holds the record's ID, all those slots get bit 31.
 *  query       per N-free frame: S seed lookups; the frame is evaluated when
 *              at least S - allowedMiss bits are set; each distinct ID among
 *              the hit slots counts once per frame (count), once more if one
 *              of its slots is unsaturated (nonSat), by slots (total,
 *              totalNonSat), as solid when all S slots hold it, and as a
 *              non-saturated frame when no hit slot of the frame is saturated
 *              (nonSatFrame).  A frame whose hit slots are all saturated adds
 *              to satCount.  A pair pools the frames of both mates; its
 *              frameCount follows the nesting of MIBFClassifier.hpp:804-805
 *              (mate 1's frames, or mate 2's when mate 1 is shorter than k).
 *  candidates  IDs 1..#IDs only (a slot holding ID 0, which no build writes,
 *              counts toward its frame but is never a candidate);
 *              count >= minCount[id] (binomial quantile at score/(#IDs) on
 *              the ID's per-frame probability, at least 1, frameCount =
 *              len - span + 1) and nonSat >= minNoSat; ranked by nonSat,
 *              count, solid, totalNonSat, total (descending) then ID; the best
 *              and every candidate with nonSat >= multiThresh * best.nonSat
 *              are reported in rank order.  bestHitAgree (-b): no result
 *              unless the best candidate leads (ties allowed) in every
 *              statistic.  The streak option (-r) stops no query early here:
 *              every frame is evaluated.
 *  frame prob  P(id) = sum_{j=S-a..S} C(S,j) occ^j (1-occ)^(S-j) (1-(1-f_id)^j)
 *              with occ = popcount/m, f_id = share of ID slots holding id.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "bbt_oracle.h"

namespace {

const uint32_t kSat = 0x80000000u;
const uint32_t kIdMask = 0x7FFFFFFFu;

/* hashes of every N-free window of s: window starts, hv[i*S+j].  Spaced-seed
 * mode (stHashIterator, one hash per seed) when seeds is set; k-mer mode
 * (ntHashIterator with S hashes per k-mer, Common/sntHashIterator.hpp:18-35)
 * when seeds is NULL. */
void seed_hashes(const char *s, uint64_t len, const char *seeds, uint32_t S, uint32_t k,
                 std::vector<uint64_t> &pos, std::vector<uint64_t> &hv) {
    pos.clear();
    hv.clear();
    if (len < k)
        return;
    int64_t n = int64_t(len);
    std::vector<uint64_t> p(static_cast<size_t>(n)), h(static_cast<size_t>(n) * S);
    int64_t cnt = seeds ? orc_sthash(s, len, seeds, S, k, p.data(), h.data(), uint64_t(n))
                        : orc_nthash(s, len, k, S, p.data(), h.data(), uint64_t(n));
    pos.assign(p.begin(), p.begin() + cnt);
    hv.assign(h.begin(), h.begin() + cnt * S);
}

inline uint32_t mix32(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return uint32_t(x >> 32);
}

struct Rank {
    std::vector<uint64_t> before; /* set bits before word w */
    explicit Rank(const uint64_t *bv, uint64_t m) : before(m / 64 + 1) {
        uint64_t c = 0;
        for (uint64_t w = 0; w < m / 64; ++w) {
            before[w] = c;
            c += uint64_t(__builtin_popcountll(bv[w]));
        }
        before[m / 64] = c;
    }
    uint64_t rank1(const uint64_t *bv, uint64_t p) const {
        uint64_t w = p >> 6, b = p & 63;
        uint64_t mask = b ? (~uint64_t(0) >> (64 - b)) : 0;
        return before[w] + uint64_t(__builtin_popcountll(bv[w] & mask));
    }
};

inline bool bit(const uint64_t *bv, uint64_t p) { return (bv[p >> 6] >> (p & 63)) & 1; }

} // namespace

extern "C" {

uint64_t orc_mibf_calc_size(uint64_t entries, uint32_t nseeds, double occupancy) {
    uint64_t v = uint64_t(-double(entries) * double(nseeds) / std::log(1.0 - occupancy));
    return v + (64 - v % 64);
}

void orc_mibf_set_bits(const char *seq, uint64_t len, const char *seeds, uint32_t S,
                       uint32_t k, uint64_t m, uint64_t *bv) {
    std::vector<uint64_t> pos, hv;
    seed_hashes(seq, len, seeds, S, k, pos, hv);
    for (uint64_t h : hv) {
        uint64_t p = h % m;
        bv[p >> 6] |= uint64_t(1) << (p & 63);
    }
}

uint64_t orc_mibf_popcount(const uint64_t *bv, uint64_t m) {
    uint64_t c = 0;
    for (uint64_t w = 0; w < m / 64; ++w)
        c += uint64_t(__builtin_popcountll(bv[w]));
    return c;
}

void orc_mibf_assign(const char *seq, const uint64_t *rec_off, const uint32_t *rec_id,
                     uint64_t nrec, const char *seeds, uint32_t S, uint32_t k, uint64_t m,
                     const uint64_t *bv, uint32_t *data) {
    Rank R(bv, m);
    const uint64_t pop = R.before[m / 64];
    std::vector<uint64_t> best(pop, 0);
    std::vector<uint64_t> pos, hv;
    for (uint64_t r = 0; r < nrec; ++r) {
        seed_hashes(seq + rec_off[r], rec_off[r + 1] - rec_off[r], seeds, S, k, pos, hv);
        const uint64_t id = rec_id[r];
        for (uint64_t h : hv) {
            uint64_t slot = R.rank1(bv, h % m);
            uint64_t key = (uint64_t(mix32(h ^ (id * 0x9E3779B97F4A7C15ULL))) << 32) | id;
            if (key > best[slot])
                best[slot] = key;
        }
    }
    for (uint64_t i = 0; i < pop; ++i)
        data[i] = uint32_t(best[i]) & kIdMask;
    /* saturation (MIBFGen.hpp:235-262) */
    for (uint64_t r = 0; r < nrec; ++r) {
        seed_hashes(seq + rec_off[r], rec_off[r + 1] - rec_off[r], seeds, S, k, pos, hv);
        const uint32_t id = rec_id[r];
        for (size_t f = 0; f < pos.size(); ++f) {
            bool rep = false;
            for (uint32_t j = 0; j < S; ++j)
                if ((data[R.rank1(bv, hv[f * S + j] % m)] & kIdMask) == id)
                    rep = true;
            if (!rep)
                for (uint32_t j = 0; j < S; ++j)
                    data[R.rank1(bv, hv[f * S + j] % m)] |= kSat;
        }
    }
}

void orc_mibf_frame_probs(const uint32_t *data, uint64_t pop, uint64_t m, uint32_t S,
                          uint32_t allowed_miss, uint32_t nids, double *probs) {
    std::vector<uint64_t> n(size_t(nids) + 1, 0);
    for (uint64_t i = 0; i < pop; ++i) {
        uint32_t id = data[i] & kIdMask;
        if (id <= nids)
            ++n[id];
    }
    const double occ = double(pop) / double(m);
    probs[0] = 0;
    for (uint32_t id = 1; id <= nids; ++id) {
        const double f = pop ? double(n[id]) / double(pop) : 0.0;
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
}

int orc_mibf_query(const uint64_t *bv, const uint32_t *data, uint64_t m, const char *seeds,
                   uint32_t S, uint32_t k, const double *probs, uint32_t nids, double score,
                   uint32_t allowed_miss, double multi_thresh, uint32_t min_nonsat,
                   int paired, int best_agree, const char *seq, const uint64_t *off,
                   uint64_t nreads, uint32_t max_hits, orc_mibf_hit *hits,
                   uint32_t *nhits, uint32_t *satcount, uint64_t *counters) {
    Rank R(bv, m);
    const uint64_t nrec = paired ? nreads / 2 : nreads;
    /* frameCount of each record; MIBFClassifier.hpp:804-805 nests the
     * conditionals, so a pair counts mate 2 only when mate 1 is shorter than k */
    std::vector<uint64_t> fcs(nrec);
    for (uint64_t r = 0; r < nrec; ++r) {
        const uint64_t a0 = paired ? 2 * r : r;
        const uint64_t len1 = off[a0 + 1] - off[a0];
        uint64_t fc = len1 >= k ? len1 - k + 1 : 0;
        if (paired && len1 < k) {
            const uint64_t len2 = off[a0 + 2] - off[a0 + 1];
            fc = len2 >= k ? len2 - k + 1 : 0;
        }
        fcs[r] = fc;
    }
    /* minCount tables (getMinCount, MIBFClassifier.hpp:930-939) per frameCount */
    std::map<uint64_t, std::vector<uint32_t>> minCount;
    for (uint64_t fc : fcs) {
        if (fc == 0 || minCount.count(fc))
            continue;
        std::vector<uint32_t> t(size_t(nids) + 1, 1);
        for (uint32_t id = 1; id <= nids; ++id) {
            uint64_t q = orc_binom_isf_up(score / double(nids), fc, probs[id]);
            t[id] = q > 1 ? uint32_t(q) : 1;
        }
        minCount.emplace(fc, std::move(t));
    }
    uint64_t bvl = 0, gathers = 0, overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bvl, gathers, overflow)
    for (uint64_t r = 0; r < nrec; ++r) {
        nhits[r] = 0;
        satcount[r] = 0;
        const uint64_t frameCount = fcs[r];
        if (frameCount == 0)
            continue;
        const std::vector<uint32_t> &mc = minCount.find(frameCount)->second;
        const uint64_t a0 = paired ? 2 * r : r;
        std::vector<uint64_t> pos, hv;
        std::map<uint32_t, orc_mibf_hit> acc;
        for (uint64_t mate = a0; mate < a0 + (paired ? 2 : 1); ++mate) {
            const char *s = seq + off[mate];
            seed_hashes(s, off[mate + 1] - off[mate], seeds, S, k, pos, hv);
            for (size_t f = 0; f < pos.size(); ++f) {
                uint32_t raw[32];
                uint32_t nh = 0;
                for (uint32_t j = 0; j < S; ++j) {
                    const uint64_t p = hv[f * S + j] % m;
                    ++bvl;
                    if (bit(bv, p)) {
                        raw[nh++] = data[R.rank1(bv, p)];
                        ++gathers;
                    }
                }
                if (nh + allowed_miss < S)
                    continue;
                bool allsat = nh > 0, anysat = false;
                for (uint32_t a = 0; a < nh; ++a) {
                    if (!(raw[a] & kSat))
                        allsat = false;
                    else
                        anysat = true;
                }
                if (allsat)
                    ++satcount[r];
                for (uint32_t a = 0; a < nh; ++a) {
                    const uint32_t id = raw[a] & kIdMask;
                    bool seen = false;
                    for (uint32_t b = 0; b < a; ++b)
                        if ((raw[b] & kIdMask) == id)
                            seen = true;
                    if (seen)
                        continue;
                    uint32_t occ = 0, ns = 0;
                    for (uint32_t b = a; b < nh; ++b)
                        if ((raw[b] & kIdMask) == id) {
                            ++occ;
                            if (!(raw[b] & kSat))
                                ++ns;
                        }
                    orc_mibf_hit &h = acc[id];
                    h.id = id;
                    h.count += 1;
                    h.nonsat += ns ? 1 : 0;
                    h.total += occ;
                    h.total_nonsat += ns;
                    h.solid += occ == S ? 1 : 0;
                    h.nonsat_frame += anysat ? 0 : 1;
                }
            }
        }
        std::vector<orc_mibf_hit> cand;
        for (auto &kv : acc) {
            const orc_mibf_hit &h = kv.second;
            if (h.id >= 1 && h.id <= nids && h.count >= mc[h.id] && h.nonsat >= min_nonsat)
                cand.push_back(h);
        }
        if (cand.empty())
            continue;
        std::sort(cand.begin(), cand.end(), [](const orc_mibf_hit &a, const orc_mibf_hit &b) {
            if (a.nonsat != b.nonsat) return a.nonsat > b.nonsat;
            if (a.count != b.count) return a.count > b.count;
            if (a.solid != b.solid) return a.solid > b.solid;
            if (a.total_nonsat != b.total_nonsat) return a.total_nonsat > b.total_nonsat;
            if (a.total != b.total) return a.total > b.total;
            return a.id < b.id;
        });
        if (best_agree) {
            /* best hit must lead (ties allowed) in every statistic */
            const orc_mibf_hit &t = cand[0];
            bool agree = true;
            for (const orc_mibf_hit &h : cand)
                if (h.count > t.count || h.nonsat > t.nonsat || h.total > t.total ||
                    h.total_nonsat > t.total_nonsat || h.solid > t.solid ||
                    h.nonsat_frame > t.nonsat_frame)
                    agree = false;
            if (!agree)
                continue;
        }
        const double bar = multi_thresh * double(cand[0].nonsat);
        uint32_t n = 0;
        for (const orc_mibf_hit &h : cand) {
            if (double(h.nonsat) < bar)
                break;
            if (n < max_hits)
                hits[r * max_hits + n] = h;
            else
                ++overflow;
            ++n;
        }
        nhits[r] = n;
    }
    if (counters) {
        counters[0] += bvl;
        counters[1] += gathers;
        counters[2] += overflow;
    }
    return 0;
}

} // extern "C"
