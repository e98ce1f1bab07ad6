/*
 * walker_host.cpp -- compiles the device walk (bbt_device.h) for the CPU so
 * the tests can check its logic against the oracle without a GPU.  Test
 * infrastructure only.
 */
#include <cstring>

#include "bbt_device.h"

using namespace bbt;

namespace {
template <int SC, bool EXH>
void run(const FilterDev &f, const uint8_t *seq, const uint64_t *off, uint64_t nreads,
         const WalkParams &wp, uint8_t *hits, double *scores, uint64_t *ctr) {
    for (uint64_t r = 0; r < nreads; ++r) {
        Walker<SC, EXH, 0> w;
        w.probes = 0;
        w.lookups = 0;
        w.init(seq + off[r], uint32_t(off[r + 1] - off[r]), &f, wp);
        while (w.advance())
            w.consume(*w.addr);
        if (EXH)
            scores[r] = w.score_value();
        else
            hits[r] = w.accepted() ? 1 : 0;
        ctr[0] += w.probes;
        ctr[1] += w.lookups;
    }
}
} // namespace

extern "C" {

/* One filter, all reads; binom[] (frame length -> threshold) optional. */
int walker_host_eval(const uint8_t *bits, uint64_t m, uint32_t h, uint32_t k,
                     const uint32_t *binom, uint32_t binom_len, const uint8_t *seq,
                     const uint64_t *off, uint64_t nreads, uint32_t scorer, int exh,
                     double threshold, uint32_t streak, uint8_t *hits, double *scores,
                     uint64_t *ctr) {
    FilterDev f;
    std::memset(&f, 0, sizeof(f));
    f.bits = bits;
    f.m = m;
    FastMod fm = fastmod_make(m);
    f.magic = fm.magic;
    f.sh1 = fm.sh1;
    f.sh2 = fm.sh2;
    f.h = h;
    f.k = k;
    f.binom = binom;
    f.binom_len = binom_len;
    for (uint32_t j = 1; j < kMaxHash; ++j)
        f.mult[j] = uint64_t(j) ^ (uint64_t(k) * kMultiSeed);
    WalkParams wp;
    wp.threshold = threshold;
    wp.streak = streak;
    wp.min_len = uint32_t(threshold + 0.5);
    switch (scorer) {
    case SIMPLE:
        exh ? run<SIMPLE, true>(f, seq, off, nreads, wp, hits, scores, ctr)
            : run<SIMPLE, false>(f, seq, off, nreads, wp, hits, scores, ctr);
        break;
    case HARMONIC:
        exh ? run<HARMONIC, true>(f, seq, off, nreads, wp, hits, scores, ctr)
            : run<HARMONIC, false>(f, seq, off, nreads, wp, hits, scores, ctr);
        break;
    case LENGTH:
        exh ? run<LENGTH, true>(f, seq, off, nreads, wp, hits, scores, ctr)
            : run<LENGTH, false>(f, seq, off, nreads, wp, hits, scores, ctr);
        break;
    case BINOMIAL:
        exh ? run<BINOMIAL, true>(f, seq, off, nreads, wp, hits, scores, ctr)
            : run<BINOMIAL, false>(f, seq, off, nreads, wp, hits, scores, ctr);
        break;
    default:
        return 1;
    }
    return 0;
}

/* n[i] % d through the device FastMod */
void walker_host_fastmod(uint64_t d, const uint64_t *n, uint64_t count, uint64_t *out) {
    FastMod fm = fastmod_make(d);
    for (uint64_t i = 0; i < count; ++i)
        out[i] = fastmod(n[i], fm.d, fm.magic, fm.sh1, fm.sh2);
}
}
