/*
 * walker_host.cpp -- compiles the device walk (bbt_device.h) for the CPU so
 * the tests can check its logic against the oracle without a GPU.  Test
 * infrastructure only.
 */
#include <algorithm>
#include <cstring>
#include <vector>

#include "bbt_device.h"

using namespace bbt;

namespace {
/* -D bitmap of the batch (bit off[read] + p), or null */
const uint32_t *g_dust = nullptr;
/* lane walk form: 0 advance()/consume(); n > 0 step()/take() with the bits of
 * a contains() call fetched n at a time, as the persistent kernel does */
unsigned g_flat = 0;
/* log2 of an interleave group (0: the filter's own array): the filter's
 * bytes sit in the last slot of each 2^g_ilv-byte group, the others hold
 * unrelated bytes */
unsigned g_ilv = 0;
std::vector<uint8_t> g_ilv_buf;

void use_interleaved(FilterDev &f) {
    if (!g_ilv)
        return;
    const size_t S = size_t(1) << g_ilv, nb = f.m / 8;
    g_ilv_buf.assign(nb * S, 0);
    for (size_t i = 0; i < nb * S; ++i)
        g_ilv_buf[i] = uint8_t(i * 2654435761u >> 13);
    for (size_t i = 0; i < nb; ++i)
        g_ilv_buf[i * S + S - 1] = f.bits[i];
    f.bits = g_ilv_buf.data() + S - 1;
    f.bshift = g_ilv;
}

template <int SC, bool EXH>
void run(const FilterDev &f, const uint8_t *seq, const uint64_t *off, uint64_t nreads,
         const WalkParams &wp, uint8_t *hits, double *scores, uint64_t *ctr) {
    for (uint64_t r = 0; r < nreads; ++r) {
        Walker<SC, EXH, 0> w;
        w.probes = 0;
        w.lookups = 0;
        w.init(seq + off[r], uint32_t(off[r + 1] - off[r]), &f, wp);
        w.dm = g_dust;
        w.dbase = off[r];
        if (g_flat > 100) {
            /* the lookahead order of the persistent kernel: roll the next
             * base and fetch its window's first bits, then take the pending
             * call's bits, then decide the new window */
            const unsigned na = g_flat - 100;
            const auto bit_of = [&](uint64_t hv, uint32_t jj) {
                const uint64_t pos = fastmod(multi_hash(hv, jj, w.kseed()), f.m, f.magic,
                                             f.sh1, f.sh2);
                return bool((f.bits[(pos >> 3) << f.bshift] >> (pos & 7)) & 1u);
            };
            bool pb[16], cb[16];
            uint32_t pn = 0;
            while (w.state == W_RUNNING) {
                uint32_t cp = 0;
                uint64_t ch = 0;
                const bool cand = w.i < w.len && w.roll_step(cp, ch);
                uint32_t cn = 0;
                if (cand)
                    for (; cn < na && cn < f.h; ++cn)
                        cb[cn] = bit_of(ch, cn);
                if (w.pending) {
                    for (uint32_t jj = 0; jj < pn && w.pending; ++jj)
                        w.take(pb[jj]);
                    while (w.pending) {
                        const uint32_t j0 = w.j;
                        uint32_t n = 0;
                        for (; n < na && j0 + n < f.h; ++n)
                            pb[n] = bit_of(w.h0, j0 + n);
                        for (uint32_t jj = 0; jj < n && w.pending; ++jj)
                            w.take(pb[jj]);
                    }
                }
                pn = 0;
                if (cand && w.state == W_RUNNING && w.decide(cp, ch)) {
                    for (uint32_t jj = 0; jj < cn; ++jj)
                        pb[jj] = cb[jj];
                    pn = cn;
                }
                if (w.state == W_RUNNING && !w.pending && w.i >= w.len)
                    w.finish_end();
            }
        } else if (g_flat) {
            while (w.state == W_RUNNING) {
                if (!w.step())
                    continue;
                for (uint32_t j0 = 0; w.pending; j0 += g_flat) {
                    bool bit[16];
                    for (uint32_t jj = 0; jj < g_flat && j0 + jj < f.h; ++jj) {
                        const uint64_t pos = fastmod(multi_hash(w.h0, j0 + jj, w.kseed()), f.m,
                                                     f.magic, f.sh1, f.sh2);
                        bit[jj] = (f.bits[(pos >> 3) << f.bshift] >> (pos & 7)) & 1u;
                    }
                    for (uint32_t jj = 0; jj < g_flat && j0 + jj < f.h && w.pending; ++jj)
                        w.take(bit[jj]);
                }
            }
        } else {
            while (w.advance())
                w.consume(*w.addr);
        }
        if (EXH)
            scores[r] = w.score_value();
        else
            hits[r] = w.accepted() ? 1 : 0;
        ctr[0] += w.probes;
        ctr[1] += w.lookups;
    }
}
/* The wave-cooperative walk of the long-read kernel, with the 64 lanes
 * emulated by loops: prefix-XOR window hashes per chunk of 64 bases,
 * speculative probing of every window the walk may probe, then the walk
 * replayed over the chunk's valid / hit masks. */
template <int SC, bool EXH>
void run_wave(const FilterDev &f, const uint8_t *seq, const uint64_t *off, uint64_t nreads,
              const WalkParams &wp, uint8_t *hits, double *scores, uint64_t *ctr) {
    for (uint64_t r = 0; r < nreads; ++r) {
        Walker<SC, EXH, 0> w;
        const uint8_t *s = seq + off[r];
        const uint32_t len = uint32_t(off[r + 1] - off[r]);
        w.init(s, len, &f, wp);
        w.dm = g_dust;
        w.dbase = off[r];
        w.probes = 0;
        w.lookups = 0;
        const unsigned kv = f.k;
        uint64_t Hc = 0, Gc = 0, Hp[64] = {0}, Gp[64] = {0};
        int64_t lastN = -1;
        for (uint32_t b0 = 0; b0 < len && w.state == W_RUNNING; b0 += 64) {
            uint64_t H[64], G[64], h0[64];
            int64_t ln[64];
            bool valid[64];
            uint32_t lk[64];
            bool allset[64];
            bool low[64];
            for (int L = 0; L < 64; ++L) {
                const uint32_t i = b0 + L;
                const unsigned c = i < len ? s[i] : 0;
                const bool ok = i < len && base_valid(c);
                H[L] = (ok ? pfx_fwd(base_code(c), i) : 0) ^ (L ? H[L - 1] : Hc);
                G[L] = (ok ? pfx_rc(base_code(c), i) : 0) ^ (L ? G[L - 1] : Gc);
                const int64_t me = ok ? -1 : int64_t(i);
                ln[L] = std::max(me, L ? ln[L - 1] : lastN);
            }
            for (int L = 0; L < 64; ++L) {
                const uint32_t i = b0 + L;
                const int src = L - int(kv);
                const uint64_t Hb = src >= 0 ? H[src] : Hp[src + 64];
                const uint64_t Gb = src >= 0 ? G[src] : Gp[src + 64];
                const int64_t pw = int64_t(i) - kv + 1;
                valid[L] = pw >= 0 && i < len && ln[L] < pw;
                const uint64_t wf = pfx_win_fwd(H[L], Hb, i);
                const uint64_t wr = pfx_win_rc(G[L], Gb, uint32_t(pw));
                h0[L] = wr < wf ? wr : wf;
                low[L] = valid[L] && w.low_complexity(uint32_t(pw));
                bool need = valid[L] && !low[L] && !(w.skipping && uint32_t(pw) < w.skipEnd);
                allset[L] = need;
                lk[L] = 0;
                for (uint32_t j = 0; need; ++j) {
                    uint64_t hv = h0[L];
                    if (j) {
                        hv = h0[L] * f.mult[j];
                        hv ^= hv >> kMultiShift;
                    }
                    const uint64_t pos = fastmod(hv, f.m, f.magic, f.sh1, f.sh2);
                    ++lk[L];
                    if (!(f.bits[(pos >> 3) << f.bshift] & (1u << (pos & 7)))) {
                        allset[L] = false;
                        need = false;
                    } else if (j + 1 >= f.h) {
                        need = false;
                    }
                }
                ctr[2] += lk[L];
            }
            Hc = H[63];
            Gc = G[63];
            lastN = ln[63];
            for (int L = 0; L < 64; ++L) {
                Hp[L] = H[L];
                Gp[L] = G[L];
            }
            uint64_t vmask = 0, hmask = 0, dmask = 0;
            uint32_t lkp[65];
            lkp[0] = 0;
            for (int L = 0; L < 64; ++L) {
                vmask |= uint64_t(valid[L]) << L;
                hmask |= uint64_t(allset[L]) << L;
                dmask |= uint64_t(low[L]) << L;
                lkp[L + 1] = lkp[L] + lk[L];
            }
            w.replay(b0, vmask, hmask, dmask,
                     [&](uint32_t x, uint32_t y) -> uint32_t { return lkp[y] - lkp[x]; });
        }
        if (w.state == W_RUNNING)
            w.finish_end();
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

int walker_host_wave_eval(const uint8_t *bits, uint64_t m, uint32_t h, uint32_t k,
                          const uint32_t *binom, uint32_t binom_len, const uint8_t *seq,
                          const uint64_t *off, uint64_t nreads, uint32_t scorer, int exh,
                          double threshold, uint32_t streak, uint8_t *hits, double *scores,
                          uint64_t *ctr, const uint32_t *dust) {
    g_dust = dust;
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
    roll_fill(f.roll, k);
    use_interleaved(f);
    WalkParams wp;
    wp.threshold = threshold;
    wp.streak = streak;
    wp.min_len = uint32_t(threshold + 0.5);
#define BBT_WAVE_CASE(S)                                                              \
    case S:                                                                           \
        exh ? run_wave<S, true>(f, seq, off, nreads, wp, hits, scores, ctr)           \
            : run_wave<S, false>(f, seq, off, nreads, wp, hits, scores, ctr);         \
        break;
    switch (scorer) {
        BBT_WAVE_CASE(SIMPLE)
        BBT_WAVE_CASE(HARMONIC)
        BBT_WAVE_CASE(LENGTH)
        BBT_WAVE_CASE(BINOMIAL)
    default:
        return 1;
    }
#undef BBT_WAVE_CASE
    return 0;
}


/* One filter, all reads; binom[] (frame length -> threshold) optional. */
int walker_host_eval(const uint8_t *bits, uint64_t m, uint32_t h, uint32_t k,
                     const uint32_t *binom, uint32_t binom_len, const uint8_t *seq,
                     const uint64_t *off, uint64_t nreads, uint32_t scorer, int exh,
                     double threshold, uint32_t streak, uint8_t *hits, double *scores,
                     uint64_t *ctr, const uint32_t *dust) {
    g_dust = dust;
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
    roll_fill(f.roll, k);
    use_interleaved(f);
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

/* RowsWalk (one read against every filter of an interleaved group, SIMPLE
 * evalRead) over G filters of one m, h and k: bits[g] their arrays; hits
 * read-major (r * G + g); ctr[0] contains() calls, ctr[1] bit lookups. */
int walker_host_rows(const uint8_t *const *bits, uint32_t G, uint64_t m, uint32_t h, uint32_t k,
                     const uint8_t *seq, const uint64_t *off, uint64_t nreads, double threshold,
                     uint32_t streak, uint8_t *hits, uint64_t *ctr, const uint32_t *dust) {
    if (G < 3 || G > 16)
        return 1;
    const unsigned SH = G <= 4 ? 2 : (G <= 8 ? 3 : 4);
    const size_t S = size_t(1) << SH, nb = m / 8;
    std::vector<uint8_t> ilv(nb * S + 16, 0);
    for (size_t i = 0; i < nb; ++i)
        for (uint32_t g = 0; g < G; ++g)
            ilv[i * S + g] = bits[g][i];
    FilterDev f;
    std::memset(&f, 0, sizeof(f));
    f.bits = bits[0];
    f.m = m;
    FastMod fm = fastmod_make(m);
    f.magic = fm.magic;
    f.sh1 = fm.sh1;
    f.sh2 = fm.sh2;
    f.h = h;
    f.k = k;
    roll_fill(f.roll, k);
    WalkParams wp;
    wp.threshold = threshold;
    wp.streak = streak;
    wp.min_len = uint32_t(threshold + 0.5);
    const uint64_t ks = uint64_t(k) * kMultiSeed;
    const auto fetch = [&](uint64_t hv, uint32_t j) {
        const uint64_t pos = fastmod(multi_hash(hv, j, ks), m, fm.magic, fm.sh1, fm.sh2);
        uint32_t w[4];
        std::memcpy(w, ilv.data() + ((pos >> 3) << SH), 4 * ((S + 3) / 4));
        uint32_t b = 0;
        if (SH == 2)
            b = group_bits<2>(w, unsigned(pos & 7));
        else if (SH == 3)
            b = group_bits<3>(w, unsigned(pos & 7));
        else
            b = group_bits<4>(w, unsigned(pos & 7));
        return b & ((1u << G) - 1u);
    };
    for (uint64_t r = 0; r < nreads; ++r) {
        const uint32_t len = uint32_t(off[r + 1] - off[r]);
        if (len >= 32768)
            return 2;
        RowsWalk<0, 16> rw;
        rw.init(seq + off[r], len, &f, G, wp);
        rw.w.dm = dust;
        rw.w.dbase = off[r];
        while (rw.step(fetch)) {
        }
        for (uint32_t g = 0; g < G; ++g)
            hits[r * G + g] = (rw.acc >> g) & 1u;
        ctr[0] += rw.probes;
        ctr[1] += rw.lookups;
    }
    return 0;
}

void walker_host_set_ilv(unsigned sh) { g_ilv = sh < 7 ? sh : 6; }

void walker_host_set_flat(unsigned n) { g_flat = n > 100 ? (n < 116 ? n : 116) : (n < 16 ? n : 16); }

/* base_valid and base_code for every byte value */
void walker_host_bases(uint8_t *valid, uint8_t *code) {
    for (unsigned c = 0; c < 256; ++c) {
        valid[c] = base_valid(c) ? 1 : 0;
        code[c] = uint8_t(base_code(c));
    }
}

/* the roll table of k */
void walker_host_roll(unsigned k, uint64_t *t) { roll_fill(t, k); }
}
