/*
 * read_pack.h -- packs a batch of reads into the BBT_READS_2BIT layout of
 * include/bbt_hip.h on host threads: 2-bit codes (A 0, C 1, T 2, G 3, i.e.
 * (ASCII >> 1) & 3), four bases per byte, then an N mask of one bit per base
 * (set for any byte that is not A/C/G/T in either case).  3/8 byte per base
 * crosses PCIe instead of 1.
 *
 * Reads are packed 32 bases at a time, with AVX2 where the host has it, else
 * eight at a time in a 64-bit word (SWAR: the codes by shifting the per-byte
 * 2-bit fields together, the mask by an exact per-byte zero test against the
 * four letters), and appended to the output at any base offset.  Work is
 * split on 32-base groups (one 8-byte code word, one 4-byte mask word), so
 * threads write disjoint bytes.
 */
#ifndef BBT_READ_PACK_H
#define BBT_READ_PACK_H

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#if defined(__x86_64__)
#include <immintrin.h>
#endif

namespace bbt_host {

/* bytes of codes for n bases: whole 32-base groups (a multiple of 8); the
 * mask of half that size follows */
inline uint64_t packed_code_bytes(uint64_t nbases) { return (nbases + 31) / 32 * 8; }
inline uint64_t packed_bytes(uint64_t nbases) { return packed_code_bytes(nbases) * 3 / 2; }

/* 8 ASCII bytes (byte i = base i) -> 16 bits of codes (base i at bits 2i)
 * and 8 mask bits (bit i = base i is not A/C/G/T) */
inline void pack8(uint64_t x, uint32_t &codes, uint32_t &mask) {
    uint64_t c = (x >> 1) & 0x0303030303030303ull;
    c |= c >> 6;
    c &= 0x000F000F000F000Full;
    c |= c >> 12;
    c &= 0x000000FF000000FFull;
    c |= c >> 24;
    codes = uint32_t(c & 0xFFFF);
    const uint64_t ones = 0x0101010101010101ull, lo7 = 0x7F7F7F7F7F7F7F7Full;
    const uint64_t u = x & 0xDFDFDFDFDFDFDFDFull; /* upper case */
    const auto zero = [&](uint64_t v) { return ~(((v & lo7) + lo7) | v | lo7); };
    const uint64_t valid = zero(u ^ (ones * 'A')) | zero(u ^ (ones * 'C')) |
                           zero(u ^ (ones * 'G')) | zero(u ^ (ones * 'T'));
    const uint64_t inv = (~valid >> 7) & ones;
    mask = uint32_t((inv * 0x0102040810204080ull) >> 56);
}

/* 32 ASCII bytes -> 64 bits of codes (base i at bits 2i) and 32 mask bits,
 * four pack8 calls */
inline void pack32_swar(const unsigned char *p, uint64_t &codes, uint32_t &mask) {
    codes = 0;
    mask = 0;
    for (uint32_t q = 0; q < 4; ++q) {
        uint64_t x;
        std::memcpy(&x, p + 8 * q, 8);
        uint32_t c16, m8;
        pack8(x, c16, m8);
        codes |= uint64_t(c16) << (16 * q);
        mask |= m8 << (8 * q);
    }
}

#if defined(__x86_64__)
/* The same with AVX2: a case fold and four byte compares give the mask
 * (movemask of the A/C/G/T matches); (byte >> 1) & 3 are the codes, four to a
 * byte by two multiply-add steps (weights 1, 4 then 1, 16), whose low bytes a
 * shuffle collects. */
__attribute__((target("avx2"))) inline void pack32_avx2(const unsigned char *p,
                                                         uint64_t &codes, uint32_t &mask) {
    const __m256i x = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(p));
    const __m256i u = _mm256_and_si256(x, _mm256_set1_epi8(char(0xDF)));
    const __m256i ac = _mm256_or_si256(_mm256_cmpeq_epi8(u, _mm256_set1_epi8('A')),
                                       _mm256_cmpeq_epi8(u, _mm256_set1_epi8('C')));
    const __m256i gt = _mm256_or_si256(_mm256_cmpeq_epi8(u, _mm256_set1_epi8('G')),
                                       _mm256_cmpeq_epi8(u, _mm256_set1_epi8('T')));
    mask = ~uint32_t(_mm256_movemask_epi8(_mm256_or_si256(ac, gt)));
    const __m256i k = _mm256_and_si256(_mm256_srli_epi16(x, 1), _mm256_set1_epi8(3));
    const __m256i p16 = _mm256_maddubs_epi16(k, _mm256_set1_epi16(0x0401));
    const __m256i p32 = _mm256_madd_epi16(p16, _mm256_set1_epi32(0x00100001));
    const __m256i sel = _mm256_setr_epi8(0, 4, 8, 12, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                                         -1, -1, 0, 4, 8, 12, -1, -1, -1, -1, -1, -1, -1, -1,
                                         -1, -1, -1, -1);
    const __m256i b = _mm256_shuffle_epi8(p32, sel);
    codes = uint64_t(uint32_t(_mm256_extract_epi32(b, 0))) |
            (uint64_t(uint32_t(_mm256_extract_epi32(b, 4))) << 32);
}

inline bool pack_avx2_available() { return __builtin_cpu_supports("avx2"); }
#else
inline void pack32_avx2(const unsigned char *p, uint64_t &codes, uint32_t &mask) {
    pack32_swar(p, codes, mask);
}
inline bool pack_avx2_available() { return false; }
#endif

/* -1: AVX2 where the host has it; 0: the SWAR form; 1: AVX2 (tests) */
inline int &pack_simd_mode() {
    static int mode = -1;
    return mode;
}

/* Output words of 32 bases: each put() appends n <= 32 packed bases (codes in
 * the low 2n bits, mask in the low n bits) and stores a code word and a mask
 * word whenever 32 bases are complete. */
struct PackSink {
    uint8_t *codes;
    uint8_t *mask;
    uint64_t g;        /* next group */
    uint64_t cacc = 0; /* codes of the group being filled */
    uint64_t macc = 0; /* its mask bits */
    unsigned fill = 0; /* bases in it */

    void put(uint64_t c, uint64_t m, unsigned n) {
        cacc |= c << (2 * fill);
        macc |= m << fill;
        fill += n;
        if (fill >= 32) {
            store();
            fill -= 32;
            const unsigned used = n - fill; /* of this call's bases, stored */
            cacc = fill ? c >> (2 * used) : 0;
            macc = fill ? m >> used : 0;
        }
    }
    void store() {
        const uint32_t mw = uint32_t(macc);
        std::memcpy(codes + 8 * g, &cacc, 8);
        std::memcpy(mask + 4 * g, &mw, 4);
        ++g;
    }
    /* the last, partial group: its missing bases as 'A' (code 0, unmasked) */
    void finish() {
        if (fill) {
            store();
            fill = 0;
            cacc = macc = 0;
        }
    }
};

/* Packs the 32-base groups [g0, g1) of the batch whose read r is
 * get(r) -> pointer at base offset off[r] (off[nreads] bases in all); bases
 * past the end of the batch are packed as 'A'.  Reads are taken 32 bases at a
 * time; a read's last, shorter piece is packed from the 32 bytes ending at its
 * end and shifted, so no byte outside a read is loaded and no copy is needed
 * (only a read shorter than 32 bases goes through a padded copy). */
template <bool AVX2, class Get>
void pack_2bit_groups(const Get &get, const uint64_t *off, uint64_t nreads, uint64_t g0,
                      uint64_t g1, uint8_t *codes, uint8_t *mask) {
    const uint64_t total = off[nreads];
    const uint64_t lo = g0 * 32, hi = std::min(g1 * 32, total);
    if (g0 >= g1 || lo >= hi)
        return;
    const auto pack32 = [](const unsigned char *p, uint64_t &c, uint32_t &m) {
        if (AVX2)
            pack32_avx2(p, c, m);
        else
            pack32_swar(p, c, m);
    };
    PackSink out;
    out.codes = codes;
    out.mask = mask;
    out.g = g0;
    uint64_t r = uint64_t(std::upper_bound(off, off + nreads + 1, lo) - off);
    r = r ? r - 1 : 0;
    for (; r < nreads && off[r] < hi; ++r) {
        const uint64_t a = std::max(off[r], lo), b = std::min(off[r + 1], hi);
        if (a >= b)
            continue;
        const unsigned char *s = reinterpret_cast<const unsigned char *>(get(r));
        uint64_t sa = a - off[r];
        const uint64_t sb = b - off[r];
        uint64_t c;
        uint32_t m;
        for (; sb - sa >= 32; sa += 32) {
            pack32(s + sa, c, m);
            out.put(c, m, 32);
        }
        if (sa < sb) {
            const unsigned n = unsigned(sb - sa);
            if (sb >= 32) {
                pack32(s + sb - 32, c, m);
                out.put(c >> (2 * (32 - n)), uint64_t(m) >> (32 - n), n);
            } else {
                alignas(32) unsigned char buf[32];
                std::memset(buf, 'A', 32);
                for (unsigned q = 0; q < n; ++q)
                    buf[q] = s[sa + q];
                pack32(buf, c, m);
                out.put(c, m, n);
            }
        }
    }
    out.finish();
}

/* The whole batch into out (packed_bytes(off[nreads]) bytes) on up to
 * `threads` threads; returns the code bytes (bbt_reads.seq_bytes). */
template <class Get>
uint64_t pack_2bit(const Get &get, const uint64_t *off, uint64_t nreads, uint8_t *out,
                   unsigned threads) {
    const uint64_t total = off[nreads];
    const uint64_t cb = packed_code_bytes(total);
    const uint64_t groups = cb / 8;
    const unsigned T = unsigned(std::max<uint64_t>(1, std::min<uint64_t>(threads, groups / 4096)));
    const int mode = pack_simd_mode();
    const bool avx2 = mode == 1 || (mode < 0 && pack_avx2_available());
    const auto work = [&](unsigned t) {
        if (avx2)
            pack_2bit_groups<true>(get, off, nreads, groups * t / T, groups * (t + 1) / T, out,
                                   out + cb);
        else
            pack_2bit_groups<false>(get, off, nreads, groups * t / T, groups * (t + 1) / T, out,
                                    out + cb);
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; ++t)
        th.emplace_back(work, t);
    work(0);
    for (std::thread &x : th)
        x.join();
    return cb;
}

} // namespace bbt_host

#endif
