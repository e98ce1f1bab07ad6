/*
 * read_pack.h -- packs a batch of reads into the BBT_READS_2BIT layout of
 * include/bbt_hip.h on host threads: 2-bit codes (A 0, C 1, T 2, G 3, i.e.
 * (ASCII >> 1) & 3), four bases per byte, then an N mask of one bit per base
 * (set for any byte that is not A/C/G/T in either case).  3/8 byte per base
 * crosses PCIe instead of 1.
 *
 * Eight bases are packed at a time in a 64-bit word (SWAR): the codes by
 * shifting the per-byte 2-bit fields together, the mask by an exact per-byte
 * zero test against the four letters.  Work is split on 32-base groups (one
 * 8-byte code word, one 4-byte mask word), so threads write disjoint bytes.
 */
#ifndef BBT_READ_PACK_H
#define BBT_READ_PACK_H

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

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

/* Packs the 32-base groups [g0, g1) of the batch whose read r is
 * get(r) -> (pointer, length) at base offset off[r] (off[nreads] bases in
 * all); bases past the end of the batch are packed as 'A'. */
template <class Get>
void pack_2bit_groups(const Get &get, const uint64_t *off, uint64_t nreads, uint64_t g0,
                      uint64_t g1, uint8_t *codes, uint8_t *mask) {
    const uint64_t total = off[nreads];
    if (g0 >= g1 || g0 * 32 >= total + 31)
        return;
    uint64_t r = uint64_t(std::upper_bound(off, off + nreads + 1, g0 * 32) - off);
    r = r ? r - 1 : 0;
    alignas(8) unsigned char buf[32];
    for (uint64_t g = g0; g < g1; ++g) {
        uint64_t p = g * 32;
        const uint64_t end = std::min(p + 32, total);
        uint32_t filled = 0;
        while (p < end) {
            while (r < nreads && off[r + 1] <= p)
                ++r;
            const char *s = get(r);
            const uint64_t take = std::min(end, off[r + 1]) - p;
            std::memcpy(buf + filled, s + (p - off[r]), size_t(take));
            filled += uint32_t(take);
            p += take;
        }
        std::memset(buf + filled, 'A', 32 - filled);
        uint64_t cw = 0;
        uint32_t mw = 0;
        for (uint32_t q = 0; q < 4; ++q) {
            uint64_t x;
            std::memcpy(&x, buf + 8 * q, 8);
            uint32_t c16, m8;
            pack8(x, c16, m8);
            cw |= uint64_t(c16) << (16 * q);
            mw |= m8 << (8 * q);
        }
        std::memcpy(codes + 8 * g, &cw, 8);
        std::memcpy(mask + 4 * g, &mw, 4);
    }
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
    const auto work = [&](unsigned t) {
        pack_2bit_groups(get, off, nreads, groups * t / T, groups * (t + 1) / T, out, out + cb);
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
