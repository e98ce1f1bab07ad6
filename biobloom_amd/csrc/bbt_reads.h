/*
 * bbt_reads.h -- 2-bit packed reads (BBT_READS_2BIT) on the device.
 *
 * The kernels read bases as ASCII; a packed batch is first turned back into
 * ASCII in device memory (HBM-bound: 3/8 byte read and 1 byte written per
 * base).  Every kernel treats a base case-insensitively and any byte other
 * than A/C/G/T as a break in the k-mers, so the unpacked 'A' 'C' 'T' 'G' 'N'
 * bytes classify exactly like the original text.
 */
#ifndef BBT_READS_H
#define BBT_READS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bbt {

/* BBT_READS_2BIT -> ASCII: thread t turns code word t (32 bases) and mask
 * word t into 32 bytes, 'A' 'C' 'T' 'G' by code and 'N' where the mask bit
 * is set.  codes: code_bytes (a multiple of 8), the mask right after. */
static __global__ __launch_bounds__(256) void k_unpack_2bit(const uint8_t *codes,
                                                            uint64_t code_bytes,
                                                            uint8_t *ascii) {
    const uint64_t nw = code_bytes / 8;
    const uint64_t *cw = reinterpret_cast<const uint64_t *>(codes);
    const uint32_t *mw = reinterpret_cast<const uint32_t *>(codes + code_bytes);
    for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nw;
         t += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t c = cw[t];
        const uint32_t m = mw[t];
        uint64_t *out = reinterpret_cast<uint64_t *>(ascii + 32 * t);
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            uint64_t v = 0;
#pragma unroll
            for (uint32_t e = 0; e < 8; ++e) {
                const uint32_t b = 8 * q + e;
                const uint32_t code = uint32_t(c >> (2 * b)) & 3u;
                const uint64_t ch = ((m >> b) & 1) ? uint64_t('N')
                                                   : (0x47544341ull >> (8 * code)) & 0xFF;
                v |= ch << (8 * e);
            }
            out[q] = v;
        }
    }
}

} // namespace bbt

#endif
