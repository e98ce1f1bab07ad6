/*
 * deflate_rt.h -- libdeflate at run time, when the system has it.
 *
 * libdeflate decodes and encodes whole raw-deflate buffers faster than zlib.
 * It is loaded with dlopen through its public API (compressor and
 * decompressor objects, whole-buffer raw deflate, CRC32), so nothing links
 * against it: without it, or with BBT_NO_LIBDEFLATE=1 in the environment,
 * the callers use zlib.  Each thread keeps its own decompressor, and its own
 * compressor per level, until it ends.
 */
#ifndef BBT_DEFLATE_RT_H
#define BBT_DEFLATE_RT_H

#include <dlfcn.h>
#include <stddef.h>
#include <stdint.h>

#include <cstdlib>
#include <map>

namespace bbt_fe {

class Deflate {
  public:
    /* the library, or null when it is absent or switched off */
    static const Deflate *get() {
        static const Deflate d;
        const char *no = std::getenv("BBT_NO_LIBDEFLATE");
        return d.m_ok && !(no && no[0] == '1') ? &d : nullptr;
    }
    /* raw deflate data in[0..n) into exactly out_n bytes at out */
    bool inflate(const void *in, size_t n, void *out, size_t out_n) const {
        struct Holder {
            void *d = nullptr;
            void (*rel)(void *) = nullptr;
            ~Holder() {
                if (d)
                    rel(d);
            }
        };
        thread_local Holder t;
        if (!t.d) {
            t.d = m_alloc_d();
            t.rel = m_free_d;
        }
        size_t got = 0;
        return t.d && m_decompress(t.d, in, n, out, out_n, &got) == 0 && got == out_n;
    }
    /* raw deflate of in[0..n) at level into out (room for bound(n) bytes);
     * the compressed size, or 0 on failure */
    size_t deflate(int level, const void *in, size_t n, void *out, size_t room) const {
        void *c = compressor(level);
        return c ? m_compress(c, in, n, out, room) : 0;
    }
    size_t bound(int level, size_t n) const {
        void *c = compressor(level);
        return c ? m_bound(c, n) : 0;
    }
    uint32_t crc32(const void *p, size_t n) const { return m_crc(0, p, n); }

  private:
    Deflate() {
        void *h = ::dlopen("libdeflate.so.0", RTLD_NOW | RTLD_LOCAL);
        if (!h)
            return;
        m_alloc_d = reinterpret_cast<void *(*)()>(::dlsym(h, "libdeflate_alloc_decompressor"));
        m_decompress = reinterpret_cast<int (*)(void *, const void *, size_t, void *, size_t,
                                                size_t *)>(
            ::dlsym(h, "libdeflate_deflate_decompress"));
        m_free_d = reinterpret_cast<void (*)(void *)>(::dlsym(h, "libdeflate_free_decompressor"));
        m_alloc_c = reinterpret_cast<void *(*)(int)>(::dlsym(h, "libdeflate_alloc_compressor"));
        m_compress = reinterpret_cast<size_t (*)(void *, const void *, size_t, void *, size_t)>(
            ::dlsym(h, "libdeflate_deflate_compress"));
        m_bound = reinterpret_cast<size_t (*)(void *, size_t)>(
            ::dlsym(h, "libdeflate_deflate_compress_bound"));
        m_free_c = reinterpret_cast<void (*)(void *)>(::dlsym(h, "libdeflate_free_compressor"));
        m_crc = reinterpret_cast<uint32_t (*)(uint32_t, const void *, size_t)>(
            ::dlsym(h, "libdeflate_crc32"));
        m_ok = m_alloc_d && m_decompress && m_free_d && m_alloc_c && m_compress && m_bound &&
               m_free_c && m_crc;
    }
    void *compressor(int level) const {
        struct Holder {
            std::map<int, void *> c;
            void (*rel)(void *) = nullptr;
            ~Holder() {
                for (auto &kv : c)
                    if (kv.second)
                        rel(kv.second);
            }
        };
        thread_local Holder t;
        t.rel = m_free_c;
        void *&c = t.c[level];
        if (!c)
            c = m_alloc_c(level);
        return c;
    }
    bool m_ok = false;
    void *(*m_alloc_d)() = nullptr;
    int (*m_decompress)(void *, const void *, size_t, void *, size_t, size_t *) = nullptr;
    void (*m_free_d)(void *) = nullptr;
    void *(*m_alloc_c)(int) = nullptr;
    size_t (*m_compress)(void *, const void *, size_t, void *, size_t) = nullptr;
    size_t (*m_bound)(void *, size_t) = nullptr;
    void (*m_free_c)(void *) = nullptr;
    uint32_t (*m_crc)(uint32_t, const void *, size_t) = nullptr;
};

} // namespace bbt_fe

#endif
