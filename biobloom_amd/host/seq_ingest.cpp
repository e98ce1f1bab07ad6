/*
 * seq_ingest.cpp -- kseq_read restated over memory, the sequential reader
 * and the block-parallel reader (seq_ingest.h).
 */
#include "seq_ingest.h"

#include "deflate_rt.h"

#include <dlfcn.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>

namespace bbt_fe {

namespace {

/* isspace in the C locale: the bytes that end a name (KS_SEP_SPACE) */
inline bool ks_space(unsigned char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

/* Headroom in front of a raw gzip block, where the unfinished record of the
 * previous block is copied. */
constexpr size_t kHeadroom = size_t(4) << 20;
/* below this many bytes a block is parsed by one thread */
constexpr size_t kMinSplit = size_t(2) << 20;

enum : uint8_t { kArenaSeq = 4, kArenaQual = 8 };

/* A field assembled from lines the way ks_getuntil2 appends them: the first
 * line stays a view into the input, a second line moves the field into the
 * arena.  After each line read, a trailing '\r' is dropped when the field is
 * longer than 1 (kseq.h:139). */
struct Field {
    const char *p = nullptr;
    size_t len = 0, off = 0;
    bool arena = false;
    void add(const char *lp, size_t ll, std::string &ar) {
        if (!arena && !p) {
            p = lp;
            len = ll;
            return;
        }
        if (!arena) {
            off = ar.size();
            ar.append(p, len);
            arena = true;
        }
        ar.append(lp, ll);
        len += ll;
    }
    void strip_cr(std::string &ar) {
        if (len > 1 && (arena ? ar.back() : p[len - 1]) == '\r') {
            --len;
            if (arena)
                ar.pop_back();
        }
    }
    /* the view pointer: an arena offset while arena is true */
    const char *ptr() const {
        return arena ? reinterpret_cast<const char *>(uintptr_t(off)) : (p ? p : "");
    }
};

} // namespace

/* ===================================================================== */
/* MemKseq                                                                */

int MemKseq::line_view(const char *&p, size_t &len) {
    /* one ks_getuntil2(ks, KS_SEP_LINE, ...) read (kseq.h:93-143) */
    if (pos >= end)
        return eof ? 0 : -1;
    const char *nl = static_cast<const char *>(std::memchr(b + pos, '\n', end - pos));
    p = b + pos;
    if (!nl) {
        if (!eof)
            return -1;
        len = end - pos;
        pos = end;
    } else {
        len = size_t(nl - p);
        pos = size_t(nl - b) + 1;
    }
    return 1;
}

int64_t MemKseq::read_view(SeqView &v, std::string &arena) {
    /* kseq_read (kseq.h:177-219) */
    if (last_char == 0) {
        while (pos < end && b[pos] != '>' && b[pos] != '@')
            ++pos;
        if (pos >= end)
            return eof ? -1 : kStarved;
        last_char = (unsigned char)b[pos++];
    }
    v = SeqView();
    /* name: ks_getuntil(ks, 0, &name, &c) */
    if (pos >= end)
        return eof ? -1 : kStarved;
    const size_t s0 = pos;
    while (pos < end && !ks_space((unsigned char)b[pos]))
        ++pos;
    int c = 0;
    if (pos >= end && !eof)
        return kStarved;
    v.name = b + s0;
    v.name_len = uint32_t(pos - s0);
    if (pos < end)
        c = (unsigned char)b[pos++];
    /* comment: the rest of the header line */
    v.comment = "";
    if (c != '\n') {
        const char *cp;
        size_t cl;
        const int r = line_view(cp, cl);
        if (r < 0)
            return kStarved;
        if (r == 1) {
            if (cl > 1 && cp[cl - 1] == '\r')
                --cl;
            v.comment = cp;
            v.comment_len = uint32_t(cl);
        }
    }
    /* sequence lines until '>', '+' or '@' starts a line */
    Field seq;
    int ch = -1;
    while (true) {
        if (pos >= end) {
            if (!eof)
                return kStarved;
            ch = -1;
            break;
        }
        ch = (unsigned char)b[pos++];
        if (ch == '>' || ch == '+' || ch == '@')
            break;
        if (ch == '\n')
            continue;
        const char *lp = b + pos - 1; /* the line, its first byte included */
        const char *rp;
        size_t rl = 0;
        const int r = line_view(rp, rl);
        if (r < 0)
            return kStarved;
        seq.add(lp, 1 + (r ? rl : 0), arena);
        if (r)
            seq.strip_cr(arena);
    }
    if (ch == '>' || ch == '@')
        last_char = ch;
    v.seq = seq.ptr();
    v.seq_len = uint32_t(seq.len);
    if (seq.arena)
        v.in_arena |= kArenaSeq;
    v.fastq = ch == '+';
    v.qual = "";
    if (!v.fastq)
        return int64_t(seq.len);
    /* the rest of the '+' line */
    const char *nl = static_cast<const char *>(std::memchr(b + pos, '\n', end - pos));
    if (!nl) {
        if (!eof)
            return kStarved;
        pos = end;
        return -2;
    }
    pos = size_t(nl - b) + 1;
    Field qual;
    while (true) {
        const char *lp;
        size_t ll;
        const int r = line_view(lp, ll);
        if (r < 0)
            return kStarved;
        if (r == 0)
            break;
        qual.add(lp, ll, arena);
        qual.strip_cr(arena);
        if (qual.len >= seq.len)
            break;
    }
    last_char = 0;
    if (seq.len != qual.len)
        return -2;
    v.qual = qual.ptr();
    v.qual_len = uint32_t(qual.len);
    if (qual.arena)
        v.in_arena |= kArenaQual;
    return int64_t(seq.len);
}

void resolve_arena(SeqView *v, size_t n, const std::string &arena) {
    for (size_t i = 0; i < n; ++i) {
        if (!v[i].in_arena)
            continue;
        if (v[i].in_arena & kArenaSeq)
            v[i].seq = arena.data() + reinterpret_cast<uintptr_t>(v[i].seq);
        if (v[i].in_arena & kArenaQual)
            v[i].qual = arena.data() + reinterpret_cast<uintptr_t>(v[i].qual);
        v[i].in_arena = 0;
    }
}

int64_t MemKseq::read(SeqRecord &rec) {
    std::string arena;
    SeqView v;
    const int64_t l = read_view(v, arena);
    if (l < 0)
        return l;
    resolve_arena(&v, 1, arena);
    v.copy_to(rec);
    return l;
}

/* ===================================================================== */
/* SeqReader                                                              */

SeqReader::SeqReader(const std::string &path) : m_buf(size_t(1) << 20) {
    m_fp = gzopen(path.c_str(), "r");
    if (m_fp)
        gzbuffer(m_fp, 1u << 20);
    m_k.b = m_buf.data();
    m_k.pos = m_k.end = 0;
    m_k.eof = false;
}

SeqReader::~SeqReader() {
    if (m_fp)
        gzclose(m_fp);
}

/* Moves the unparsed bytes to the front and reads more (the buffer doubles
 * when it is full of one unfinished record). */
void SeqReader::fill() {
    const size_t keep = m_k.end - m_k.pos;
    std::memmove(m_buf.data(), m_buf.data() + m_k.pos, keep);
    if (keep * 2 > m_buf.size())
        m_buf.resize(m_buf.size() * 2);
    size_t n = keep;
    while (n < m_buf.size()) {
        const int r = gzread(m_fp, m_buf.data() + n,
                             unsigned(std::min<size_t>(m_buf.size() - n, size_t(1) << 30)));
        if (r <= 0) {
            m_eof = true;
            if (r < 0)
                m_status = -3;
            break;
        }
        n += size_t(r);
    }
    m_k.b = m_buf.data();
    m_k.pos = 0;
    m_k.end = n;
    m_k.eof = m_eof;
}

int64_t SeqReader::read(SeqRecord &rec) {
    if (!m_fp)
        return -3;
    while (true) {
        const size_t start = m_k.resume();
        const int64_t l = m_k.read(rec);
        if (l != MemKseq::kStarved) {
            if (l == -1 && m_status)
                return m_status;
            return l;
        }
        /* unfinished record: parse it again from its start with more bytes */
        m_k.pos = start;
        m_k.last_char = 0;
        fill();
    }
}

/* ===================================================================== */
/* ParallelSeqReader                                                      */

/* ---- BGZF ---------------------------------------------------------------- */

/* A BGZF stream (RFC 1952 members, each with the extra subfield 'B','C' of
 * length 2 holding the member's size - 1, at most 64 KiB of text each).
 * read() fills a buffer by inflating whole members on threads straight into
 * it (sizes from each member's ISIZE, contents checked against its CRC32);
 * the part of a member that does not fit waits in a small buffer.  A member
 * without the subfield switches the rest of the stream to sequential zlib,
 * as gzread would read it. */
class BgzfSource {
  public:
    BgzfSource(int fd, unsigned threads)
        : m_fd(fd), m_threads(std::max(1u, threads)), m_fast(Deflate::get()) {}
    ~BgzfSource() {
        if (m_zinit)
            inflateEnd(&m_z);
        ::close(m_fd);
    }
    /* true when the file at fd starts with a BGZF member */
    static bool detect(int fd) {
        unsigned char h[18];
        if (::pread(fd, h, sizeof(h), 0) != ssize_t(sizeof(h)))
            return false;
        size_t sz = 0;
        return member_size(h, sizeof(h), sz) == 1;
    }
    /* up to want bytes of text into out; eof once the stream is done */
    size_t read(char *out, size_t want, bool &eof, bool &err) {
        compact();
        size_t got = 0;
        while (got < want && !m_err && !m_done) {
            if (m_pend_at < m_pend.size()) {
                const size_t t = std::min(want - got, m_pend.size() - m_pend_at);
                std::memcpy(out + got, m_pend.data() + m_pend_at, t);
                m_pend_at += t;
                got += t;
                continue;
            }
            if (m_stream) {
                got += inflate_stream(out + got, want - got);
                continue;
            }
            /* the members that fit whole, inflated on threads */
            std::vector<Member> ms;
            const size_t room = want - got;
            size_t cpos = m_cpos, total = 0;
            Member m;
            int r = 1;
            while (ms.size() < 4096 && (r = next_member(cpos, m)) == 1 && m.isize <= room - total) {
                m.out = total;
                total += m.isize;
                ms.push_back(m);
                cpos += m.csize;
            }
            if (!ms.empty()) {
                /* the text stops before the first member that fails */
                std::atomic<size_t> first_bad{ms.size()};
                const unsigned T =
                    unsigned(std::min<size_t>(m_threads, std::max<size_t>(1, ms.size() / 8)));
                const auto work = [&](unsigned t) {
                    for (size_t i = t; i < ms.size(); i += T)
                        if (!inflate_member(ms[i], out + got + ms[i].out)) {
                            size_t cur = first_bad.load();
                            while (i < cur && !first_bad.compare_exchange_weak(cur, i)) {
                            }
                        }
                };
                std::vector<std::thread> th;
                for (unsigned t = 1; t < T; ++t)
                    th.emplace_back(work, t);
                work(0);
                for (std::thread &x : th)
                    x.join();
                if (first_bad < ms.size()) {
                    m_err = true;
                    got += ms[first_bad].out;
                } else {
                    got += total;
                }
                m_cpos = cpos;
            } else if (r == 1) {
                /* the next member does not fit: its text waits to the side */
                m_pend.resize(m.isize);
                m_pend_at = 0;
                if (!inflate_member(m, m_pend.data())) {
                    m_err = true;
                    m_pend.clear();
                }
                m_cpos = cpos + m.csize;
            } else if (r < 0) {
                /* not a BGZF member.  Bytes after a member that do not start
                 * another gzip member are trailing garbage, which gzread
                 * ignores: the input ends there.  Any other gzip member is
                 * inflated sequentially from here. */
                if (cpos > 0 && !gzip_magic(cpos))
                    m_done = true;
                else
                    m_stream = true;
            } else {
                m_done = true; /* no more input */
            }
        }
        if (m_err)
            err = true;
        eof = m_err || (m_done && m_pend_at == m_pend.size());
        return got;
    }

  private:
    struct Member {
        size_t at = 0, csize = 0, isize = 0, payload = 0, out = 0;
        uint32_t crc = 0;
    };
    /* 1: a BGZF member header (flags exactly FEXTRA, a 'B','C' subfield) of
     * total member size sz at h; 0: more header bytes needed; -1: not one */
    static int member_size(const unsigned char *h, size_t n, size_t &sz) {
        if (n < 12)
            return 0;
        if (h[0] != 0x1f || h[1] != 0x8b || h[2] != 8 || h[3] != 4)
            return -1;
        const size_t xlen = size_t(h[10]) | (size_t(h[11]) << 8);
        if (n < 12 + xlen)
            return 0;
        for (size_t x = 12; x + 4 <= 12 + xlen;) {
            const size_t slen = size_t(h[x + 2]) | (size_t(h[x + 3]) << 8);
            if (h[x] == 'B' && h[x + 1] == 'C' && slen == 2 && x + 6 <= 12 + xlen) {
                sz = (size_t(h[x + 4]) | (size_t(h[x + 5]) << 8)) + 1;
                return sz >= 12 + xlen + 8 ? 1 : -1;
            }
            x += 4 + slen;
        }
        return -1;
    }
    /* drops consumed compressed bytes once they are most of the buffer (only
     * between reads: collected members hold offsets into it) */
    void compact() {
        if (m_cpos < (size_t(8) << 20) || m_cpos * 2 < m_clen)
            return;
        std::memmove(m_cbuf.data(), m_cbuf.data() + m_cpos, m_clen - m_cpos);
        m_clen -= m_cpos;
        m_cpos = 0;
    }
    /* reads compressed bytes until need of them follow from */
    bool fill(size_t from, size_t need) {
        while (m_clen - from < need && !m_src_eof) {
            if (m_cbuf.size() < m_clen + (size_t(4) << 20))
                m_cbuf.resize(m_clen + (size_t(4) << 20));
            const ssize_t r = ::read(m_fd, m_cbuf.data() + m_clen, m_cbuf.size() - m_clen);
            if (r < 0) {
                m_err = true;
                return false;
            }
            if (r == 0)
                m_src_eof = true;
            m_clen += size_t(r);
        }
        return m_clen - from >= need;
    }
    /* true when the compressed bytes at pos start a gzip member (two magic
     * bytes; a lone last byte does not) */
    bool gzip_magic(size_t pos) {
        if (!fill(pos, 2))
            return false;
        const unsigned char *h = reinterpret_cast<unsigned char *>(m_cbuf.data() + pos);
        return h[0] == 0x1f && h[1] == 0x8b;
    }
    /* the member at cpos: 1 found, 0 end of input, -1 not BGZF (or cut short) */
    int next_member(size_t cpos, Member &m) {
        if (!fill(cpos, 12))
            return m_clen > cpos ? -1 : 0;
        const unsigned char *h0 = reinterpret_cast<unsigned char *>(m_cbuf.data() + cpos);
        const size_t xlen = size_t(h0[10]) | (size_t(h0[11]) << 8);
        if (!fill(cpos, 12 + xlen))
            return -1;
        size_t sz = 0;
        if (member_size(reinterpret_cast<unsigned char *>(m_cbuf.data() + cpos), m_clen - cpos,
                        sz) != 1 ||
            !fill(cpos, sz))
            return -1;
        const unsigned char *h = reinterpret_cast<unsigned char *>(m_cbuf.data() + cpos);
        m.at = cpos;
        m.csize = sz;
        m.payload = 12 + xlen;
        const unsigned char *t = h + sz - 8;
        m.crc = uint32_t(t[0]) | (uint32_t(t[1]) << 8) | (uint32_t(t[2]) << 16) |
                (uint32_t(t[3]) << 24);
        m.isize = size_t(t[4]) | (size_t(t[5]) << 8) | (size_t(t[6]) << 16) | (size_t(t[7]) << 24);
        return m.isize > (size_t(1) << 16) ? -1 : 1;
    }
    bool inflate_member(const Member &m, char *out) const {
        const char *in = m_cbuf.data() + m.at + m.payload;
        const size_t in_n = m.csize - m.payload - 8;
        if (m_fast)
            return m_fast->inflate(in, in_n, out, m.isize) &&
                   m_fast->crc32(out, m.isize) == m.crc;
        z_stream z;
        std::memset(&z, 0, sizeof(z));
        if (inflateInit2(&z, -15) != Z_OK)
            return false;
        z.next_in =
            reinterpret_cast<Bytef *>(const_cast<char *>(m_cbuf.data() + m.at + m.payload));
        z.avail_in = uInt(m.csize - m.payload - 8);
        z.next_out = reinterpret_cast<Bytef *>(out);
        z.avail_out = uInt(m.isize);
        const int r = inflate(&z, Z_FINISH);
        const bool ok = r == Z_STREAM_END && z.total_out == m.isize;
        inflateEnd(&z);
        return ok && uint32_t(crc32(0L, reinterpret_cast<Bytef *>(out), uInt(m.isize))) == m.crc;
    }
    /* sequential zlib over the rest of the stream (gzip members of any kind);
     * fills out to want unless the input ends (m_done) or is bad (m_err) */
    size_t inflate_stream(char *out, size_t want) {
        if (!m_zinit) {
            std::memset(&m_z, 0, sizeof(m_z));
            if (inflateInit2(&m_z, 15 + 16) != Z_OK) {
                m_err = true;
                return 0;
            }
            m_zinit = true;
        }
        size_t got = 0;
        while (got < want) {
            if (m_cpos >= m_clen && !fill(m_cpos, 1)) {
                /* the input ends (inside a member too: a stream cut short ends
                 * there with the text inflated so far, as gzread ends it) */
                m_done = true;
                break;
            }
            m_z.next_in = reinterpret_cast<Bytef *>(m_cbuf.data() + m_cpos);
            m_z.avail_in = uInt(std::min<size_t>(m_clen - m_cpos, size_t(1) << 30));
            m_z.next_out = reinterpret_cast<Bytef *>(out + got);
            m_z.avail_out = uInt(std::min<size_t>(want - got, size_t(1) << 30));
            const uInt in0 = m_z.avail_in, out0 = m_z.avail_out;
            const int r = inflate(&m_z, Z_NO_FLUSH);
            m_cpos += in0 - m_z.avail_in;
            got += out0 - m_z.avail_out;
            if (r == Z_STREAM_END) {
                inflateReset(&m_z);
                /* trailing garbage after a member ends the input (gzread) */
                if (!gzip_magic(m_cpos)) {
                    m_done = true;
                    break;
                }
            } else if (r != Z_OK && r != Z_BUF_ERROR) {
                m_err = true;
                break;
            }
        }
        return got;
    }

    int m_fd;
    unsigned m_threads;
    const Deflate *m_fast = nullptr;
    std::vector<char> m_cbuf; /* compressed bytes [0, m_clen), unread from m_cpos */
    size_t m_clen = 0, m_cpos = 0;
    std::vector<char> m_pend; /* text of a member that did not fit, from m_pend_at */
    size_t m_pend_at = 0;
    bool m_src_eof = false, m_done = false, m_err = false, m_stream = false;
    z_stream m_z;
    bool m_zinit = false;
};

ParallelSeqReader::ParallelSeqReader(const std::string &path, unsigned threads,
                                     size_t block_bytes)
    : m_threads(threads), m_block(std::max<size_t>(block_bytes, size_t(1) << 16)) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const bool chosen = m_threads != 0;
    if (!chosen)
        m_threads = std::min(16u, hw);
    m_fd = ::open(path.c_str(), O_RDONLY);
    if (m_fd < 0)
        return;
    /* a plain regular file is parsed in place from a read-only mapping;
     * anything else (gzip, pipes) goes through zlib, which also passes plain
     * data through */
    struct stat sb;
    unsigned char magic[2] = {0, 0};
    if (::fstat(m_fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size > 0 &&
        ::pread(m_fd, magic, 2, 0) >= 1 && !(magic[0] == 0x1f && magic[1] == 0x8b)) {
        void *p = ::mmap(nullptr, size_t(sb.st_size), PROT_READ, MAP_PRIVATE, m_fd, 0);
        if (p != MAP_FAILED) {
            m_map = static_cast<const char *>(p);
            m_map_len = size_t(sb.st_size);
            ::madvise(p, m_map_len, MADV_SEQUENTIAL);
        }
    }
    /* parsing mapped text is fast: past 8 threads the shares mostly wait
     * for memory and take cores from the rest of the pipeline */
    if (m_map && !chosen)
        m_threads = std::min(8u, hw);
    if (!m_map && S_ISREG(sb.st_mode) && BgzfSource::detect(m_fd)) {
        m_bgzf.reset(new BgzfSource(m_fd, m_threads)); /* takes the descriptor */
        m_fd = -1;
    } else if (!m_map) {
        m_fp = gzdopen(m_fd, "r"); /* takes the descriptor */
        m_fd = -1;
        if (!m_fp)
            return;
        gzbuffer(m_fp, 1u << 20);
    }
    m_ok = true;
    m_thread = std::thread([this] { run(); });
}

ParallelSeqReader::~ParallelSeqReader() {
    m_cur.reset();
    {
        std::lock_guard<std::mutex> g(m_mu);
        m_stop = true;
    }
    m_cv.notify_all();
    if (m_thread.joinable())
        m_thread.join();
    for (SeqBlock *b : m_full)
        delete b;
    for (SeqBlock *b : m_free)
        delete b;
    if (m_fp)
        gzclose(m_fp);
    if (m_map)
        ::munmap(const_cast<char *>(m_map), m_map_len);
    if (m_fd >= 0)
        ::close(m_fd);
}

size_t ParallelSeqReader::fill_raw(std::vector<char> &buf, size_t at, size_t want, bool &eof,
                                   bool &err) {
    if (m_bgzf)
        return m_bgzf->read(buf.data() + at, want, eof, err);
    size_t got = 0;
    while (got < want) {
        const size_t ask = std::min<size_t>(want - got, size_t(1) << 30);
        const long r = long(gzread(m_fp, buf.data() + at + got, unsigned(ask)));
        if (r < 0) {
            err = eof = true;
            break;
        }
        if (r == 0) {
            eof = true;
            break;
        }
        got += size_t(r);
    }
    return got;
}

namespace {

/* First position >= from where a record plausibly starts: a '>' at a line
 * start, or an '@' at a line start two lines above a line that starts with
 * '+'.  n when there is none. */
size_t guess_start(const char *b, size_t n, size_t from) {
    size_t p = from;
    while (p < n) {
        if (p > 0) {
            const char *nl = static_cast<const char *>(std::memchr(b + p - 1, '\n', n - p + 1));
            if (!nl)
                return n;
            p = size_t(nl - b) + 1;
            if (p >= n)
                return n;
        }
        if (b[p] == '>')
            return p;
        if (b[p] == '@') {
            const char *e1 = static_cast<const char *>(std::memchr(b + p, '\n', n - p));
            const char *e2 = e1 ? static_cast<const char *>(
                                      std::memchr(e1 + 1, '\n', size_t(b + n - (e1 + 1))))
                                : nullptr;
            if (e2 && e2 + 1 < b + n && e2[1] == '+')
                return p;
        }
        ++p;
    }
    return n;
}

struct ShareResult {
    size_t next = 0;    /* resume position after the share's last record */
    int64_t status = 0; /* 0 handed over at next >= stop; kStarved; or a kseq status < 0 */
};

/* Parses records that start in [start, stop) (a record may end beyond stop)
 * into out[0..count), multi-line fields into arena. */
ShareResult parse_share(const char *b, size_t start, size_t stop, size_t n, bool eof,
                        std::vector<SeqView> &out, size_t &count, std::string &arena) {
    MemKseq k;
    k.b = b;
    k.pos = start;
    k.end = n;
    k.eof = eof;
    count = 0;
    arena.clear();
    ShareResult res;
    while (true) {
        const size_t r = k.resume();
        /* hand over at stop; a share that reaches the end of the input reads
         * on to kseq's end-of-file return */
        if (r >= stop && !(eof && stop >= n)) {
            res.next = r;
            break;
        }
        if (count == out.size())
            out.resize(out.size() + out.size() / 2 + 1024);
        const int64_t l = k.read_view(out[count], arena);
        if (l < 0) {
            res.next = r;
            res.status = l;
            break;
        }
        ++count;
    }
    resolve_arena(out.data(), count, arena);
    return res;
}

} // namespace

void ParallelSeqReader::parse_block(const char *b, size_t n, bool eof, SeqBlock &out,
                                    size_t &carry_from) {
    const unsigned T = n < std::min(kMinSplit, m_block / 4) ? 1u : m_threads;
    if (out.parts.size() < T) {
        out.parts.resize(T);
        out.arena.resize(T);
    }
    out.count.assign(out.parts.size(), 0);
    std::vector<size_t> g(T + 1);
    g[0] = 0;
    for (unsigned w = 1; w < T; ++w)
        g[w] = std::max(g[w - 1], guess_start(b, n, n / T * w));
    g[T] = n;
    std::vector<ShareResult> res(T);
    std::vector<std::thread> th;
    for (unsigned w = 1; w < T; ++w)
        th.emplace_back([&, w] {
            res[w] = parse_share(b, g[w], g[w + 1], n, eof, out.parts[w], out.count[w],
                                 out.arena[w]);
        });
    res[0] = parse_share(b, g[0], g[1], n, eof, out.parts[0], out.count[0], out.arena[0]);
    for (std::thread &t : th)
        t.join();
    /* keep each share only if the previous one handed over exactly at its start */
    size_t expected = 0;
    int64_t st = 0;
    for (unsigned w = 0; w < T; ++w) {
        bool rest = false;
        if (g[w] != expected) {
            /* a wrong guess: this share parses the rest of the block */
            ++m_reparsed;
            res[w] = parse_share(b, expected, n, n, eof, out.parts[w], out.count[w],
                                 out.arena[w]);
            rest = true;
        }
        expected = res[w].next;
        st = res[w].status;
        if (st != 0 || rest) {
            for (unsigned x = w + 1; x < T; ++x)
                out.count[x] = 0;
            break;
        }
    }
    carry_from = n;
    out.status = 0;
    if (st == MemKseq::kStarved)
        carry_from = expected;
    else if (st < 0)
        out.status = st;
}

SeqBlock *ParallelSeqReader::take_free() {
    std::unique_lock<std::mutex> lk(m_mu);
    m_cv.wait(lk, [&] { return m_stop || !m_free.empty() || m_live < ParallelSeqReader::kMaxBlocks; });
    if (m_stop)
        return nullptr;
    if (m_free.empty()) {
        ++m_live;
        return new SeqBlock();
    }
    SeqBlock *b = m_free.front();
    m_free.pop_front();
    return b;
}

void ParallelSeqReader::publish(SeqBlock *b) {
    {
        std::lock_guard<std::mutex> g(m_mu);
        m_full.push_back(b);
    }
    m_cv.notify_all();
}

void ParallelSeqReader::recycle(SeqBlock *b) {
    {
        std::lock_guard<std::mutex> g(m_mu);
        m_free.push_back(b);
    }
    m_cv.notify_all();
}

void ParallelSeqReader::run() {
    if (m_map)
        run_mapped();
    else
        run_stream();
}

/* A mapped plain file: blocks are windows of the mapping; an unfinished
 * record at a window's end starts the next window (the window doubles when
 * one record fills it). */
void ParallelSeqReader::run_mapped() {
    size_t base = 0, win = m_block;
    while (true) {
        SeqBlock *blk = take_free();
        if (!blk)
            return;
        const size_t n = std::min(win, m_map_len - base);
        const bool eof = base + n == m_map_len;
        size_t carry_from = 0;
        parse_block(m_map + base, n, eof, *blk, carry_from);
        ++m_blocks;
        const bool last = blk->status < 0;
        publish(blk);
        if (last)
            return;
        win = carry_from == 0 && !eof ? win * 2 : m_block;
        base += carry_from;
    }
}

/* A zlib stream (gzip, a pipe, or a file that cannot be mapped): the block's
 * bytes are the unfinished record of the previous block followed by the next
 * decompressed bytes; the block keeps its bytes, which its views point into. */
void ParallelSeqReader::run_stream() {
    std::vector<char> A(kHeadroom + m_block), B, next;
    bool eof = false, err = false;
    size_t aoff = kHeadroom;
    size_t an = fill_raw(A, kHeadroom, m_block, eof, err);
    while (true) {
        SeqBlock *blk = take_free();
        if (!blk)
            return;
        /* parse A on the share threads while the next raw block is read */
        size_t carry_from = 0, ngot = 0;
        bool neof = eof, nerr = false;
        std::thread reader;
        if (!eof) {
            if (B.size() < kHeadroom + m_block)
                B.resize(kHeadroom + m_block);
            reader = std::thread([&] { ngot = fill_raw(B, kHeadroom, m_block, neof, nerr); });
        }
        parse_block(A.data() + aoff, an, eof, *blk, carry_from);
        if (reader.joinable())
            reader.join();
        if (err && blk->status == -1)
            blk->status = -3;
        ++m_blocks;
        /* the next work buffer: the unfinished record, then the new raw bytes */
        const size_t carry = an - carry_from;
        const char *cp = A.data() + aoff + carry_from;
        size_t naoff = kHeadroom;
        if (eof) {
            next.assign(kHeadroom + carry, 0);
            std::memcpy(next.data() + kHeadroom, cp, carry);
        } else if (carry <= kHeadroom) {
            std::memcpy(B.data() + kHeadroom - carry, cp, carry);
            next.swap(B);
            naoff = kHeadroom - carry;
        } else {
            next.resize(kHeadroom + carry + ngot);
            std::memcpy(next.data() + kHeadroom, cp, carry);
            std::memcpy(next.data() + kHeadroom + carry, B.data() + kHeadroom, ngot);
        }
        /* the block owns the bytes its views point into; its old buffer is
         * reused for the next raw read */
        blk->bytes.swap(A);
        const bool last = blk->status < 0;
        publish(blk);
        if (last)
            return;
        B.swap(A);
        A.swap(next);
        aoff = naoff;
        an = carry + (eof ? 0 : ngot);
        eof = neof;
        err = err || nerr;
    }
}

ParallelSeqReader::BlockRef ParallelSeqReader::next_block() {
    if (!m_ok || m_ended)
        return nullptr;
    SeqBlock *b = nullptr;
    {
        std::unique_lock<std::mutex> lk(m_mu);
        m_cv.wait(lk, [&] { return !m_full.empty(); });
        b = m_full.front();
        m_full.pop_front();
    }
    if (b->status < 0)
        m_ended = true;
    return BlockRef(b, [this](SeqBlock *x) { recycle(x); });
}

int64_t ParallelSeqReader::read(SeqRecord &rec) {
    if (!m_ok)
        return -3;
    while (true) {
        if (m_done)
            return m_final;
        if (m_cur) {
            while (m_part < m_cur->parts.size()) {
                if (m_idx < m_cur->count[m_part]) {
                    const SeqView &v = m_cur->parts[m_part][m_idx++];
                    v.copy_to(rec);
                    return int64_t(v.seq_len);
                }
                ++m_part;
                m_idx = 0;
            }
            const int64_t st = m_cur->status;
            m_cur.reset();
            if (st < 0) {
                m_done = true;
                m_final = st;
                return st;
            }
        }
        m_cur = next_block();
        m_part = m_idx = 0;
    }
}

} // namespace bbt_fe
