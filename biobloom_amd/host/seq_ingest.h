/*
 * seq_ingest.h -- FASTA/FASTQ(.gz) ingest for the front ends.
 *
 * Both readers deliver exactly the records, fields and return codes of the
 * reference's kseq_read (Common/kseq.h:177-219, line reads :93-143):
 *   name     up to the first isspace byte; the rest of the header line is
 *            the comment (its trailing '\r' dropped when it is longer than 1)
 *   seq      the following lines up to a line that starts with '>', '+' or
 *            '@', empty lines skipped, each line's '\r' dropped once the
 *            sequence is longer than 1
 *   qual     after the '+' line, whole lines until it is as long as seq
 *   return   the sequence length; -1 at end of file; -2 for a missing or
 *            wrong-length quality; -3 for a stream error
 * A caller stops at the first negative return, as the reference's loops do
 * (BioBloomClassifier.cpp:61-70).
 *
 * SeqReader is the sequential restatement.  ParallelSeqReader splits large
 * blocks of the input (windows of a mapped plain file, or decompressed
 * blocks of a gzip stream) among threads: each thread starts at a guessed
 * record start and parses its share with the same state machine; a share is
 * kept only when the previous share ended exactly where it began (otherwise
 * the rest of the block is parsed again from the true position), so its
 * records are always the sequential reader's.  Records are handed out as
 * views into the input bytes (a field that spans lines is assembled in a
 * per-share arena), so parsing copies nothing for one-line records.
 * Decompression of the next block overlaps the parse of the current one, and
 * parsed blocks queue up ahead of the caller, so parsing also overlaps the
 * caller's device work.  A BGZF file (blocked gzip, as bgzip writes: every
 * member at most 64 KiB and its compressed size in the header) is inflated
 * member by member on the parse threads.
 */
#ifndef BBT_SEQ_INGEST_H
#define BBT_SEQ_INGEST_H

#include <stdint.h>
#include <zlib.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <string_view>
#include <thread>
#include <utility>
#include <vector>

namespace bbt_fe {

/* One FASTA/FASTQ record as kseq_read delivers it. */
struct SeqRecord {
    std::string name, comment, seq, qual;
    bool fastq = false;
};

/* A record's fields in place: pointers into the input bytes or a share's
 * arena, valid while the block that holds the record is alive. */
struct SeqView {
    const char *name = nullptr, *comment = nullptr, *seq = nullptr, *qual = nullptr;
    uint32_t name_len = 0, comment_len = 0, seq_len = 0, qual_len = 0;
    bool fastq = false;
    uint8_t in_arena = 0; /* while parsing: bit f set = field f is an arena offset */
    std::string_view name_sv() const { return {name, name_len}; }
    std::string_view comment_sv() const { return {comment, comment_len}; }
    std::string_view seq_sv() const { return {seq, seq_len}; }
    std::string_view qual_sv() const { return {qual, qual_len}; }
    void copy_to(SeqRecord &r) const {
        r.name.assign(name, name_len);
        r.comment.assign(comment, comment_len);
        r.seq.assign(seq, seq_len);
        r.qual.assign(qual, qual_len);
        r.fastq = fastq;
    }
};

/* kseq_read over bytes in memory, b[pos..end).  When eof is false the bytes
 * are a prefix of the stream: a record that reaches end is not finished and
 * the read returns kStarved instead. */
struct MemKseq {
    static constexpr int64_t kStarved = -100;
    const char *b = nullptr;
    size_t pos = 0, end = 0;
    bool eof = true;
    int last_char = 0; /* header byte already consumed (kseq_t::last_char) */
    /* where parsing with last_char == 0 would resume to the same effect */
    size_t resume() const { return last_char ? pos - 1 : pos; }
    /* fields that span lines are appended to arena; their view pointers then
     * hold arena offsets and in_arena marks them (resolve_arena fixes them) */
    int64_t read_view(SeqView &v, std::string &arena);
    int64_t read(SeqRecord &rec);

  private:
    int line_view(const char *&p, size_t &len); /* 1 read, 0 nothing at eof, -1 starved */
};

/* Turns the arena offsets of views parsed into arena into pointers. */
void resolve_arena(SeqView *v, size_t n, const std::string &arena);

/* Sequential FASTA/FASTQ(.gz) reader. */
class SeqReader {
  public:
    explicit SeqReader(const std::string &path);
    ~SeqReader();
    bool ok() const { return m_fp != nullptr; }
    int64_t read(SeqRecord &rec);

  private:
    void fill();
    gzFile m_fp = nullptr;
    std::vector<char> m_buf;
    MemKseq m_k;
    bool m_eof = false;
    int64_t m_status = 0;
};

/* Parsed records of one block of input. */
struct SeqBlock {
    std::vector<std::vector<SeqView>> parts; /* per share, in input order */
    std::vector<size_t> count;               /* views used in each part */
    std::vector<std::string> arena;          /* per share */
    std::vector<char> bytes;                 /* the block's input bytes (gzip input) */
    int64_t status = 0; /* 0: more blocks follow; < 0: kseq status after these records */
    size_t size() const {
        size_t n = 0;
        for (size_t c : count)
            n += c;
        return n;
    }
};

class BgzfSource;

/* Multi-threaded reader with SeqReader's exact output. */
class ParallelSeqReader {
  public:
    using BlockRef = std::shared_ptr<SeqBlock>;
    /* parsed blocks alive at once, queued ahead or held by the caller: a
     * caller that holds this many blocks waits forever for the next one */
    static constexpr unsigned kMaxBlocks = 48;
    /* threads = parse threads (0: the hardware's, at most 8 for a mapped
     * plain file and 16 for compressed or streamed input) */
    explicit ParallelSeqReader(const std::string &path, unsigned threads = 0,
                               size_t block_bytes = size_t(16) << 20);
    /* every BlockRef must be released before the reader is destroyed */
    ~ParallelSeqReader();
    bool ok() const { return m_ok; }
    /* The next block of records, in input order; null after the block whose
     * status is negative.  Views stay valid while the reference is held. */
    BlockRef next_block();
    /* Record by record (copies the fields); kseq_read contract, and after a
     * negative return every call returns it again. */
    int64_t read(SeqRecord &rec);
    /* parse statistics: blocks, shares re-parsed after a wrong guess */
    uint64_t blocks() const { return m_blocks; }
    uint64_t reparsed() const { return m_reparsed; }

  private:
    void run();
    void run_mapped();
    void run_stream();
    SeqBlock *take_free();
    void publish(SeqBlock *b);
    void recycle(SeqBlock *b);
    size_t fill_raw(std::vector<char> &buf, size_t at, size_t want, bool &eof, bool &err);
    void parse_block(const char *b, size_t n, bool eof, SeqBlock &out, size_t &carry_from);

    gzFile m_fp = nullptr;
    std::unique_ptr<BgzfSource> m_bgzf; /* BGZF file, inflated on threads */
    const char *m_map = nullptr; /* plain file, mapped */
    size_t m_map_len = 0;
    bool m_ok = false;
    int m_fd = -1;
    unsigned m_threads;
    size_t m_block;
    std::thread m_thread;
    std::mutex m_mu;
    std::condition_variable m_cv;
    std::deque<SeqBlock *> m_full, m_free;
    unsigned m_live = 0; /* blocks allocated */
    bool m_stop = false, m_ended = false;
    /* read(): the current block */
    BlockRef m_cur;
    size_t m_part = 0, m_idx = 0;
    int64_t m_final = 0;
    bool m_done = false;
    std::atomic<uint64_t> m_blocks{0}, m_reparsed{0};
};

/* Record-by-record walk over a ParallelSeqReader's blocks, without copying:
 * next() returns a view that stays valid while block() (or a copy of that
 * reference) is held. */
class SeqCursor {
  public:
    explicit SeqCursor(ParallelSeqReader &r) : m_r(r) {}
    /* the next record, or null at the end (status() is then kseq's return) */
    const SeqView *next() {
        while (true) {
            if (m_blk) {
                while (m_part < m_blk->parts.size()) {
                    if (m_idx < m_blk->count[m_part])
                        return &m_blk->parts[m_part][m_idx++];
                    ++m_part;
                    m_idx = 0;
                }
                const int64_t st = m_blk->status;
                m_blk.reset();
                if (st < 0) {
                    m_status = st;
                    m_end = true;
                }
            }
            if (m_end)
                return nullptr;
            m_blk = m_r.next_block();
            m_part = m_idx = 0;
            if (!m_blk) {
                m_end = true;
                if (m_status == 0)
                    m_status = -3; /* the reader could not open its input */
                return nullptr;
            }
        }
    }
    const ParallelSeqReader::BlockRef &block() const { return m_blk; }
    int64_t status() const { return m_status; }

  private:
    ParallelSeqReader &m_r;
    ParallelSeqReader::BlockRef m_blk;
    size_t m_part = 0, m_idx = 0;
    bool m_end = false;
    int64_t m_status = 0;
};

/* A view of a record held in SeqRecord strings. */
inline SeqView view_of(const SeqRecord &r) {
    SeqView v;
    v.name = r.name.data();
    v.name_len = uint32_t(r.name.size());
    v.comment = r.comment.data();
    v.comment_len = uint32_t(r.comment.size());
    v.seq = r.seq.data();
    v.seq_len = uint32_t(r.seq.size());
    v.qual = r.qual.data();
    v.qual_len = uint32_t(r.qual.size());
    v.fastq = r.fastq;
    return v;
}

} // namespace bbt_fe

#endif
