/*
 * front_end.cpp -- BioBloomClassifier over libbbt_hip (see front_end.h).
 *
 * Reads are parsed by ParallelSeqReader (parsing threads run ahead of the
 * device), gathered into device batches (pinned staging -> HBM) and
 * classified by bbt_eval_batch in a two-slot pipeline; the ResultsManager
 * counters accumulate on the device across batches.  Output text follows
 * BioBloomClassifier.h:115-529 and ResultsManager.hpp:191-237.
 */
#include "front_end.h"
#include "read_pack.h"

#include "deflate_rt.h"

#include <hip/hip_runtime.h>
#include <unistd.h>

#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <iostream>
#include <mutex>
#include <sstream>
#include <thread>
#include <unordered_map>

namespace bbt_fe {

static void die(const std::string &msg) {
    std::cerr << msg << std::endl;
    std::exit(1);
}

#define HIP_CHECK(expr)                                                           \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            die(std::string("HIP error: ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

static void bbt_check(bbt_status s) {
    if (s != BBT_OK)
        die(bbt_last_error());
}

/* ===================================================================== */
/* device batches                                                         */

/* One of the two pipeline slots: pinned staging, device buffers, a stream and
 * a workspace for one batch of reads.  The host packs and emits one slot while
 * the device works on the other; both add to the same summary counters. */
struct BioBloomClassifier::Batch {
    hipStream_t st = nullptr;
    hipEvent_t walked = nullptr; /* recorded after the slot's kernels */
    hipEvent_t started = nullptr; /* before them, once the other slot's are done */
    bool walk_recorded = false;
    bbt_workspace *ws = nullptr;
    uint32_t nf = 0;
    bool scores = false;
    uint64_t cap_reads = 0, cap_bytes = 0;
    uint64_t n = 0; /* reads packed / in flight */
    uint64_t max_len = 0;
    uint64_t code_bytes = 0; /* 2-bit batch: bytes of codes (the mask follows) */
    bool paired = false;
    char *d_seq = nullptr, *h_seq = nullptr;
    uint64_t *d_off = nullptr, *h_off = nullptr;
    uint8_t *d_hits = nullptr, *h_hits = nullptr;
    double *d_scores = nullptr, *h_scores = nullptr;
    uint32_t *d_cls = nullptr, *h_cls = nullptr;

    void release() {
        (void)hipFree(d_seq);
        (void)hipFree(d_off);
        (void)hipFree(d_hits);
        (void)hipFree(d_scores);
        (void)hipFree(d_cls);
        (void)hipHostFree(h_seq);
        (void)hipHostFree(h_off);
        (void)hipHostFree(h_hits);
        (void)hipHostFree(h_scores);
        (void)hipHostFree(h_cls);
        d_seq = h_seq = nullptr;
        d_off = h_off = nullptr;
        d_hits = h_hits = nullptr;
        d_scores = h_scores = nullptr;
        d_cls = h_cls = nullptr;
    }
    void reserve(uint64_t reads, uint64_t bytes) {
        bytes = (bytes + 7) / 8 * 8 + 8;
        if (reads <= cap_reads && bytes <= cap_bytes)
            return;
        HIP_CHECK(hipStreamSynchronize(st));
        release();
        cap_reads = std::max<uint64_t>(reads, cap_reads);
        cap_bytes = std::max<uint64_t>(bytes, cap_bytes);
        HIP_CHECK(hipMalloc(&d_seq, cap_bytes));
        HIP_CHECK(hipMalloc(&d_off, (cap_reads + 1) * 8));
        HIP_CHECK(hipMalloc(&d_hits, cap_reads * nf));
        HIP_CHECK(hipMalloc(&d_cls, cap_reads * 4));
        HIP_CHECK(hipHostMalloc(&h_seq, cap_bytes, hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc(&h_off, (cap_reads + 1) * 8, hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc(&h_hits, cap_reads * nf, hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc(&h_cls, cap_reads * 4, hipHostMallocDefault));
        if (scores) {
            HIP_CHECK(hipMalloc(&d_scores, cap_reads * nf * 8));
            HIP_CHECK(hipHostMalloc(&h_scores, cap_reads * nf * 8, hipHostMallocDefault));
        }
    }
};

BioBloomClassifier::BioBloomClassifier(const std::vector<std::string> &paths,
                                       double scoreThreshold, const std::string &prefix,
                                       const std::string &postfix, const Options &opt)
    : m_threshold(scoreThreshold), m_prefix(prefix), m_postfix(postfix), m_opt(opt) {
    /* loadFilters (BioBloomClassifier.cpp:950-971) */
    std::cerr << "Starting to Load Filters." << std::endl;
    std::vector<const char *> cp;
    for (const std::string &p : paths)
        cp.push_back(p.c_str());
    bbt_check(bbt_filters_load(cp.data(), uint32_t(cp.size()), &m_fs));
    m_nf = bbt_filterset_size(m_fs);
    for (uint32_t i = 0; i < m_nf; ++i) {
        m_ids.push_back(bbt_filter_id(m_fs, i));
        std::cerr << "Loaded Filter: " + m_ids.back();
        if (m_opt.scorer == BBT_BINOMIAL) {
            bbt_filter_info inf;
            bbt_check(bbt_filter_info_get(m_fs, i, &inf));
            std::cerr << " FPR: " << inf.fpr;
        }
        std::cerr << std::endl;
    }
    std::cerr << "Filter Loading Complete." << std::endl;
    if (!m_opt.tune_file.empty() && access(m_opt.tune_file.c_str(), R_OK) == 0) {
        /* knobs tuned earlier for this filter size and read length */
        bbt_check(bbt_tuning_load(m_opt.tune_file.c_str()));
        std::cerr << "Loaded tuning table: " << m_opt.tune_file << std::endl;
    }
    /* each slot has its own stream and workspace: one slot's copies overlap
     * the other's kernels */
    for (Batch *&b : m_slot) {
        b = new Batch();
        b->nf = m_nf;
        b->scores = m_opt.mode == BBT_BESTHIT || m_opt.mode == BBT_SCORES;
        HIP_CHECK(hipStreamCreate(&b->st));
        HIP_CHECK(hipEventCreate(&b->walked));
        HIP_CHECK(hipEventCreate(&b->started));
        bbt_check(bbt_workspace_create(&b->ws));
    }
    HIP_CHECK(hipMalloc(&m_dsummary, (2 * m_nf + 2) * 8));
    HIP_CHECK(hipMalloc(&m_dctr, 4 * 8));
    HIP_CHECK(hipMemset(m_dsummary, 0, (2 * m_nf + 2) * 8));
    HIP_CHECK(hipMemset(m_dctr, 0, 4 * 8));
}

BioBloomClassifier::~BioBloomClassifier() {
    finish_writes();
    for (Batch *&b : m_slot) {
        if (!b)
            continue;
        (void)hipStreamSynchronize(b->st);
        b->release();
        (void)hipEventDestroy(b->walked);
        (void)hipEventDestroy(b->started);
        bbt_workspace_free(b->ws);
        (void)hipStreamDestroy(b->st);
        delete b;
        b = nullptr;
    }
    (void)hipFree(m_dsummary);
    (void)hipFree(m_dctr);
    bbt_filters_free(m_fs);
}

void BioBloomClassifier::setOrderedFilter() {
    /* BioBloomClassifier.h:67-75 */
    if (m_opt.mode == BBT_BESTHIT) {
        std::cerr << "Best Hit mode and Ordered mode detected. Not yet supported."
                  << std::endl;
        std::exit(1);
    }
    m_opt.mode = BBT_ORDERED;
}

/* Packs recs into slot s's pinned staging (the device may still be working
 * on the other slot). */
void BioBloomClassifier::pack(int s, const RecBatch &recs, bool paired, uint64_t scale) {
    Batch &b = *m_slot[s];
    const uint64_t n = recs.v.size();
    b.n = n;
    b.paired = paired;
    if (n == 0)
        return;
    uint64_t bytes = 0, maxlen = 0;
    for (const SeqView &r : recs.v) {
        bytes += r.seq_len;
        maxlen = std::max<uint64_t>(maxlen, r.seq_len);
    }
    /* reserve for scale times this batch (the full batch size while the
     * batches ramp up): growing the pinned and device buffers synchronizes
     * the device */
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t need = m_opt.pack2bit ? bbt_host::packed_bytes(bytes) : bytes;
    b.reserve(n * scale, need * scale);
    m_phase[4] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    uint64_t o = 0;
    for (uint64_t i = 0; i < n; ++i) {
        b.h_off[i] = o;
        o += recs.v[i].seq_len;
    }
    b.h_off[n] = o;
    b.max_len = maxlen;
    if (m_opt.pack2bit) {
        /* 2-bit codes and the N mask straight from the records' views */
        const auto get = [&](uint64_t r) { return recs.v[r].seq; };
        b.code_bytes = bbt_host::pack_2bit(get, b.h_off, n, reinterpret_cast<uint8_t *>(b.h_seq),
                                           8);
        return;
    }
    /* gather the bases into the pinned staging buffer on a few threads */
    const unsigned T = o < (uint64_t(8) << 20) ? 1u : 8u;
    const auto gather = [&](unsigned t) {
        for (uint64_t i = n * t / T, e = n * (t + 1) / T; i < e; ++i)
            std::memcpy(b.h_seq + b.h_off[i], recs.v[i].seq, recs.v[i].seq_len);
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; ++t)
        th.emplace_back(gather, t);
    gather(0);
    for (std::thread &t : th)
        t.join();
    b.max_len = maxlen;
}

/* Enqueues copy-in, classification and copy-out of slot s; returns without
 * waiting.  The copy-in overlaps the other slot's kernels; the kernels start
 * after the other slot's, so launches run one at a time in submission order
 * and each keeps the whole device (and its L2) to itself. */
void BioBloomClassifier::submit(int s) {
    Batch &b = *m_slot[s];
    const Batch &other = *m_slot[1 - s];
    const uint64_t n = b.n;
    const bool paired = b.paired;
    if (n == 0)
        return;
    const uint64_t o = b.h_off[n];
    const uint64_t copy = m_opt.pack2bit ? b.code_bytes * 3 / 2 : (o + 7) / 8 * 8;
    HIP_CHECK(hipMemcpyAsync(b.d_seq, b.h_seq, copy, hipMemcpyHostToDevice, b.st));
    HIP_CHECK(hipMemcpyAsync(b.d_off, b.h_off, (n + 1) * 8, hipMemcpyHostToDevice, b.st));
    if (other.walk_recorded)
        HIP_CHECK(hipStreamWaitEvent(b.st, other.walked, 0));
    HIP_CHECK(hipEventRecord(b.started, b.st));
    bbt_reads in;
    std::memset(&in, 0, sizeof(in));
    in.seq = b.d_seq;
    in.seq_bytes = m_opt.pack2bit ? b.code_bytes : b.cap_bytes;
    in.format = m_opt.pack2bit ? BBT_READS_2BIT : BBT_READS_ASCII;
    in.off = b.d_off;
    in.nreads = n;
    in.max_len = uint32_t(b.max_len);
    bbt_eval_params p;
    std::memset(&p, 0, sizeof(p));
    p.mode = m_opt.mode;
    p.scorer = m_opt.scorer;
    p.threshold = m_threshold;
    p.streak = m_opt.streak;
    p.paired = paired;
    p.inclusive = m_opt.inclusive;
    p.path = m_opt.path;
    p.dust = m_opt.dust ? 1u : 0u;
    p.dust_t = m_opt.dust_t;
    p.dust_w = m_opt.dust_w;
    bbt_eval_out out;
    std::memset(&out, 0, sizeof(out));
    out.hits = b.d_hits;
    out.scores = b.d_scores;
    out.cls = b.d_cls;
    out.summary = m_dsummary;
    out.counters = m_dctr;
    if (m_opt.autotune && !m_autotuned && m_opt.path == BBT_PATH_BATCHED &&
        m_opt.mode != BBT_ORDERED) {
        /* the knobs timed on this first batch; outputs only (the trials
         * must not add to the summary) */
        m_autotuned = true;
        bbt_eval_out t_out;
        std::memset(&t_out, 0, sizeof(t_out));
        t_out.hits = b.d_hits;
        t_out.scores = b.d_scores;
        bbt_tuning best;
        bbt_check(bbt_autotune(m_fs, &in, &p, &t_out, b.st, &best));
        std::cerr << "Autotuned: reads of " << best.wave_min_len
                  << "+ bases per wavefront, refill at " << best.refill_min
                  << " idle lanes, " << best.blocks_per_cu << " blocks per CU, probe unroll "
                  << best.probe_ahead << std::endl;
        if (!m_opt.tune_file.empty())
            bbt_check(bbt_tuning_save(m_opt.tune_file.c_str()));
    }
    bbt_check(bbt_eval_batch_ws(m_fs, b.ws, &in, &p, &out, b.st));
    HIP_CHECK(hipEventRecord(b.walked, b.st));
    b.walk_recorded = true;
    const uint64_t nrec = paired ? n / 2 : n;
    HIP_CHECK(hipMemcpyAsync(b.h_cls, b.d_cls, nrec * 4, hipMemcpyDeviceToHost, b.st));
    HIP_CHECK(hipMemcpyAsync(b.h_hits, b.d_hits, n * m_nf, hipMemcpyDeviceToHost, b.st));
    if (b.scores)
        HIP_CHECK(hipMemcpyAsync(b.h_scores, b.d_scores, n * m_nf * 8,
                                 hipMemcpyDeviceToHost, b.st));
}

/* Waits for slot s's launch. */
void BioBloomClassifier::finish(int s) {
    Batch &b = *m_slot[s];
    if (!b.n)
        return;
    HIP_CHECK(hipStreamSynchronize(b.st));
    /* the device time of the batch's kernels (from the copy-in or the end of
     * the other slot's kernels, whichever is later, to the walk's end) */
    float ms = 0;
    if (hipEventElapsedTime(&ms, b.started, b.walked) == hipSuccess)
        m_phase[5] += ms / 1e3;
}

/* Blocks one batch may hold: the three batches alive at once (one on the
 * device, one being packed, one being gathered) stay well inside the reader's
 * pool, so the reader can always parse ahead. */
constexpr size_t kBatchBlocks = (ParallelSeqReader::kMaxBlocks - 4) / 3;

/* Pipeline over a batch source: next(recs) fills the next batch and returns
 * false once the input is exhausted (recs may still hold records).  While the
 * device classifies batch i, a helper thread packs and submits batch i+1 and
 * this thread gathers batch i+2 (parsing runs ahead on the reader's threads);
 * batch i is then emitted, so outputs keep input order and the device always
 * has the next launch queued.  The first batch is an eighth of the batch size
 * and each next one twice the last, so the device starts while the rest of
 * the input is still being parsed. */
void BioBloomClassifier::pipeline(const std::function<bool(RecBatch &)> &next, bool paired,
                                  const std::string &type) {
    using clk = std::chrono::steady_clock;
    const auto since = [](clk::time_point t) {
        return std::chrono::duration<double>(clk::now() - t).count();
    };
    RecBatch recs[3];
    uint64_t scale[3] = {1, 1, 1};
    bool more = true;
    uint64_t limit = std::max<uint64_t>(1, m_opt.batch_reads / 8);
    /* gathers the next batch into recs[k] (this thread only) */
    const auto gather = [&](int k) {
        const clk::time_point t = clk::now();
        recs[k].clear();
        m_batch_limit = limit;
        /* while the batches ramp up, pack() reserves for the full size */
        scale[k] = std::max<uint64_t>(1, m_opt.batch_reads / limit);
        if (more)
            more = next(recs[k]);
        limit = std::min(m_opt.batch_reads, 2 * limit);
        m_phase[0] += since(t);
    };
    gather(0);
    clk::time_point t = clk::now();
    pack(0, recs[0], paired, scale[0]);
    submit(0);
    m_phase[1] += since(t);
    gather(1);
    for (uint64_t i = 0;; ++i) {
        /* batch i is on the device (slot i % 2); batch i - 1 has been emitted,
         * so slot (i + 1) % 2 and recs[(i + 2) % 3] are free */
        const int b0 = int(i % 3), b1 = int((i + 1) % 3), b2 = int((i + 2) % 3);
        const int s0 = int(i % 2), s1 = int((i + 1) % 2);
        double pack_s = 0;
        std::thread packer([&] {
            const clk::time_point tp = clk::now();
            pack(s1, recs[b1], paired, scale[b1]);
            submit(s1);
            pack_s = since(tp);
        });
        gather(b2);
        packer.join();
        m_phase[1] += pack_s;
        t = clk::now();
        finish(s0);
        m_phase[2] += since(t);
        t = clk::now();
        emit_batch(s0, recs[b0], paired, type);
        recs[b0].clear(); /* releases its blocks to the reader */
        m_phase[3] += since(t);
        if (recs[b1].v.empty())
            break;
    }
}

/* iostream default formatting of a double, as the reference prints scores */
static std::string fmt(double v) {
    std::ostringstream s;
    s << v;
    return s.str();
}

/* text compressed as BGZF members (gzip members of at most 65280 bytes of
 * text, each with the 'B','C' extra subfield holding the member's size - 1),
 * at zlib's default level (6) as gzopen "wb" writes, with libdeflate when the
 * system has it: any gzip reader returns the concatenated text, and BGZF
 * readers (this one included) can inflate the members in parallel */
static void bgzf_members(const std::string &in, std::string &out) {
    constexpr size_t kText = 65280;
    constexpr int kLevel = 6; /* Z_DEFAULT_COMPRESSION */
    const Deflate *fast = Deflate::get();
    out.clear();
    for (size_t i = 0; i < in.size(); i += kText) {
        const size_t n = std::min(kText, in.size() - i);
        const size_t h = out.size();
        size_t total = 0;
        if (fast) {
            const size_t room = fast->bound(kLevel, n);
            out.resize(h + 18 + room + 8);
            const size_t c = fast->deflate(kLevel, in.data() + i, n, &out[h + 18], room);
            if (!c)
                die("libdeflate: compression failed");
            total = 18 + c + 8;
        } else {
            z_stream z;
            std::memset(&z, 0, sizeof(z));
            if (deflateInit2(&z, kLevel, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK)
                die("zlib: deflateInit2 failed");
            out.resize(h + 18 + deflateBound(&z, uLong(n)) + 8);
            z.next_in = reinterpret_cast<Bytef *>(const_cast<char *>(in.data() + i));
            z.avail_in = uInt(n);
            z.next_out = reinterpret_cast<Bytef *>(&out[h + 18]);
            z.avail_out = uInt(out.size() - h - 26);
            if (deflate(&z, Z_FINISH) != Z_STREAM_END)
                die("zlib: deflate failed");
            total = 18 + z.total_out + 8;
            deflateEnd(&z);
        }
        static const unsigned char kHead[16] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0,
                                                0,    0xff, 6, 0, 'B', 'C', 2, 0};
        std::memcpy(&out[h], kHead, 16);
        out[h + 16] = char((total - 1) & 0xff);
        out[h + 17] = char((total - 1) >> 8);
        const uint32_t crc =
            fast ? fast->crc32(in.data() + i, n)
                 : uint32_t(crc32(0L, reinterpret_cast<const Bytef *>(in.data() + i), uInt(n)));
        unsigned char *t = reinterpret_cast<unsigned char *>(&out[h + total - 8]);
        for (int b = 0; b < 4; ++b) {
            t[b] = (unsigned char)(crc >> (8 * b));
            t[4 + b] = (unsigned char)(uint32_t(n) >> (8 * b));
        }
        out.resize(h + total);
    }
}

/* the empty member that ends a BGZF file */
static const char kBgzfEof[28] = {'\x1f', '\x8b', 8, 4, 0, 0, 0, 0, 0, '\xff', 6, 0, 'B', 'C',
                                  2, 0, 27, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0};

namespace {
/* an output file; gzip files receive whole BGZF members */
struct Out {
    FILE *f = nullptr;
    bool gz = false;
    void write(const std::string &s) {
        if (!s.empty() && std::fwrite(s.data(), 1, s.size(), f) != s.size())
            die("write failed");
    }
    void close() {
        /* a gzip file ends with the empty member, so even an unwritten one is
         * a valid (empty) gzip file */
        if (gz && std::fwrite(kBgzfEof, 1, sizeof(kBgzfEof), f) != sizeof(kBgzfEof))
            die("write failed");
        std::fclose(f);
    }
};
} // namespace

/* Writes the formatted text of finished batches, in order, on its own
 * thread: buf[t][k] is thread range t's text for output k (the partition
 * files, then the read status, then stdout).  At most one batch waits while
 * another is written. */
struct BioBloomClassifier::Writer {
    using Text = std::vector<std::vector<std::string>>;
    std::vector<void *> outs;
    void *status = nullptr;
    bool to_stdout = false;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Text> q;
    bool busy = false, stop = false;

    Writer(const std::vector<void *> &o, void *st, bool so) : outs(o), status(st), to_stdout(so) {
        th = std::thread([this] { loop(); });
    }
    ~Writer() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv.notify_all();
        th.join();
    }
    void push(Text &&t) {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&] { return q.empty(); });
        q.push_back(std::move(t));
        cv.notify_all();
    }
    void drain() {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&] { return q.empty() && !busy; });
    }
    void loop() {
        while (true) {
            Text t;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return stop || !q.empty(); });
                if (q.empty())
                    return;
                t = std::move(q.front());
                q.pop_front();
                busy = true;
            }
            cv.notify_all();
            const size_t n = outs.size();
            for (const std::vector<std::string> &b : t) {
                for (size_t k = 0; k < n; ++k)
                    static_cast<Out *>(outs[k])->write(b[k]);
                if (status)
                    static_cast<Out *>(status)->write(b[n]);
                if (to_stdout)
                    std::cout.write(b[n + 1].data(), std::streamsize(b[n + 1].size()));
            }
            {
                std::lock_guard<std::mutex> g(mu);
                busy = false;
            }
            cv.notify_all();
        }
    }
};

/* waits for the writer and ends it (before outputs are closed) */
void BioBloomClassifier::finish_writes() {
    if (m_writer) {
        m_writer->drain();
        delete m_writer;
        m_writer = nullptr;
    }
}

void BioBloomClassifier::close_outputs(const std::string &) {
    finish_writes();
    for (void *v : m_outs) {
        Out *o = static_cast<Out *>(v);
        o->close();
        delete o;
    }
    m_outs.clear();
}

static Out *open_one(const std::string &path, bool gz) {
    Out *o = new Out();
    o->f = std::fopen(path.c_str(), "wb");
    o->gz = gz;
    if (!o->f)
        die("cannot open output file " + path);
    return o;
}

/* one record of a partition file appended to o (header with the extra
 * score text, sequence, and for FASTQ the quality) */
static void append_record(std::string &o, const SeqView &r, const std::string &type,
                          const std::string &extra) {
    o += type == "fa" ? '>' : '@';
    o.append(r.name, r.name_len);
    o += ' ';
    o.append(r.comment, r.comment_len);
    o += extra;
    o += '\n';
    o.append(r.seq, r.seq_len);
    o += '\n';
    if (type != "fa") {
        o += "+\n";
        o.append(r.qual, r.qual_len);
        o += '\n';
    }
}

/* The output text of a finished batch: records in ranges are formatted on
 * threads into one buffer per output (partition files, read status,
 * stdout), gzip outputs compressed there too (one member per buffer), then
 * the buffers are written in input order. */
void BioBloomClassifier::emit_batch(int s, const RecBatch &batch, bool paired,
                                    const std::string &type) {
    const Batch &b = *m_slot[s];
    const std::vector<SeqView> &recs = batch.v;
    const uint64_t nrec = paired ? recs.size() / 2 : recs.size();
    const uint32_t multi = m_nf + 1;
    if (!m_stdout && m_outs.empty() && !m_status)
        return; /* summary only: the counters are on the device */
    const size_t nouts = m_outs.size(), sink_status = nouts, sink_stdout = nouts + 1;
    const unsigned T =
        nrec < 4096 ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<std::string>> buf(T, std::vector<std::string>(nouts + 2));
    const auto work = [&](unsigned t) {
        std::vector<std::string> &o = buf[t];
        for (uint64_t r = nrec * t / T, re = nrec * (t + 1) / T; r < re; ++r) {
            const uint32_t cls = b.h_cls[r];
            const uint64_t first = paired ? 2 * r : r;
            const int nm = paired ? 2 : 1;
            std::string extra[2];
            std::string scoreStr[2];
            for (int m = 0; m < nm; ++m) {
                const uint64_t rd = first + m;
                if (!b.scores)
                    continue;
                const double *srow = b.h_scores + rd * m_nf;
                const uint8_t *hrow = b.h_hits + rd * m_nf;
                if (m_opt.mode == BBT_BESTHIT) {
                    /* evaluateReadBestHit: max score (from 0), repeated per best hit */
                    double mx = 0;
                    for (uint32_t i = 0; i < m_nf; ++i)
                        if (mx < srow[i])
                            mx = srow[i];
                    scoreStr[m] = fmt(mx);
                    for (uint32_t i = 0; i < m_nf; ++i)
                        if (hrow[i])
                            extra[m] += " " + fmt(mx);
                } else {
                    for (uint32_t i = 0; i < m_nf; ++i)
                        extra[m] += " " + fmt(srow[i]);
                }
            }
            /* -d / -n: printSingle / printPair (BioBloomClassifier.h:115-142, 211-250) */
            if (m_stdout) {
                const bool print = m_opt.inverse ? cls != 0 : cls == 0;
                if (print)
                    for (int m = 0; m < nm; ++m) {
                        const SeqView &rec = recs[first + m];
                        std::string &d = o[sink_stdout];
                        d += '@';
                        d.append(rec.name, rec.name_len);
                        d += ' ';
                        d.append(rec.comment, rec.comment_len);
                        if (!m_opt.inverse && m_opt.mode == BBT_BESTHIT) {
                            d += ' ';
                            d += scoreStr[m];
                        }
                        d += '\n';
                        d.append(rec.seq, rec.seq_len);
                        d += "\n+\n";
                        d.append(rec.qual, rec.qual_len);
                        d += '\n';
                    }
            }
            /* printSingleToFile / printPairToFile (BioBloomClassifier.h:144-209, 252-529) */
            if (nouts) {
                for (int m = 0; m < nm; ++m) {
                    const SeqView &rec = recs[first + m];
                    std::string add;
                    if (m_opt.mode == BBT_SCORES && cls == multi)
                        add = extra[m];
                    else if (m_opt.mode == BBT_BESTHIT)
                        add = (type == "fa" && cls == multi) ? extra[m] : " " + scoreStr[m];
                    append_record(o[size_t(cls) * nm + m], rec, type, add);
                }
            }
            if (m_status) {
                std::string &d = o[sink_status];
                d.append(recs[first].name, recs[first].name_len);
                d += '\t';
                d += cls < m_nf ? m_ids[cls] : (cls == m_nf ? "noMatch" : "multiMatch");
                d += '\n';
            }
        }
        if (m_gz)
            for (size_t k = 0; k < nouts; ++k)
                if (!o[k].empty()) {
                    std::string z;
                    bgzf_members(o[k], z);
                    o[k].swap(z);
                }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; ++t)
        th.emplace_back(work, t);
    work(0);
    for (std::thread &x : th)
        x.join();
    if (!m_writer)
        m_writer = new Writer(m_outs, m_status, m_stdout);
    m_writer->push(std::move(buf));
}

void BioBloomClassifier::progress(uint64_t &total, uint64_t add) {
    for (uint64_t i = 0; i < add; ++i)
        if (++total % m_opt.file_interval == 0)
            std::cerr << "Currently Reading Read Number: " << total << std::endl;
}

void BioBloomClassifier::write_summary(uint64_t totalReads) {
    finish_writes();
    std::vector<uint64_t> sm(2 * m_nf + 2);
    std::vector<uint64_t> ctr(4);
    for (Batch *b : m_slot)
        HIP_CHECK(hipStreamSynchronize(b->st));
    HIP_CHECK(hipMemcpy(sm.data(), m_dsummary, sm.size() * 8, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(ctr.data(), m_dctr, 32, hipMemcpyDeviceToHost));
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - m_t0).count();
    std::cerr << "Filtering time: " << secs << " s (gather " << m_phase[0] << ", pack+launch "
              << m_phase[1] << " incl. staging allocation " << m_phase[4] << ", device wait "
              << m_phase[2] << ", output " << m_phase[3] << "; device kernels " << m_phase[5]
              << ")" << std::endl;
    m_probes = ctr[0];
    m_lookups = ctr[1];
    uint64_t need = 0;
    bbt_check(bbt_summary_text(m_fs, sm.data(), totalReads, nullptr, 0, &need));
    std::vector<char> buf(need + 1);
    bbt_check(bbt_summary_text(m_fs, sm.data(), totalReads, buf.data(), need + 1, nullptr));
    const std::string path = m_prefix + "_summary.tsv";
    std::cerr << "Writing file: " << path << std::endl;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        die("cannot open " + path);
    std::fwrite(buf.data(), 1, need, f);
    std::fclose(f);
    if (m_status) {
        Out *o = static_cast<Out *>(m_status);
        o->close();
        delete o;
        m_status = nullptr;
    }
    std::cout.flush();
}

void BioBloomClassifier::filter(const std::vector<std::string> &inputFiles) {
    filterPrint(inputFiles, "");
}

void BioBloomClassifier::open_outputs(const std::string &type, bool paired) {
    if (!type.empty()) {
        m_gz = m_postfix == ".gz";
        std::vector<std::string> names = m_ids;
        names.push_back("noMatch");
        names.push_back("multiMatch");
        for (const std::string &n : names) {
            if (paired) {
                m_outs.push_back(open_one(m_prefix + "_" + n + "_1." + type + m_postfix, m_gz));
                m_outs.push_back(open_one(m_prefix + "_" + n + "_2." + type + m_postfix, m_gz));
            } else {
                m_outs.push_back(open_one(m_prefix + "_" + n + "." + type + m_postfix, m_gz));
            }
        }
    }
    if (m_opt.read_status)
        m_status = open_one(m_prefix + "_readStatus.tsv", false);
}

void BioBloomClassifier::filterPrint(const std::vector<std::string> &inputFiles,
                                     const std::string &type) {
    open_outputs(type, false);
    std::cerr << "Filtering Start" << std::endl;
    m_t0 = std::chrono::steady_clock::now();
    uint64_t total = 0;
    for (const std::string &path : inputFiles) {
        ParallelSeqReader rd(path, m_opt.threads, m_opt.ingest_block);
        if (!rd.ok())
            die("file " + path + " cannot be opened");
        SeqCursor cur(rd);
        pipeline(
            [&](RecBatch &batch) {
                bool more = true;
                while (batch.v.size() < m_batch_limit &&
                       batch.holds.size() < kBatchBlocks) {
                    /* stops at end of file, or a truncated record as kseq */
                    const SeqView *v = cur.next();
                    if (!v) {
                        more = false;
                        break;
                    }
                    batch.v.push_back(*v);
                    batch.hold(cur.block());
                }
                progress(total, batch.v.size());
                return more;
            },
            false, type);
    }
    std::cerr << "Total Reads: " << total << "\n";
    if (!type.empty()) {
        for (size_t i = 0; i < m_outs.size(); ++i) {
            std::string n = i < m_nf ? m_ids[i] : (i == m_nf ? "noMatch" : "multiMatch");
            std::cerr << "File written to: " << m_prefix + "_" + n + "." + type + m_postfix
                      << std::endl;
        }
        close_outputs(type);
    }
    write_summary(total);
}

void BioBloomClassifier::pair_loop(const std::string &f1, const std::string &f2,
                                   const std::string &type) {
    ParallelSeqReader r1(f1, m_opt.threads, m_opt.ingest_block),
        r2(f2, m_opt.threads, m_opt.ingest_block);
    if (!r1.ok())
        die("file " + f1 + " cannot be opened");
    if (!r2.ok())
        die("file " + f2 + " cannot be opened");
    SeqCursor c1(r1), c2(r2);
    pipeline(
        [&](RecBatch &batch) {
            bool more = true;
            while (batch.v.size() < 2 * m_batch_limit &&
                   batch.holds.size() < kBatchBlocks) {
                /* the loop stops when either file ends (BioBloomClassifier.cpp:269, 291) */
                const SeqView *a = c1.next();
                const SeqView *b = c2.next();
                if (!a || !b) {
                    more = false;
                    break;
                }
                batch.v.push_back(*a);
                batch.hold(c1.block());
                batch.v.push_back(*b);
                batch.hold(c2.block());
            }
            m_pairs += batch.v.size() / 2;
            return more;
        },
        true, type);
}

/* Single-file ("smart pair") input, BioBloomClassifier.cpp:464-562 and
 * 572-690: a record's key is its name less the last two characters when the
 * name ends in '1' or '2' (the key is also the name printed); the first record
 * of a key waits, the next record with that key completes the pair as mate 2,
 * and records still waiting at the end of the file are not classified.  Pairs
 * are classified in the order they complete, the reference's order on one
 * thread.  A name shorter than two characters is kept whole (the reference
 * would take a negative length there). */
void BioBloomClassifier::smart_pair_loop(const std::string &file, const std::string &type) {
    ParallelSeqReader rd(file, m_opt.threads, m_opt.ingest_block);
    if (!rd.ok())
        die("file " + file + " cannot be opened");
    SeqCursor cur(rd);
    std::unordered_map<std::string, SeqRecord> waiting;
    pipeline(
        [&](RecBatch &batch) {
            bool more = true;
            while (batch.v.size() < 2 * m_batch_limit &&
                   batch.holds.size() < kBatchBlocks) {
                const SeqView *v = cur.next();
                if (!v) {
                    more = false;
                    break;
                }
                uint32_t L = v->name_len;
                if (L >= 2 && (v->name[L - 1] == '1' || v->name[L - 1] == '2'))
                    L -= 2;
                std::string key(v->name, L);
                auto it = waiting.find(key);
                if (it == waiting.end()) {
                    /* the first mate waits, copied out of its block */
                    SeqRecord &r = waiting[key];
                    v->copy_to(r);
                    r.name = key;
                } else {
                    batch.owned.push_back(std::move(it->second));
                    waiting.erase(it);
                    batch.v.push_back(view_of(batch.owned.back()));
                    SeqView second = *v;
                    second.name_len = L;
                    batch.v.push_back(second);
                    batch.hold(cur.block());
                }
            }
            m_pairs += batch.v.size() / 2;
            return more;
        },
        true, type);
}

void BioBloomClassifier::filterPair(const std::string &file1, const std::string &file2) {
    filterPairPrint(file1, file2, "");
}

void BioBloomClassifier::filterPair(const std::string &file) {
    pair_run(file, nullptr, "");
}

void BioBloomClassifier::filterPairPrint(const std::string &file, const std::string &type) {
    pair_run(file, nullptr, type);
}

void BioBloomClassifier::filterPairPrint(const std::string &file1, const std::string &file2,
                                         const std::string &type) {
    pair_run(file1, &file2, type);
}

void BioBloomClassifier::pair_run(const std::string &file1, const std::string *file2,
                                  const std::string &type) {
    open_outputs(type, true);
    std::cerr << "Filtering Start" << "\n";
    m_t0 = std::chrono::steady_clock::now();
    m_pairs = 0;
    if (file2)
        pair_loop(file1, *file2, type);
    else
        smart_pair_loop(file1, type);
    close_outputs(type);
    std::cerr << "Total Reads:" << m_pairs << std::endl;
    write_summary(m_pairs);
}

void BioBloomClassifier::filterPair(const std::vector<std::string> &files1,
                                    const std::vector<std::string> &files2) {
    /* file-list mode (BioBloomClassifier.cpp:744-942): pairs of files in turn */
    open_outputs("", true);
    std::cerr << "Filtering Start" << "\n";
    m_t0 = std::chrono::steady_clock::now();
    m_pairs = 0;
    for (size_t i = 0; i < files1.size() && i < files2.size(); ++i)
        pair_loop(files1[i], files2[i], "");
    std::cerr << "Total Reads:" << m_pairs << std::endl;
    write_summary(m_pairs);
}

} // namespace bbt_fe
