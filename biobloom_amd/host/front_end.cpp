/*
 * front_end.cpp -- BioBloomClassifier over libbbt_hip (see front_end.h).
 *
 * Reads are parsed on the host in input order, gathered into device batches
 * (pinned staging -> HBM) and classified by bbt_eval_batch; the
 * ResultsManager counters accumulate on the device across batches.  Output
 * text follows BioBloomClassifier.h:115-529 and ResultsManager.hpp:191-237.
 */
#include "front_end.h"

#include <hip/hip_runtime.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
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
/* SeqReader                                                              */

SeqReader::SeqReader(const std::string &path) : m_buf(1 << 16) {
    m_fp = gzopen(path.c_str(), "r");
}

SeqReader::~SeqReader() {
    if (m_fp)
        gzclose(m_fp);
}

int SeqReader::getc_() {
    if (m_err)
        return -3;
    if (m_begin >= m_end) {
        if (m_eof)
            return -1;
        m_begin = 0;
        m_end = gzread(m_fp, m_buf.data(), unsigned(m_buf.size()));
        if (m_end == 0) {
            m_eof = true;
            return -1;
        }
        if (m_end < 0) {
            m_eof = m_err = true;
            return -3;
        }
    }
    return (unsigned char)m_buf[m_begin++];
}

/* Appends bytes up to a delimiter ('\n', or any isspace byte when
 * stop_at_space); the delimiter is consumed and reported.  Returns false
 * when nothing could be read at end of file. */
bool SeqReader::getline_(std::string &out, bool stop_at_space, int *delim) {
    bool got = false;
    if (delim)
        *delim = 0;
    while (true) {
        int c = getc_();
        if (c < 0)
            break;
        got = true;
        if (stop_at_space ? std::isspace(c) != 0 : c == '\n') {
            if (delim)
                *delim = c;
            return true;
        }
        out.push_back(char(c));
    }
    return got;
}

int64_t SeqReader::read(SeqRecord &rec) {
    int c;
    if (m_last == 0) {
        while ((c = getc_()) >= 0 && c != '>' && c != '@') {
        }
        if (c < 0)
            return c;
        m_last = c;
    }
    rec.name.clear();
    rec.comment.clear();
    rec.seq.clear();
    rec.qual.clear();
    int d = 0;
    if (!getline_(rec.name, true, &d))
        return m_err ? -3 : -1;
    if (d != '\n')
        getline_(rec.comment, false, nullptr);
    while ((c = getc_()) >= 0 && c != '>' && c != '+' && c != '@') {
        if (c == '\n')
            continue;
        rec.seq.push_back(char(c));
        getline_(rec.seq, false, nullptr);
    }
    if (c == '>' || c == '@')
        m_last = c;
    rec.fastq = (c == '+');
    if (!rec.fastq)
        return int64_t(rec.seq.size());
    while ((c = getc_()) >= 0 && c != '\n') {
    }
    if (c == -1)
        return -2;
    while (rec.qual.size() < rec.seq.size() && getline_(rec.qual, false, nullptr)) {
    }
    if (m_err)
        return -3;
    m_last = 0;
    if (rec.seq.size() != rec.qual.size())
        return -2;
    return int64_t(rec.seq.size());
}

/* ===================================================================== */
/* device batches                                                         */

struct BioBloomClassifier::Batch {
    hipStream_t st = nullptr;
    uint32_t nf = 0;
    bool scores = false;
    uint64_t cap_reads = 0, cap_bytes = 0;
    char *d_seq = nullptr, *h_seq = nullptr;
    uint64_t *d_off = nullptr, *h_off = nullptr;
    uint8_t *d_hits = nullptr, *h_hits = nullptr;
    double *d_scores = nullptr, *h_scores = nullptr;
    uint32_t *d_cls = nullptr, *h_cls = nullptr;
    uint64_t *d_summary = nullptr, *d_ctr = nullptr;

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
    m_b = new Batch();
    m_b->nf = m_nf;
    m_b->scores = m_opt.mode == BBT_BESTHIT || m_opt.mode == BBT_SCORES;
    HIP_CHECK(hipStreamCreate(&m_b->st));
    HIP_CHECK(hipMalloc(&m_b->d_summary, (2 * m_nf + 2) * 8));
    HIP_CHECK(hipMalloc(&m_b->d_ctr, 4 * 8));
    HIP_CHECK(hipMemset(m_b->d_summary, 0, (2 * m_nf + 2) * 8));
    HIP_CHECK(hipMemset(m_b->d_ctr, 0, 4 * 8));
}

BioBloomClassifier::~BioBloomClassifier() {
    if (m_b) {
        m_b->release();
        (void)hipFree(m_b->d_summary);
        (void)hipFree(m_b->d_ctr);
        (void)hipStreamDestroy(m_b->st);
        delete m_b;
    }
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

void BioBloomClassifier::run_batch(std::vector<SeqRecord> &recs, bool paired) {
    Batch &b = *m_b;
    const uint64_t n = recs.size();
    if (n == 0)
        return;
    uint64_t bytes = 0, maxlen = 0;
    for (const SeqRecord &r : recs) {
        bytes += r.seq.size();
        maxlen = std::max<uint64_t>(maxlen, r.seq.size());
    }
    if (maxlen > 0xFFFFFFFFull)
        die("read longer than 4 Gbp");
    b.reserve(n, bytes);
    uint64_t o = 0;
    for (uint64_t i = 0; i < n; ++i) {
        b.h_off[i] = o;
        std::memcpy(b.h_seq + o, recs[i].seq.data(), recs[i].seq.size());
        o += recs[i].seq.size();
    }
    b.h_off[n] = o;
    HIP_CHECK(hipMemcpyAsync(b.d_seq, b.h_seq, (o + 7) / 8 * 8, hipMemcpyHostToDevice, b.st));
    HIP_CHECK(hipMemcpyAsync(b.d_off, b.h_off, (n + 1) * 8, hipMemcpyHostToDevice, b.st));
    bbt_reads in;
    std::memset(&in, 0, sizeof(in));
    in.seq = b.d_seq;
    in.seq_bytes = b.cap_bytes;
    in.off = b.d_off;
    in.nreads = n;
    in.max_len = uint32_t(maxlen);
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
    out.summary = b.d_summary;
    out.counters = b.d_ctr;
    bbt_check(bbt_eval_batch(m_fs, &in, &p, &out, b.st));
    const uint64_t nrec = paired ? n / 2 : n;
    HIP_CHECK(hipMemcpyAsync(b.h_cls, b.d_cls, nrec * 4, hipMemcpyDeviceToHost, b.st));
    HIP_CHECK(hipMemcpyAsync(b.h_hits, b.d_hits, n * m_nf, hipMemcpyDeviceToHost, b.st));
    if (b.scores)
        HIP_CHECK(hipMemcpyAsync(b.h_scores, b.d_scores, n * m_nf * 8,
                                 hipMemcpyDeviceToHost, b.st));
    HIP_CHECK(hipStreamSynchronize(b.st));
}

/* iostream default formatting of a double, as the reference prints scores */
static std::string fmt(double v) {
    std::ostringstream s;
    s << v;
    return s.str();
}

namespace {
struct Out {
    gzFile gz = nullptr;
    FILE *f = nullptr;
    void write(const std::string &s) {
        if (gz)
            gzwrite(gz, s.data(), unsigned(s.size()));
        else if (f)
            std::fwrite(s.data(), 1, s.size(), f);
    }
};
} // namespace

void BioBloomClassifier::close_outputs(const std::string &) {
    for (void *v : m_outs) {
        Out *o = static_cast<Out *>(v);
        if (o->gz)
            gzclose(o->gz);
        if (o->f)
            std::fclose(o->f);
        delete o;
    }
    m_outs.clear();
}

static Out *open_one(const std::string &path, bool gz) {
    Out *o = new Out();
    if (gz)
        o->gz = gzopen(path.c_str(), "wb");
    else
        o->f = std::fopen(path.c_str(), "wb");
    if (!o->gz && !o->f)
        die("cannot open output file " + path);
    return o;
}

/* header line of one record for the partition files / stdout */
static std::string record_text(const SeqRecord &r, const std::string &type,
                               const std::string &extra) {
    std::string s;
    if (type == "fa") {
        s = ">" + r.name + " " + r.comment + extra + "\n" + r.seq + "\n";
    } else {
        s = "@" + r.name + " " + r.comment + extra + "\n" + r.seq + "\n+\n" + r.qual + "\n";
    }
    return s;
}

void BioBloomClassifier::emit_batch(const std::vector<SeqRecord> &recs, bool paired,
                                    const std::string &type) {
    const Batch &b = *m_b;
    const uint64_t nrec = paired ? recs.size() / 2 : recs.size();
    const uint32_t multi = m_nf + 1;
    for (uint64_t r = 0; r < nrec; ++r) {
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
            bool print = m_opt.inverse ? cls != 0 : cls == 0;
            if (print)
                for (int m = 0; m < nm; ++m) {
                    const SeqRecord &rec = recs[first + m];
                    std::string sc = (!m_opt.inverse && m_opt.mode == BBT_BESTHIT)
                                         ? " " + scoreStr[m] : "";
                    std::cout << "@" << rec.name << " " << rec.comment << sc << "\n"
                              << rec.seq << "\n+\n" << rec.qual << "\n";
                }
        }
        /* printSingleToFile / printPairToFile (BioBloomClassifier.h:144-209, 252-529) */
        if (!m_outs.empty()) {
            for (int m = 0; m < nm; ++m) {
                const SeqRecord &rec = recs[first + m];
                std::string add;
                if (m_opt.mode == BBT_SCORES && cls == multi)
                    add = extra[m];
                else if (m_opt.mode == BBT_BESTHIT)
                    add = (type == "fa" && cls == multi) ? extra[m] : " " + scoreStr[m];
                Out *o = static_cast<Out *>(m_outs[size_t(cls) * nm + m]);
                o->write(record_text(rec, type, add));
            }
        }
        if (m_status) {
            const std::string name = cls < m_nf ? m_ids[cls]
                                                : (cls == m_nf ? "noMatch" : "multiMatch");
            static_cast<Out *>(m_status)->write(recs[first].name + "\t" + name + "\n");
        }
    }
}

void BioBloomClassifier::progress(uint64_t &total, uint64_t add) {
    for (uint64_t i = 0; i < add; ++i)
        if (++total % m_opt.file_interval == 0)
            std::cerr << "Currently Reading Read Number: " << total << std::endl;
}

void BioBloomClassifier::write_summary(uint64_t totalReads) {
    std::vector<uint64_t> sm(2 * m_nf + 2);
    std::vector<uint64_t> ctr(4);
    HIP_CHECK(hipMemcpy(sm.data(), m_b->d_summary, sm.size() * 8, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(ctr.data(), m_b->d_ctr, 32, hipMemcpyDeviceToHost));
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
        if (o->f)
            std::fclose(o->f);
        delete o;
        m_status = nullptr;
    }
    std::cout.flush();
}

void BioBloomClassifier::filter(const std::vector<std::string> &inputFiles) {
    filterPrint(inputFiles, "");
}

void BioBloomClassifier::filterPrint(const std::vector<std::string> &inputFiles,
                                     const std::string &type) {
    if (!type.empty()) {
        m_gz = m_postfix == ".gz";
        std::vector<std::string> names = m_ids;
        names.push_back("noMatch");
        names.push_back("multiMatch");
        for (const std::string &n : names)
            m_outs.push_back(open_one(m_prefix + "_" + n + "." + type + m_postfix, m_gz));
    }
    if (m_opt.read_status)
        m_status = open_one(m_prefix + "_readStatus.tsv", false);
    std::cerr << "Filtering Start" << std::endl;
    uint64_t total = 0;
    std::vector<SeqRecord> recs;
    for (const std::string &path : inputFiles) {
        SeqReader rd(path);
        if (!rd.ok())
            die("file " + path + " cannot be opened");
        bool more = true;
        while (more) {
            recs.resize(m_opt.batch_reads);
            uint64_t n = 0;
            while (n < m_opt.batch_reads) {
                if (rd.read(recs[n]) < 0) {
                    more = false; /* end of file, or a truncated record as kseq */
                    break;
                }
                ++n;
            }
            recs.resize(n);
            progress(total, n);
            run_batch(recs, false);
            emit_batch(recs, false, type);
        }
    }
    std::cerr << "Total Reads: " << total << "\n";
    if (!type.empty()) {
        for (size_t i = 0; i < m_outs.size(); ++i) {
            Out *o = static_cast<Out *>(m_outs[i]);
            if (o->gz)
                gzclose(o->gz);
            if (o->f)
                std::fclose(o->f);
            delete o;
            std::string n = i < m_nf ? m_ids[i] : (i == m_nf ? "noMatch" : "multiMatch");
            std::cerr << "File written to: " << m_prefix + "_" + n + "." + type + m_postfix
                      << std::endl;
        }
        m_outs.clear();
    }
    write_summary(total);
}

void BioBloomClassifier::pair_loop(const std::string &f1, const std::string &f2,
                                   const std::string &type) {
    SeqReader r1(f1), r2(f2);
    if (!r1.ok())
        die("file " + f1 + " cannot be opened");
    if (!r2.ok())
        die("file " + f2 + " cannot be opened");
    std::vector<SeqRecord> recs;
    bool more = true;
    while (more) {
        recs.resize(2 * m_opt.batch_reads);
        uint64_t n = 0;
        while (n < m_opt.batch_reads) {
            /* the loop stops when either file ends (BioBloomClassifier.cpp:269, 291) */
            int64_t l1 = r1.read(recs[2 * n]);
            int64_t l2 = r2.read(recs[2 * n + 1]);
            if (l1 < 0 || l2 < 0) {
                more = false;
                break;
            }
            ++n;
        }
        recs.resize(2 * n);
        m_pairs += n;
        run_batch(recs, true);
        emit_batch(recs, true, type);
    }
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
    SeqReader rd(file);
    if (!rd.ok())
        die("file " + file + " cannot be opened");
    std::unordered_map<std::string, SeqRecord> waiting;
    std::vector<SeqRecord> recs;
    SeqRecord rec;
    bool more = true;
    while (more) {
        recs.clear();
        while (recs.size() < 2 * m_opt.batch_reads) {
            if (rd.read(rec) < 0) {
                more = false;
                break;
            }
            const size_t L = rec.name.size();
            if (L >= 2 && (rec.name[L - 1] == '1' || rec.name[L - 1] == '2'))
                rec.name.resize(L - 2);
            auto it = waiting.find(rec.name);
            if (it == waiting.end()) {
                std::string key = rec.name;
                waiting.emplace(std::move(key), std::move(rec));
            } else {
                recs.push_back(std::move(it->second));
                recs.push_back(std::move(rec));
                waiting.erase(it);
            }
            rec = SeqRecord();
        }
        m_pairs += recs.size() / 2;
        run_batch(recs, true);
        emit_batch(recs, true, type);
    }
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
    if (!type.empty()) {
        m_gz = m_postfix == ".gz";
        std::vector<std::string> names = m_ids;
        names.push_back("noMatch");
        names.push_back("multiMatch");
        for (const std::string &n : names) {
            m_outs.push_back(open_one(m_prefix + "_" + n + "_1." + type + m_postfix, m_gz));
            m_outs.push_back(open_one(m_prefix + "_" + n + "_2." + type + m_postfix, m_gz));
        }
    }
    if (m_opt.read_status)
        m_status = open_one(m_prefix + "_readStatus.tsv", false);
    std::cerr << "Filtering Start" << "\n";
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
    if (m_opt.read_status)
        m_status = open_one(m_prefix + "_readStatus.tsv", false);
    std::cerr << "Filtering Start" << "\n";
    m_pairs = 0;
    for (size_t i = 0; i < files1.size() && i < files2.size(); ++i)
        pair_loop(files1[i], files2[i], "");
    std::cerr << "Total Reads:" << m_pairs << std::endl;
    write_summary(m_pairs);
}

} // namespace bbt_fe
