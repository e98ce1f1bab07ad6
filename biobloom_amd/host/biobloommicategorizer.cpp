/*
 * biobloommicategorizer -- the reference's multi-index Bloom filter
 * categorizer (BioBloomCategorizer/BioBloomMICategorizer.cpp:120-344,
 * MIBFClassifier.hpp:47-939) over libbbt_hip.  Options, the per-read output
 * lines on stdout (TSV, or FASTA/FASTQ with the hits in the header,
 * MIBFClassifier.hpp:652-736) and <prefix>_summary.tsv (ResultsManager over
 * "repeat" and the IDs) follow the reference; reads are classified in device
 * batches by bbt_mibf_query_batch.
 *
 * Differences kept on purpose: reads are counted for the summary in every
 * thread setting (the reference counts them only in its multi-threaded
 * reader); -r (streak) does not stop a query early, every frame is
 * evaluated, and a warning says so when -r is given (the streak rule lives in
 * the absent btl_bloomfilter MIBFQuerySupport); --debug is not available.
 */
#include <getopt.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "front_end.h"
#include "read_pack.h"

using namespace std;

namespace {

enum OutType { TSV = 0, FASTA = 1, FASTQ = 2 };

struct Opt {
    string prefix, filter;
    bool paired = false, hitOnly = false, bestAgree = false;
    double score = std::pow(10.0, -10.0);
    unsigned frameMatches = 1, streak = 3, minNoSat = 0;
    double multi = 1.0;
    uint64_t interval = 10000000;
    int out = TSV;
    int verbose = 0;
    unsigned threads = 0; /* parse threads per file (0: the reader's default) */
    uint64_t batch = uint64_t(1) << 17; /* reads (or pairs) per device batch */
    bool pack2bit = true; /* send reads 2-bit packed (false: ASCII) */
};

/* candidates a batch's query keeps per record; the few records with more
 * are queried again with room for all of them */
constexpr uint32_t kFirstHits = 4;

void usage() {
    cerr << "Usage: biobloommicategorizer [OPTION]... -f [FILTER] [FILE]...\n"
            "Usage: biobloommicategorizer [OPTION]... -f [FILTER1] -e [FILE1.fq] [FILE2.fq]\n"
            "The input format may be FASTA, FASTQ, and compressed with gz. Output will be to\n"
            "stdout with a summary file outputted to a prefix.\n"
            "  -p, --prefix=N         Output prefix to use. Otherwise will output to current\n"
            "                         directory.\n"
            "  -f, --filter=N         Path of miBF '.bf' file\n"
            "  -e, --paired_mode      Uses paired-end information.\n"
            "  -s, --min_FPR=N        Minimum -10*log(FPR) threshold for a match. [100.0]\n"
            "  -t, --threads=N        Threads that parse each input file. [all; at most 8\n"
            "                         for plain files, 16 for gzip]\n"
            "  -i, --hitOnly          Return only results that hit the filter.\n"
            "      --fa               Output categorized reads in Fasta output.\n"
            "      --fq               Output categorized reads in Fastq output.\n"
            "      --version          Display version information.\n"
            "  -h, --help             Display this dialog.\n"
            "  -v, --verbose          Display verbose output\n"
            "  -I, --interval         the interval to report file processing status [10000000]\n"
            "      --batch_reads=N    Reads (or pairs) per device batch. [131072]\n"
            "      --ascii_reads      Send reads to the device as text (default: 2-bit\n"
            "                         codes and an N mask, 3/8 of the bytes).\n"
            "Advanced options:\n"
            "  -a, --frameMatches=N   Min number seed matches needed in a frame to match [1]\n"
            "                         Ignored if k-mers used when indexing.\n"
            "  -m, --multi=N          Multi Match threshold. [1.0]\n"
            "  -r, --streak=N         Number of additional hits needed to skip classification. [3]\n"
            "                         Accepted and checked against -c; it has no other effect\n"
            "                         here: every frame of a read is evaluated.\n"
            "  -c, --minNoSat=N       Minimum count of non saturated matches. Increasing this value\n"
            "                         will filter out repetitive or low quality sequences.[0]\n"
            "  -b, --bestHitAgree     Filters out all matches where best hit is ambiguous because\n"
            "                         match count metrics do not agree.\n";
    exit(EXIT_SUCCESS);
}

bool fexists(const string &f) {
    ifstream in(f.c_str());
    return in.good();
}

/* One of two pipeline slots: pinned staging and device buffers for a batch
 * of reads, and pinned host copies of its results. */
struct Slot {
    char *dseq = nullptr, *hseq = nullptr;
    uint64_t *doff = nullptr, *hoff = nullptr;
    uint32_t *dnh = nullptr, *dhits = nullptr, *dsat = nullptr;
    uint32_t *hnh = nullptr, *hhits = nullptr, *hsat = nullptr;
    uint64_t *dctr = nullptr, *hctr = nullptr;
    size_t seq_cap = 0, rec_cap = 0, hit_cap = 0;
    size_t nreads = 0, bytes = 0;
    size_t code_bytes = 0; /* 2-bit batch: bytes of codes (the mask follows); 0: ASCII */
    uint32_t stride = 0; /* candidates per record in dhits / hhits */
    uint32_t max_len = 0;

    static void die(const string &m) {
        cerr << "Error: " << m << endl;
        exit(1);
    }
    void release() {
        (void)hipFree(dseq);
        (void)hipFree(doff);
        (void)hipFree(dnh);
        (void)hipFree(dhits);
        (void)hipFree(dsat);
        (void)hipFree(dctr);
        (void)hipHostFree(hseq);
        (void)hipHostFree(hoff);
        (void)hipHostFree(hnh);
        (void)hipHostFree(hhits);
        (void)hipHostFree(hsat);
        (void)hipHostFree(hctr);
        dseq = hseq = nullptr;
        doff = hoff = nullptr;
        dnh = dhits = dsat = hnh = hhits = hsat = nullptr;
        dctr = hctr = nullptr;
        seq_cap = rec_cap = hit_cap = 0;
    }
    /* bytes: the padded sequence bytes of the batch; n reads */
    void reserve(size_t nbytes, size_t n) {
        if (nbytes > seq_cap) {
            (void)hipFree(dseq);
            (void)hipHostFree(hseq);
            seq_cap = std::max(nbytes, seq_cap * 2);
            if (hipMalloc(&dseq, seq_cap) != hipSuccess ||
                hipHostMalloc(&hseq, seq_cap, hipHostMallocDefault) != hipSuccess)
                die("device allocation failed");
        }
        if (n > rec_cap) {
            (void)hipFree(doff);
            (void)hipFree(dnh);
            (void)hipFree(dsat);
            (void)hipHostFree(hoff);
            (void)hipHostFree(hnh);
            (void)hipHostFree(hsat);
            rec_cap = std::max(n, rec_cap * 2);
            if (hipMalloc(&doff, (rec_cap + 1) * 8) != hipSuccess ||
                hipMalloc(&dnh, rec_cap * 4) != hipSuccess ||
                hipMalloc(&dsat, rec_cap * 4) != hipSuccess ||
                hipHostMalloc(&hoff, (rec_cap + 1) * 8, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(&hnh, rec_cap * 4, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(&hsat, rec_cap * 4, hipHostMallocDefault) != hipSuccess)
                die("device allocation failed");
        }
    }
    /* room for max_hits candidates of each of n records */
    void reserve_hits(size_t n, uint32_t max_hits) {
        const size_t words = n * max_hits * 7;
        if (words > hit_cap) {
            (void)hipFree(dhits);
            (void)hipHostFree(hhits);
            hit_cap = std::max(words, hit_cap * 2);
            if (hipMalloc(&dhits, hit_cap * 4) != hipSuccess ||
                hipHostMalloc(&hhits, hit_cap * 4, hipHostMallocDefault) != hipSuccess)
                die("device allocation failed");
        }
        stride = max_hits;
        if (!dctr && (hipMalloc(&dctr, 32) != hipSuccess ||
                      hipHostMalloc(&hctr, 32, hipHostMallocDefault) != hipSuccess))
            die("device allocation failed");
    }
};

/* blocks of parsed input one batch may hold (the three batches alive at
 * once stay well inside the reader's pool) */
constexpr size_t kBatchBlocks = (bbt_fe::ParallelSeqReader::kMaxBlocks - 4) / 3;

class MIBFClassifier {
  public:
    MIBFClassifier(const Opt &o) : m_o(o) {
        /* MIBFClassifier.hpp:49-95 */
        if (bbt_mibf_load(o.filter.c_str(), &m_mf, nullptr) != BBT_OK) {
            cerr << "Error: " << bbt_last_error() << endl;
            exit(1);
        }
        bbt_mibf_info(m_mf, &m_info);
        for (uint32_t i = 0; i <= m_info.nids; ++i)
            m_names.push_back(bbt_mibf_name(m_mf, i));
        m_allowedMiss = (o.frameMatches > m_info.nseeds || m_info.kmer_mode)
                            ? 0
                            : m_info.nseeds - o.frameMatches;
        m_above.assign(m_info.nids + 1, 0);
        m_unique.assign(m_info.nids + 1, 0);
        /* one stream: the filter's query state is shared by its launches */
        if (hipStreamCreateWithFlags(&m_st, hipStreamNonBlocking) != hipSuccess)
            Slot::die("stream creation failed");
        for (hipEvent_t &e : m_done)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                Slot::die("event creation failed");
    }
    ~MIBFClassifier() {
        (void)hipStreamSynchronize(m_st);
        for (Slot &s : m_slot)
            s.release();
        m_redo.release();
        for (hipEvent_t &e : m_done)
            (void)hipEventDestroy(e);
        (void)hipStreamDestroy(m_st);
        bbt_mibf_free(m_mf);
    }

    void filter(const vector<string> &files) {
        cerr << "Filtering Start" << endl;
        m_t0 = std::chrono::steady_clock::now();
        for (const string &f : files) {
            bbt_fe::ParallelSeqReader rd(f, m_o.threads);
            if (!rd.ok()) {
                cerr << "file " << f << " cannot be opened" << endl;
                exit(1);
            }
            bbt_fe::SeqCursor cur(rd);
            pipeline(
                [&](bbt_fe::RecBatch &b) {
                    while (b.v.size() < m_o.batch && b.holds.size() < kBatchBlocks) {
                        const bbt_fe::SeqView *v = cur.next();
                        if (!v)
                            return false;
                        b.v.push_back(*v);
                        b.hold(cur.block());
                        if (++m_numRead % m_o.interval == 0)
                            cerr << "Currently Reading Read Number: " << m_numRead << endl;
                    }
                    return true;
                },
                false);
        }
        finish();
    }

    void filterPair(const string &f1, const string &f2) {
        m_t0 = std::chrono::steady_clock::now();
        bbt_fe::ParallelSeqReader r1(f1, m_o.threads), r2(f2, m_o.threads);
        if (!r1.ok() || !r2.ok()) {
            cerr << "file " << (r1.ok() ? f2 : f1) << " cannot be opened" << endl;
            exit(1);
        }
        bbt_fe::SeqCursor c1(r1), c2(r2);
        pipeline(
            [&](bbt_fe::RecBatch &b) {
                while (b.v.size() < 2 * m_o.batch && b.holds.size() < kBatchBlocks) {
                    const bbt_fe::SeqView *a = c1.next();
                    if (!a)
                        return false;
                    const bbt_fe::SeqView *m = c2.next();
                    if (!m)
                        return false;
                    b.v.push_back(*a);
                    b.hold(c1.block());
                    b.v.push_back(*m);
                    b.hold(c2.block());
                    if (++m_numRead % m_o.interval == 0)
                        cerr << "Currently Reading Read Number: " << m_numRead << endl;
                }
                return true;
            },
            true);
        finish();
    }

  private:
    /* Pipeline over batches: while the device queries batch i, one helper
     * thread packs batch i+1 and queues it behind batch i on the stream,
     * another waits for batch i and formats it, and this thread gathers
     * batch i+2; outputs keep input order. */
    void pipeline(const std::function<bool(bbt_fe::RecBatch &)> &next, bool paired) {
        using clk = std::chrono::steady_clock;
        const auto since = [](clk::time_point t) {
            return std::chrono::duration<double>(clk::now() - t).count();
        };
        bbt_fe::RecBatch recs[3];
        bool more = true;
        const auto gather = [&](int k) {
            const clk::time_point t = clk::now();
            recs[k].clear();
            if (more)
                more = next(recs[k]);
            m_phase[0] += since(t);
        };
        gather(0);
        clk::time_point t = clk::now();
        pack(m_slot[0], recs[0].v);
        launch(0, paired);
        m_phase[1] += since(t);
        gather(1);
        for (uint64_t i = 0;; ++i) {
            /* slot (i + 1) % 2 and recs[(i + 2) % 3] held batch i - 1, which
             * has been emitted */
            const int b0 = int(i % 3), b1 = int((i + 1) % 3), b2 = int((i + 2) % 3);
            const int s0 = int(i % 2), s1 = int((i + 1) % 2);
            std::thread packer([&] {
                const clk::time_point tp = clk::now();
                pack(m_slot[s1], recs[b1].v);
                launch(s1, paired);
                m_phase[1] += since(tp);
            });
            std::thread emitter([&] {
                clk::time_point te = clk::now();
                if (m_slot[s0].nreads && hipEventSynchronize(m_done[s0]) != hipSuccess)
                    Slot::die("query failed");
                m_phase[2] += since(te);
                te = clk::now();
                emit(s0, recs[b0], paired);
                m_phase[3] += since(te);
            });
            gather(b2);
            packer.join();
            emitter.join();
            recs[b0].clear();
            if (recs[b1].v.empty())
                break;
        }
    }

    void pack(Slot &S, const vector<bbt_fe::SeqView> &v) {
        const size_t n = v.size();
        S.nreads = n;
        if (!n)
            return;
        size_t bytes = 0;
        uint32_t max_len = 0;
        for (const auto &r : v) {
            bytes += r.seq_len;
            max_len = std::max(max_len, r.seq_len);
        }
        const size_t padded = m_o.pack2bit ? bbt_host::packed_bytes(bytes)
                                           : (bytes + 7) / 8 * 8 + 8;
        S.reserve(padded, n);
        uint64_t o = 0;
        for (size_t i = 0; i < n; ++i) {
            S.hoff[i] = o;
            o += v[i].seq_len;
        }
        S.hoff[n] = o;
        S.max_len = max_len;
        if (m_o.pack2bit) {
            /* 2-bit codes and the N mask straight from the records' views */
            const auto get = [&](uint64_t r) { return v[r].seq; };
            S.code_bytes = bbt_host::pack_2bit(get, S.hoff, n,
                                               reinterpret_cast<uint8_t *>(S.hseq), 8);
            S.bytes = padded;
            return;
        }
        S.code_bytes = 0;
        const unsigned T = o < (size_t(8) << 20) ? 1u : 8u;
        const auto gather = [&](unsigned t) {
            for (size_t i = n * t / T, e = n * (t + 1) / T; i < e; ++i)
                memcpy(S.hseq + S.hoff[i], v[i].seq, v[i].seq_len);
        };
        vector<std::thread> th;
        for (unsigned t = 1; t < T; ++t)
            th.emplace_back(gather, t);
        gather(0);
        for (auto &t : th)
            t.join();
        memset(S.hseq + o, 0, padded - o);
        S.bytes = padded;
        S.max_len = max_len;
    }

    /* queues slot si's batch and an event after its copy-out */
    void launch(int si, bool paired) {
        if (submit(m_slot[si], paired, kFirstHits) &&
            hipEventRecord(m_done[si], m_st) != hipSuccess)
            Slot::die("event record failed");
    }

    /* queues copy-in, query (keeping max_hits candidates per record) and
     * copy-out of S's batch; false when the batch is empty */
    bool submit(Slot &S, bool paired, uint32_t max_hits) {
        const size_t n = S.nreads, nrec = paired ? n / 2 : n;
        if (!n)
            return false;
        S.reserve_hits(nrec, max_hits);
        memset(S.hctr, 0, 32);
        if (hipMemcpyAsync(S.dseq, S.hseq, S.bytes, hipMemcpyHostToDevice, m_st) != hipSuccess ||
            hipMemcpyAsync(S.doff, S.hoff, (n + 1) * 8, hipMemcpyHostToDevice, m_st) !=
                hipSuccess ||
            hipMemcpyAsync(S.dctr, S.hctr, 32, hipMemcpyHostToDevice, m_st) != hipSuccess)
            Slot::die("upload failed");
        bbt_reads in;
        std::memset(&in, 0, sizeof(in));
        in.seq = S.dseq;
        in.seq_bytes = S.code_bytes ? S.code_bytes : S.bytes;
        in.format = S.code_bytes ? BBT_READS_2BIT : BBT_READS_ASCII;
        in.off = S.doff;
        in.nreads = n;
        in.max_len = S.max_len;
        bbt_mibf_params p;
        std::memset(&p, 0, sizeof(p));
        p.score = m_o.score;
        p.multi_thresh = m_o.multi;
        p.allowed_miss = m_allowedMiss;
        p.min_nonsat = m_o.minNoSat;
        p.max_hits = max_hits;
        p.flags = (paired ? BBT_MIBF_PAIRED : 0) | (m_o.bestAgree ? BBT_MIBF_BEST_AGREE : 0);
        bbt_mibf_out out{S.dnh, S.dhits, S.dsat, S.dctr};
        if (bbt_mibf_query_batch(m_mf, &in, &p, &out, m_st) != BBT_OK) {
            cerr << "Error: " << bbt_last_error() << endl;
            exit(1);
        }
        if (hipMemcpyAsync(S.hnh, S.dnh, nrec * 4, hipMemcpyDeviceToHost, m_st) != hipSuccess ||
            hipMemcpyAsync(S.hhits, S.dhits, nrec * max_hits * 7 * 4, hipMemcpyDeviceToHost,
                           m_st) != hipSuccess ||
            hipMemcpyAsync(S.hsat, S.dsat, nrec * 4, hipMemcpyDeviceToHost, m_st) != hipSuccess ||
            hipMemcpyAsync(S.hctr, S.dctr, 32, hipMemcpyDeviceToHost, m_st) != hipSuccess)
            Slot::die("download failed");
        return true;
    }

    /* the output lines and summary counts of a finished batch: records in
     * ranges formatted on threads, then written in input order */
    void emit(int si, const bbt_fe::RecBatch &b, bool paired) {
        Slot &S = m_slot[si];
        const size_t n = S.nreads, nrec = paired ? n / 2 : n;
        if (!n)
            return;
        /* records with more candidates than the batch's query kept are
         * queried again, as one small batch, with room for all of them */
        vector<size_t> big;
        uint32_t need = 0;
        for (size_t r = 0; r < nrec; ++r) {
            if (S.hnh[r] == 0xFFFFFFFFu)
                Slot::die("a record was not evaluated (per-record store overflow)");
            if (S.hnh[r] > S.stride) {
                big.push_back(r);
                need = std::max(need, S.hnh[r]);
            }
        }
        vector<int64_t> redo_of;
        if (!big.empty()) {
            vector<bbt_fe::SeqView> rv;
            for (size_t r : big) {
                if (paired) {
                    rv.push_back(b.v[2 * r]);
                    rv.push_back(b.v[2 * r + 1]);
                } else {
                    rv.push_back(b.v[r]);
                }
            }
            pack(m_redo, rv);
            submit(m_redo, paired, need);
            if (hipStreamSynchronize(m_st) != hipSuccess)
                Slot::die("query failed");
            redo_of.assign(nrec, -1);
            for (size_t j = 0; j < big.size(); ++j) {
                if (m_redo.hnh[j] != S.hnh[big[j]])
                    Slot::die("the repeated query of a record differs");
                redo_of[big[j]] = int64_t(j);
            }
        }
        const unsigned T = nrec < 4096 ? 1u : 8u;
        struct Part {
            string out;
            vector<uint64_t> above, unique;
            uint64_t multi = 0, noMatch = 0;
        };
        vector<Part> parts(T);
        const auto work = [&](unsigned t) {
            Part &P = parts[t];
            P.above.assign(m_info.nids + 1, 0);
            P.unique.assign(m_info.nids + 1, 0);
            P.out.reserve((nrec * (t + 1) / T - nrec * t / T) * (paired ? 2 : 1) * 96);
            for (size_t r = nrec * t / T, e = nrec * (t + 1) / T; r < e; ++r) {
                const uint32_t nh = S.hnh[r];
                const uint32_t *h =
                    nh > S.stride ? m_redo.hhits + size_t(redo_of[r]) * m_redo.stride * 7
                                  : S.hhits + r * S.stride * 7;
                if (!paired) {
                    /* filterSingleRead (MIBFClassifier.hpp:738-751) */
                    count(P, h, nh);
                    if (m_o.hitOnly && nh == 0)
                        continue;
                    format(b.v[r], S.hsat[r], h, nh, P.out);
                } else {
                    /* filterPairedRead (:753-771): hitOnly returns before the
                     * summary update */
                    if (m_o.hitOnly && nh == 0)
                        continue;
                    count(P, h, nh);
                    format(b.v[2 * r], S.hsat[r], h, nh, P.out);
                    format(b.v[2 * r + 1], S.hsat[r], h, nh, P.out);
                }
            }
        };
        vector<std::thread> th;
        for (unsigned t = 1; t < T; ++t)
            th.emplace_back(work, t);
        work(0);
        for (auto &t : th)
            t.join();
        vector<string> text;
        for (Part &P : parts) {
            text.push_back(std::move(P.out));
            for (uint32_t i = 0; i <= m_info.nids; ++i) {
                m_above[i] += P.above[i];
                m_unique[i] += P.unique[i];
            }
            m_multi += P.multi;
            m_noMatch += P.noMatch;
        }
        m_writer.push(std::move(text));
    }

    /* Writes formatted batches to stdout in order on its own thread, so a
     * batch is written while the next one is formatted; at most one batch
     * waits. */
    struct Writer {
        std::thread th;
        std::mutex mu;
        std::condition_variable cv;
        std::deque<vector<string>> q;
        bool busy = false, stop = false;
        Writer() {
            th = std::thread([this] {
                while (true) {
                    vector<string> t;
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
                    for (const string &x : t)
                        cout.write(x.data(), std::streamsize(x.size()));
                    {
                        std::lock_guard<std::mutex> g(mu);
                        busy = false;
                    }
                    cv.notify_all();
                }
            });
        }
        ~Writer() {
            {
                std::lock_guard<std::mutex> g(mu);
                stop = true;
            }
            cv.notify_all();
            th.join();
        }
        void push(vector<string> &&t) {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] { return q.empty(); });
            q.push_back(std::move(t));
            cv.notify_all();
        }
        void drain() {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] { return q.empty() && !busy; });
        }
    };

    /* ResultsManager::updateSummaryData over QueryResults (ResultsManager.hpp:40-64) */
    template <class P>
    static void count(P &p, const uint32_t *h, uint32_t n) {
        for (uint32_t i = 0; i < n; ++i)
            ++p.above[h[i * 7]];
        if (n == 0)
            ++p.noMatch;
        else if (n == 1)
            ++p.unique[h[0]];
        else
            ++p.multi;
    }

    /* decimal digits of v appended to o */
    static void put_u32(string &o, uint32_t v) {
        char b[10];
        int n = 0;
        do {
            b[n++] = char('0' + v % 10);
            v /= 10;
        } while (v);
        while (n)
            o.push_back(b[--n]);
    }

    /* appendResults (MIBFClassifier.hpp:652-673): name, nonSatFrameCount,
     * solidCount, count, nonSatCount, totalNonSatCount, totalCount */
    void append(const uint32_t *h, uint32_t n, string &o) const {
        static const int kField[6] = {6, 5, 1, 2, 4, 3};
        o.push_back('\t');
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t *c = h + i * 7;
            if (i)
                o.push_back(';');
            o += m_names[c[0]];
            for (int f : kField) {
                o.push_back(',');
                put_u32(o, c[f]);
            }
        }
    }

    /* formatOutStr (MIBFClassifier.hpp:675-736) */
    void format(const bbt_fe::SeqView &rd, uint32_t sat, const uint32_t *h, uint32_t n,
                string &o) const {
        const auto head = [&] {
            o.append(rd.name, rd.name_len);
            if (rd.comment_len) {
                o.push_back(' ');
                o.append(rd.comment, rd.comment_len);
            }
        };
        if (m_o.out == FASTA || m_o.out == FASTQ) {
            o.push_back(m_o.out == FASTA ? '>' : '@');
            head();
            if (n) {
                o.push_back('\t');
                o += m_names[h[0]];
                append(h, n, o);
            }
            o.push_back('\n');
            o.append(rd.seq, rd.seq_len);
            if (m_o.out == FASTQ) {
                o += "\n+\n";
                o.append(rd.qual, rd.qual_len);
            }
        } else if (n == 0) {
            o += "*\t";
            head();
            o.push_back('\t');
            put_u32(o, sat);
            o += "\t*";
        } else {
            o += m_names[h[0]];
            o.push_back('\t');
            head();
            o.push_back('\t');
            put_u32(o, sat);
            append(h, n, o);
        }
        o.push_back('\n');
    }

    void finish() {
        m_writer.drain();
        cout.flush();
        cerr << "Filtering time: "
             << std::chrono::duration<double>(std::chrono::steady_clock::now() - m_t0).count()
             << " s (gather " << m_phase[0] << ", pack+launch " << m_phase[1]
             << ", device wait " << m_phase[2] << ", output " << m_phase[3] << ")" << endl;
        cerr << "Total Reads:" << m_numRead << endl;
        cerr << "Writing file: " << m_o.prefix << "_summary.tsv" << endl;
        vector<uint64_t> sm(2 * (m_info.nids + 1) + 2);
        for (uint32_t i = 0; i <= m_info.nids; ++i) {
            sm[i] = m_above[i];
            sm[m_info.nids + 1 + i] = m_unique[i];
        }
        sm[2 * (m_info.nids + 1)] = m_multi;
        sm[2 * (m_info.nids + 1) + 1] = m_noMatch;
        uint64_t need = 0;
        bbt_mibf_summary_text(m_mf, sm.data(), m_numRead, nullptr, 0, &need);
        string text(need + 1, '\0');
        bbt_mibf_summary_text(m_mf, sm.data(), m_numRead, &text[0], need + 1, nullptr);
        text.resize(need);
        ofstream so(m_o.prefix + "_summary.tsv");
        so << text;
        cout.flush();
    }

    Opt m_o;
    bbt_mibf *m_mf = nullptr;
    bbt_mibf_info_t m_info{};
    vector<string> m_names;
    uint32_t m_allowedMiss = 0;
    hipStream_t m_st = nullptr;
    hipEvent_t m_done[2] = {nullptr, nullptr}; /* recorded after each slot's copy-out */
    double m_phase[4] = {0, 0, 0, 0}; /* s: gather, pack + launch, device wait, output */
    Slot m_slot[2];
    Slot m_redo; /* records queried again with room for all their candidates */
    Writer m_writer;
    std::chrono::steady_clock::time_point m_t0;
    uint64_t m_numRead = 0, m_multi = 0, m_noMatch = 0;
    vector<uint64_t> m_above, m_unique;
};

} // namespace

int main(int argc, char *argv[]) {
    Opt o;
    int fq = 0, fa = 0, tsv = 0, version = 0, debug = 0;
    bool die = false, streak_given = false;
    static struct option lo[] = {{"prefix", required_argument, NULL, 'p'},
                                 {"filter", required_argument, NULL, 'f'},
                                 {"paired", no_argument, NULL, 'e'},
                                 {"min_FPR", required_argument, NULL, 's'},
                                 {"help", no_argument, NULL, 'h'},
                                 {"interval", required_argument, NULL, 'I'},
                                 {"threads", required_argument, NULL, 't'},
                                 {"frameMatches", required_argument, NULL, 'a'},
                                 {"fq", no_argument, &fq, 1},
                                 {"fa", no_argument, &fa, 1},
                                 {"tsv", no_argument, &tsv, 1},
                                 {"hitOnly", no_argument, NULL, 'i'},
                                 {"version", no_argument, &version, 1},
                                 {"multi", required_argument, NULL, 'm'},
                                 {"streak", required_argument, NULL, 'r'},
                                 {"minNoSat", required_argument, NULL, 'c'},
                                 {"bestHitAgree", no_argument, NULL, 'b'},
                                 {"stdout", no_argument, NULL, 'd'},
                                 {"inverse", no_argument, NULL, 'n'},
                                 {"debug", no_argument, &debug, 1},
                                 {"verbose", no_argument, NULL, 'v'},
                                 {"batch_reads", required_argument, NULL, 1000},
                                 {"ascii_reads", no_argument, NULL, 1001},
                                 {NULL, 0, NULL, 0}};
    int c, idx = 0;
    while ((c = getopt_long(argc, argv, "p:f:es:hI:t:f:m:r:dnvc:a:bi", lo, &idx)) != -1) {
        stringstream cv(optarg ? optarg : "");
        switch (c) {
        case 's': {
            double raw = 0;
            cv >> raw;
            o.score = std::pow(10.0, -(raw / 10.0));
            break;
        }
        case 'f': o.filter = optarg; break;
        case 'p': o.prefix = optarg; break;
        case 1000:
            if (!(cv >> o.batch) || o.batch == 0) {
                cerr << "Error - Invalid parameters! batch_reads: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        case 1001: o.pack2bit = false; break;
        case 'h': usage(); break;
        case 'I':
            if (!(cv >> o.interval)) {
                cerr << "Error - Invalid parameters! I: " << optarg << endl;
                return 0;
            }
            break;
        case 'e': o.paired = true; break;
        case 'a':
            if (!(cv >> o.frameMatches)) {
                cerr << "Error - Invalid parameter! a: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        case 'b': o.bestAgree = true; break;
        case 't':
            if (!(cv >> o.threads)) {
                cerr << "Error - Invalid parameter! t: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        case 'v': ++o.verbose; break;
        case 'r':
            if (!(cv >> o.streak)) {
                cerr << "Error - Invalid parameter! r: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            streak_given = true;
            break;
        case 'c':
            if (!(cv >> o.minNoSat)) {
                cerr << "Error - Invalid parameter! c: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        case 'm':
            if (!(cv >> o.multi)) {
                cerr << "Error - Invalid parameter! m: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        case 'd':
        case 'n': break; /* parsed, unused, as in the reference */
        case 'i': o.hitOnly = true; break;
        case 0: break;
        case '?': die = true; break;
        default: die = true;
        }
    }
    if (version) {
        cerr << "biobloommicategorizer (biobloom_amd)" << endl;
        return 0;
    }
    vector<string> inputs;
    while (optind < argc)
        inputs.push_back(argv[optind++]);
    if (o.paired && inputs.size() != 2) {
        cerr << "Usage of paired end mode:\n"
             << "BioBloomCategorizer [OPTION]... -f \"[FILTER1]...\" [FILEPAIR1] [FILEPAIR2]\n"
             << endl;
        exit(1);
    }
    if (o.streak < o.minNoSat) {
        cerr << "-r should be greater or equalt to -c" << endl;
        exit(1);
    }
    if (streak_given)
        cerr << "Warning: -r (streak) does not stop a query early here; every frame of "
                "a read is evaluated (the reference's streak rule is not available)"
             << endl;
    if (inputs.empty()) {
        cerr << "Error: Need Input File" << endl;
        die = true;
    }
    if (o.filter.empty()) {
        cerr << "Error: Need Filter File (-f)" << endl;
        die = true;
    }
    if (die) {
        cerr << "Try '--help' for more information.\n";
        exit(EXIT_FAILURE);
    }
    if (o.prefix.find('/') != string::npos) {
        const string dir = o.prefix.substr(0, o.prefix.find_last_of('/'));
        struct stat sb;
        if (stat(dir.c_str(), &sb) != 0) {
            cerr << "Error: Output folder does not exist. " << dir << endl;
            exit(1);
        } else if (!S_ISDIR(sb.st_mode)) {
            cerr << "Error: Output folder - file exists with this name. " << dir << endl;
            exit(1);
        }
    }
    if (fa && fq) {
        cerr << "Error: fasta (--fa) and fastq (--fq) outputs types cannot be both set" << endl;
        exit(1);
    }
    o.out = fq ? FASTQ : (fa ? FASTA : TSV);
    (void)tsv;
    if (debug) {
        cerr << "Error: --debug output is not available in this build" << endl;
        exit(1);
    }
    if (!fexists(o.filter)) {
        cerr << "Error: " + o.filter + " File cannot be opened" << endl;
        exit(1);
    }
    MIBFClassifier BMC(o);
    if (o.paired)
        BMC.filterPair(inputs[0], inputs[1]);
    else
        BMC.filter(inputs);
    return 0;
}
