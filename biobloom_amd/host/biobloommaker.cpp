/*
 * biobloommaker -- the reference maker (BioBloomMaker/BioBloomMaker.cpp:
 * 110-547) over libbbt_hip: the same sizing (BloomFilterInfo.h:57-76),
 * k-mer insertion on the GPU, the .bf/.txt outputs, -s (a filter whose
 * k-mers are left out, BloomFilterGenerator.cpp:1175-1199) and the
 * progressive builders (-r, BloomFilterGenerator.cpp:85-1170).
 *
 * Progressive growth is sequential by definition: a read pair is tested
 * against the filter as the pairs before it left it.  The reference runs the
 * loop under OpenMP, so its result depends on thread timing; this builder
 * gives the one-thread result.  It tests a window of pending pairs in one
 * device launch against the current filter; the pairs before the first
 * tagged pair are untagged in the sequential order too (the filter does not
 * change before a tag), the tagged pair is inserted, and the pairs after it
 * are tested again against the grown filter.
 */
#include <getopt.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <deque>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "front_end.h"

using namespace std;

static void usage() {
    cerr << "Usage: biobloommaker -p [FILTERID] [OPTION]... [FILE]...\n"
            "Usage: biobloommaker -p [FILTERID] -r 0.2 [OPTION]... [FILE]... [FASTQ1] [FASTQ2]\n"
            "  -p, --file_prefix=N    Filter prefix and filter ID. Required option.\n"
            "  -o, --output_dir=N     Output location of the filter and filter info files.\n"
            "  -f, --fal_pos_rate=N   Maximum false positive rate to use in filter. [0.0078125]\n"
            "  -g, --hash_num=N       Set number of hash functions to use in filter.\n"
            "  -k, --kmer_size=N      K-mer size to use to create filter. [25]\n"
            "  -n, --num_ele=N        Set the number of expected elements.\n"
            "  -s, --subtract=N       Path to a filter whose k-mers are not added (with -r:\n"
            "                         k-mers it holds do not score).\n"
            "  -r, --progressive=N    Progressive filter creation; N is a score in [0, 1] or a\n"
            "                         minimum match length (a positive integer > 1).\n"
            "  -d, --no_rep_kmer      Progressive mode: load the seed with the -s filter\n"
            "                         subtracted.\n"
            "  -b, --baitScore=N      Score threshold when considering only bait. [r]\n"
            "  -e, --iterations=N     Pass through files N times if threshold is not met.\n"
            "  -i, --inclusive        If one paired read matches, both reads will be included.\n"
            "  -l, --file_list=N      A file of list of file pairs (or single files).\n"
            "  -I, --interval         The interval to report file processing status\n"
            "  -P, --print_reads      Print tagged reads to STDOUT in FASTQ format.\n"
            "  -t, --threads=N        Accepted for compatibility.\n"
            "  -h, --help             Display this dialog.\n";
    exit(EXIT_SUCCESS);
}

namespace {

void fatal(const string &m) {
    cerr << m << endl;
    exit(1);
}

void check(bbt_status s) {
    if (s != BBT_OK)
        fatal(bbt_last_error());
}

#define HIP_OR_DIE(expr)                                                        \
    do {                                                                        \
        if ((expr) != hipSuccess)                                               \
            fatal(string("HIP error: ") + #expr);                               \
    } while (0)

/* Device staging of one window of reads. */
struct Window {
    hipStream_t st = nullptr;
    size_t cap_reads = 0, cap_bytes = 0;
    char *d_seq = nullptr, *h_seq = nullptr;
    uint64_t *d_off = nullptr, *h_off = nullptr;
    uint8_t *d_pass = nullptr, *h_pass = nullptr; /* filter, then bait */
    Window() { HIP_OR_DIE(hipStreamCreate(&st)); }
    ~Window() {
        release();
        (void)hipStreamDestroy(st);
    }
    void release() {
        (void)hipFree(d_seq);
        (void)hipFree(d_off);
        (void)hipFree(d_pass);
        (void)hipHostFree(h_seq);
        (void)hipHostFree(h_off);
        (void)hipHostFree(h_pass);
        d_seq = h_seq = nullptr;
        d_off = h_off = nullptr;
        d_pass = h_pass = nullptr;
    }
    void reserve(size_t reads, size_t bytes) {
        bytes = (bytes + 7) / 8 * 8 + 8;
        if (reads <= cap_reads && bytes <= cap_bytes)
            return;
        release();
        cap_reads = max(reads, cap_reads * 2);
        cap_bytes = max(bytes, cap_bytes * 2);
        HIP_OR_DIE(hipMalloc(&d_seq, cap_bytes));
        HIP_OR_DIE(hipMalloc(&d_off, (cap_reads + 1) * 8));
        HIP_OR_DIE(hipMalloc(&d_pass, 2 * cap_reads));
        HIP_OR_DIE(hipHostMalloc(&h_seq, cap_bytes, hipHostMallocDefault));
        HIP_OR_DIE(hipHostMalloc(&h_off, (cap_reads + 1) * 8, hipHostMallocDefault));
        HIP_OR_DIE(hipHostMalloc(&h_pass, 2 * cap_reads, hipHostMallocDefault));
    }
};

/* BloomFilterGenerator's progressive loops (BloomFilterGenerator.cpp:117-274
 * two files, 343-509 with a bait filter, 821-1034 file pairs, 1066-1160
 * single-end files), in the one-thread order. */
class Progressive {
  public:
    bbt_filterset *fs = nullptr, *bait = nullptr, *sub = nullptr;
    double score = 0, bait_score = 0;
    uint32_t scorer = BBT_SIMPLE, k = 25, h = 3;
    bool inclusive = false, print = false;
    uint64_t expected = 0, total = 0, total_reads = 0, tagged = 0;
    uint64_t interval = 10000000;

    ~Progressive() {
        if (scratch)
            bbt_filters_free(scratch);
    }

    /* one pass over a pair of read files */
    void pairs(const string &f1, const string &f2, const char *tag) {
        bbt_fe::ParallelSeqReader r1(f1), r2(f2);
        if (!r1.ok())
            fatal("file " + f1 + " cannot be opened");
        cerr << tag << f1 << endl;
        if (!r2.ok())
            fatal("file " + f2 + " cannot be opened");
        cerr << tag << f2 << endl;
        deque<bbt_fe::SeqRecord> q; /* pending reads, mates adjacent */
        bool eof = false;
        size_t want = 1024;
        while (true) {
            while (!eof && q.size() < 2 * want) {
                bbt_fe::SeqRecord a, b;
                if (r1.read(a) < 0 || r2.read(b) < 0) {
                    eof = true;
                    break;
                }
                q.push_back(std::move(a));
                q.push_back(std::move(b));
            }
            if (q.empty())
                return;
            /* the next pair is read and counted, then the loop stops once
             * the filter holds the expected k-mers */
            if (total >= expected) {
                count(1);
                return;
            }
            evaluate(q);
            const size_t n = q.size() / 2;
            size_t t = 0;
            for (; t < n; ++t)
                if (pair_tagged(q[2 * t], q[2 * t + 1], t))
                    break;
            if (t == n) {
                count(n);
                q.clear();
                want = min<size_t>(want * 2, 1 << 16);
                continue;
            }
            count(t + 1);
            ++tagged;
            const uint64_t n1 = insert(2 * t), n2 = insert(2 * t + 1);
            if (print) {
                const uint64_t rep1 = repeats(2 * t), rep2 = repeats(2 * t + 1);
                debug(q[2 * t], n1, rep1);
                debug(q[2 * t + 1], n2, rep2);
            }
            q.erase(q.begin(), q.begin() + 2 * (t + 1));
            want = min<size_t>(max<size_t>(4 * (t + 1), 256), 1 << 16);
        }
    }

    /* one pass over a single-end read file */
    void singles(const string &f) {
        bbt_fe::ParallelSeqReader r(f);
        if (!r.ok())
            fatal("file " + f + " cannot be opened");
        cerr << "Reading File: " << f << endl;
        deque<bbt_fe::SeqRecord> q;
        bool eof = false;
        size_t want = 2048;
        while (true) {
            while (!eof && q.size() < want) {
                bbt_fe::SeqRecord a;
                if (r.read(a) < 0) {
                    eof = true;
                    break;
                }
                q.push_back(std::move(a));
            }
            /* single-end: a read is counted only while the filter is short */
            if (q.empty() || total >= expected)
                return;
            evaluate(q);
            const size_t n = q.size();
            size_t t = 0;
            for (; t < n; ++t)
                if (kmers(q[t]) > score && w.h_pass[t])
                    break;
            if (t == n) {
                count(n);
                q.clear();
                want = min<size_t>(want * 2, 1 << 17);
                continue;
            }
            count(t + 1);
            ++tagged;
            const uint64_t n1 = insert(t);
            if (print)
                debug(q[t], n1, repeats(t));
            q.erase(q.begin(), q.begin() + t + 1);
            want = min<size_t>(max<size_t>(4 * (t + 1), 256), 1 << 17);
        }
    }

  private:
    Window w;
    size_t nreads = 0;
    bbt_filterset *scratch = nullptr;

    double kmers(const bbt_fe::SeqRecord &r) const {
        return r.seq.size() > k ? double(r.seq.size() - k + 1) : 0.0;
    }

    /* whether pending pair t is tagged (PROG_INC / PROG_STD) */
    bool pair_tagged(const bbt_fe::SeqRecord &a, const bbt_fe::SeqRecord &b, size_t t) const {
        const bool p1 = passes(2 * t), p2 = passes(2 * t + 1);
        if (inclusive)
            return (kmers(a) > score && p1) || (kmers(b) > score && p2);
        return p1 && p2;
    }

    bool passes(size_t i) const { return w.h_pass[i] || (bait && w.h_pass[w.cap_reads + i]); }

    void count(uint64_t n) {
        for (uint64_t i = 0; i < n; ++i)
            if (++total_reads % interval == 0)
                cerr << "Currently Reading Read Number: " << total_reads
                     << "\tUnique k-mers Added: " << total
                     << "\tReads Used in Tagging: " << tagged << endl;
    }

    /* evalRead of every pending read against the filter (and the bait
     * filter), in one launch each */
    void evaluate(const deque<bbt_fe::SeqRecord> &q) {
        nreads = q.size();
        size_t bytes = 0;
        for (const auto &r : q)
            bytes += r.seq.size();
        w.reserve(nreads, bytes);
        uint64_t o = 0;
        for (size_t i = 0; i < nreads; ++i) {
            w.h_off[i] = o;
            memcpy(w.h_seq + o, q[i].seq.data(), q[i].seq.size());
            o += q[i].seq.size();
        }
        w.h_off[nreads] = o;
        HIP_OR_DIE(hipMemcpyAsync(w.d_seq, w.h_seq, (o + 7) / 8 * 8, hipMemcpyHostToDevice, w.st));
        HIP_OR_DIE(hipMemcpyAsync(w.d_off, w.h_off, (nreads + 1) * 8, hipMemcpyHostToDevice,
                                  w.st));
        bbt_reads in;
        memset(&in, 0, sizeof(in));
        in.seq = w.d_seq;
        in.seq_bytes = w.cap_bytes;
        in.off = w.d_off;
        in.nreads = nreads;
        check(bbt_eval_subtract(fs, 0, sub, 0, &in, scorer, score, 3, w.d_pass, w.st));
        HIP_OR_DIE(hipMemcpyAsync(w.h_pass, w.d_pass, nreads, hipMemcpyDeviceToHost, w.st));
        if (bait) {
            check(bbt_eval_subtract(bait, 0, sub, 0, &in, scorer, bait_score, 3,
                                    w.d_pass + w.cap_reads, w.st));
            HIP_OR_DIE(hipMemcpyAsync(w.h_pass + w.cap_reads, w.d_pass + w.cap_reads, nreads,
                                      hipMemcpyDeviceToHost, w.st));
        }
        HIP_OR_DIE(hipStreamSynchronize(w.st));
    }

    /* loadFilter(filter, read) for pending read i: its new k-mers */
    uint64_t insert(size_t i) {
        uint64_t c[2] = {0, 0};
        check(bbt_filter_insert_ordered(fs, 0, w.d_seq + w.h_off[i], w.h_off[i + 1] - w.h_off[i],
                                        c, w.st));
        total += c[0];
        return c[0];
    }

    /* checkFilter(filterSub, read): the read's k-mers the subtract filter
     * holds (counted as the removed k-mers of an insert into a scratch filter) */
    uint64_t repeats(size_t i) {
        if (!sub)
            return 0;
        if (!scratch) {
            uint64_t m = 64;
            check(bbt_filters_create(1, &m, &h, &k, &scratch));
        }
        uint64_t c[3] = {0, 0, 0};
        check(bbt_filter_insert_subtract(scratch, 0, sub, 0, w.d_seq + w.h_off[i],
                                         w.h_off[i + 1] - w.h_off[i], c, w.st));
        return c[2];
    }

    /* printDebug (BloomFilterGenerator.h:101-106) */
    void debug(const bbt_fe::SeqRecord &r, uint64_t kmers_added, uint64_t rep) {
        cout << "@" << r.name << " " << kmers_added << " " << total << " " << rep << " "
             << tagged << " " << total_reads << "\n"
             << r.seq << "\n+\n"
             << r.qual << "\n";
    }
};

} // namespace

int main(int argc, char *argv[]) {
    string prefix, outDir, subtractFilter, fileList;
    double fpr = 0.0078125;
    unsigned hashNum = 0, k = 25, iterations = 1;
    uint64_t entryNum = 0, interval = 10000000;
    double progressive = -1, baitThreshold = -1;
    bool die = false, inclusive = false, printReads = false, noRep = false;
    uint32_t scorer = BBT_SIMPLE;
    static struct option lo[] = {{"file_prefix", required_argument, NULL, 'p'},
                                 {"output_dir", required_argument, NULL, 'o'},
                                 {"fal_pos_rate", required_argument, NULL, 'f'},
                                 {"hash_num", required_argument, NULL, 'g'},
                                 {"kmer_size", required_argument, NULL, 'k'},
                                 {"num_ele", required_argument, NULL, 'n'},
                                 {"help", no_argument, NULL, 'h'},
                                 {"verbose", no_argument, NULL, 'v'},
                                 {"threads", required_argument, NULL, 't'},
                                 {"progressive", required_argument, NULL, 'r'},
                                 {"subtract", required_argument, NULL, 's'},
                                 {"no_rep_kmer", no_argument, NULL, 'd'},
                                 {"baitScore", required_argument, NULL, 'b'},
                                 {"iterations", required_argument, NULL, 'e'},
                                 {"inclusive", no_argument, NULL, 'i'},
                                 {"file_list", required_argument, NULL, 'l'},
                                 {"interval", required_argument, NULL, 'I'},
                                 {"print_reads", no_argument, NULL, 'P'},
                                 {NULL, 0, NULL, 0}};
    int c, idx = 0;
    while ((c = getopt_long(argc, argv, "f:p:o:k:n:g:hvs:t:Pr:ib:e:l:da:I:", lo, &idx)) !=
           -1) {
        stringstream cv(optarg ? optarg : "");
        switch (c) {
        case 'p': prefix = optarg; break;
        case 'o':
            outDir = optarg;
            if (!outDir.empty() && outDir.back() != '/')
                outDir += "/";
            break;
        case 'f':
            if (!(cv >> fpr)) {
                cerr << "Error - Invalid set of bloom filter parameters! f: " << optarg << endl;
                return 0;
            }
            if (fpr > 1) {
                cerr << "Error -f cannot be greater than 1 " << optarg << endl;
                return 0;
            }
            break;
        case 'g': cv >> hashNum; break;
        case 'k': cv >> k; break;
        case 'n': cv >> entryNum; break;
        case 'h': usage(); break;
        case 'v':
        case 't': break;
        case 's':
            if (!(cv >> subtractFilter)) {
                cerr << "Error - Invalid set of bloom filter parameters! s: " << optarg << endl;
                return 0;
            }
            break;
        case 'd': noRep = true; break;
        case 'i': inclusive = true; break;
        case 'P': printReads = true; break;
        case 'l': cv >> fileList; break;
        case 'I': cv >> interval; break;
        case 'r': {
            /* BioBloomMaker.cpp:262-289: an integer > 1 is a match length */
            unsigned matchLen;
            if ((cv >> matchLen) && matchLen > 1) {
                progressive = matchLen;
                scorer = BBT_LENGTH;
            } else {
                stringstream cv2(optarg);
                if (!(cv2 >> progressive)) {
                    cerr << "Error - Invalid set of bloom filter parameters! r: " << optarg
                         << endl;
                    return 0;
                }
                if (progressive < 0 || progressive > 1) {
                    cerr << "Error - r must be a positive integer or a floating "
                         << "point between 0 and 1. Input given:" << optarg << endl;
                    exit(EXIT_FAILURE);
                }
            }
            break;
        }
        case 'b':
            if (!(cv >> baitThreshold)) {
                cerr << "Error - Invalid set of bloom filter parameters! b: " << optarg << endl;
                return 0;
            }
            if (baitThreshold < 0) {
                cerr << "Error - b must be a positive integer or a floating "
                     << "point between 0 and 1. Input given:" << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        case 'e':
            if (!(cv >> iterations)) {
                cerr << "Error - Invalid set of bloom filter parameters! b: " << optarg << endl;
                return 0;
            }
            if (iterations < 1) {
                cerr << "Error - e must be > 1" << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        default: die = true;
        }
    }
    vector<string> inputs;
    while (optind < argc)
        inputs.push_back(argv[optind++]);
    if (prefix.empty()) {
        cerr << "Error: Need Filter Prefix ID" << endl;
        die = true;
    }
    if (inputs.empty()) {
        cerr << "Error: Need Input File" << endl;
        die = true;
    }
    if (prefix.find('/') != string::npos) {
        cerr << "Prefix ID cannot have '/' characters" << endl;
        die = true;
    }
    if (die) {
        cerr << "Try '--help' for more information.\n";
        exit(EXIT_FAILURE);
    }
    if (hashNum == 0)
        hashNum = bbt_calc_optimal_hashnum(fpr);
    /* BioBloomMaker.cpp:378-411: progressive input checks */
    string file1, file2;
    if (progressive != -1) {
        if (baitThreshold == -1) {
            baitThreshold = progressive;
        } else if ((baitThreshold < 1 && progressive > 1) ||
                   (baitThreshold > 1 && progressive < 1)) {
            cerr << "Both the bait threshold and progressive bloom filter must be either < 1 "
                    "or > 1"
                 << endl;
            exit(1);
        }
        if (inputs.size() > 2 && fileList.empty()) {
            file2 = inputs.back();
            inputs.pop_back();
            file1 = inputs.back();
            inputs.pop_back();
        } else if (fileList.empty()) {
            cerr << "at least 3 inputs are required when using progressive mode\n"
                    "biobloommaker <options> seed_seq reads1 reads2"
                 << endl;
            exit(1);
        }
        cerr << "Building Bloom filter in progressive mode. ";
        if (scorer == BBT_LENGTH)
            cerr << "Min match length = " << unsigned(std::round(progressive)) << " bp" << endl;
        else
            cerr << "Score threshold = " << progressive << endl;
    }
    /* read every seed record; records joined by '\n' so no k-mer spans two */
    string all;
    uint64_t expected = 0;
    string sources;
    for (const string &f : inputs) {
        /* records as the sequential kseq reader gives them, parsed on threads */
        bbt_fe::ParallelSeqReader rd(f);
        if (!rd.ok()) {
            cerr << "file " << f << " cannot be opened" << endl;
            exit(1);
        }
        cerr << "Opening File " << f << endl;
        bbt_fe::SeqCursor cur(rd);
        while (const bbt_fe::SeqView *v = cur.next()) {
            expected += uint64_t(v->seq_len) - k + 1; /* size_t arithmetic */
            all.append(v->seq, v->seq_len);
            all.push_back('\n');
        }
        sources += f + " ";
    }
    const uint64_t fileEntries = expected;
    if (entryNum == 0)
        entryNum = expected;
    const uint64_t m = bbt_calc_optimal_size(entryNum, fpr, hashNum);
    cerr << "Allocating " << m
         << " bits of space for filter and will output filter this size (plus header)"
         << endl;
    bbt_filterset *fs = nullptr;
    uint32_t hh = hashNum, kk = k;
    uint64_t mm = m;
    check(bbt_filters_create(1, &mm, &hh, &kk, &fs));
    void *d = nullptr;
    if (hipMalloc(&d, all.size() + 8) != hipSuccess ||
        hipMemcpy(d, all.data(), all.size(), hipMemcpyHostToDevice) != hipSuccess) {
        cerr << "device upload failed" << endl;
        exit(1);
    }
    bbt_filterset *sub = nullptr;
    if (!subtractFilter.empty()) {
        /* generate(filename, subtractFilter) / the progressive builders load
         * the subtract filter with its .txt */
        const char *sp = subtractFilter.c_str();
        check(bbt_filters_load(&sp, 1, &sub));
        if (progressive != -1) {
            /* checkFilters (BloomFilterGenerator.h:83-99) */
            bbt_filter_info si;
            check(bbt_filter_info_get(sub, 0, &si));
            if (si.h != hashNum)
                fatal("Error: Subtraction filter's hash number " + to_string(si.h) +
                      " is a different size than output filter's hash number " +
                      to_string(hashNum));
            if (si.k != k)
                fatal("Error: Subtraction filter's k-mer size " + to_string(si.k) +
                      " is a different size than output filter's k-mer size " + to_string(k));
        }
    }
    uint64_t counts[3] = {0, 0, 0};
    const bool subtractSeed = sub && (progressive == -1 || noRep);
    if (subtractSeed) {
        check(bbt_filter_insert_subtract(fs, 0, sub, 0, static_cast<const char *>(d), all.size(),
                                         counts, nullptr));
        cerr << "Total Number of K-mers not added: " << counts[2] << endl;
    } else {
        check(bbt_filter_insert(fs, 0, static_cast<const char *>(d), all.size(), counts,
                                nullptr));
    }
    uint64_t totalEntries = counts[0];
    if (progressive != -1) {
        cerr << "Approximated (due to false positives) total unique k-mers in reference files "
             << totalEntries << endl;
        Progressive pg;
        pg.fs = fs;
        pg.sub = sub;
        pg.score = progressive;
        pg.scorer = scorer;
        pg.k = k;
        pg.h = hashNum;
        pg.inclusive = inclusive;
        pg.print = printReads;
        pg.expected = entryNum;
        pg.total = totalEntries;
        pg.interval = interval ? interval : 1;
        /* the bait builders: a second filter of the seed, fixed, with its own
         * threshold (BloomFilterGenerator.cpp:321-328) */
        vector<string> list1, list2;
        bool listPaired = false;
        if (!fileList.empty()) {
            ifstream lf(fileList.c_str());
            if (!lf.is_open()) {
                cout << "Unable to open file";
                return 0;
            }
            string line;
            bool first = true;
            while (getline(lf, line)) {
                stringstream ss(line);
                string a, b;
                ss >> a;
                if (first)
                    listPaired = bool(ss >> b);
                else if (listPaired)
                    ss >> b;
                first = false;
                list1.push_back(a);
                if (listPaired)
                    list2.push_back(b);
            }
            cerr << (listPaired ? "Using file list paired mode" : "Using file list single end mode")
                 << endl;
            if (!listPaired && baitThreshold != progressive)
                fatal("single end bait mode not implemented. If needed feature request "
                      "cjustin@bcgsc.ca.");
        }
        if (baitThreshold != progressive) {
            uint64_t bm = bbt_calc_optimal_size(fileEntries, fpr, hashNum);
            check(bbt_filters_create(1, &bm, &hh, &kk, &pg.bait));
            uint64_t bc[2];
            check(bbt_filter_insert(pg.bait, 0, static_cast<const char *>(d), all.size(), bc,
                                    nullptr));
            pg.bait_score = baitThreshold;
        }
        /* the file-list builders skip everything once the seed alone is enough */
        const bool run = fileList.empty() || pg.total < pg.expected;
        for (unsigned it = 0; run && it < iterations; ++it) {
            cerr << "Iteration " << it + 1 << endl;
            if (fileList.empty()) {
                pg.pairs(file1, file2, "Reading file ");
            } else if (listPaired) {
                for (size_t i = 0; i < list1.size(); ++i)
                    pg.pairs(list1[i], list2[i], "Reading File: ");
            } else {
                for (const string &f : list1)
                    pg.singles(f);
            }
        }
        if (pg.total >= pg.expected)
            cerr << "K-mer threshold reached at read " << pg.total_reads << endl;
        else
            cerr << "K-mer threshold not reached, number of k-mers: " << pg.total << endl;
        totalEntries = pg.total;
        if (pg.bait)
            bbt_filters_free(pg.bait);
        cout.flush();
    } else {
        cerr << "Approximated (due to false positives) total unique k-mers in reference files "
             << totalEntries << endl;
    }
    if (sub)
        bbt_filters_free(sub);
    (void)hipFree(d);
    check(bbt_filter_refresh(fs, 0, nullptr));
    if (bbt_filter_set_id(fs, 0, prefix.c_str()) != BBT_OK)
        fatal(bbt_last_error());
    if (bbt_filter_store(fs, 0, (outDir + prefix).c_str(), fpr, entryNum, totalEntries, counts[1],
                         sources.c_str()) != BBT_OK)
        fatal(bbt_last_error());
    bbt_filters_free(fs);
    cerr << "Filter Creation Complete." << endl;
    return 0;
}
