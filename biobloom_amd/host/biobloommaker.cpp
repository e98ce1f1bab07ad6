/*
 * biobloommaker -- the reference maker (BioBloomMaker/BioBloomMaker.cpp:
 * 110-547, standard mode) over libbbt_hip: the same sizing
 * (BloomFilterInfo.h:57-76), k-mer insertion on the GPU, and the .bf/.txt
 * outputs, and -s (a filter whose k-mers are left out, BloomFilterGenerator.cpp:
 * 1175-1199).  Progressive and bait modes are not on this path and are
 * rejected.
 */
#include <getopt.h>
#include <hip/hip_runtime.h>

#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "front_end.h"

using namespace std;

static void usage() {
    cerr << "Usage: biobloommaker -p [FILTERID] [OPTION]... [FILE]...\n"
            "  -p, --file_prefix=N    Filter prefix and filter ID. Required option.\n"
            "  -o, --output_dir=N     Output location of the filter and filter info files.\n"
            "  -f, --fal_pos_rate=N   Maximum false positive rate to use in filter. [0.0078125]\n"
            "  -g, --hash_num=N       Set number of hash functions to use in filter.\n"
            "  -k, --kmer_size=N      K-mer size to use to create filter. [25]\n"
            "  -n, --num_ele=N        Set the number of expected elements.\n"
            "  -s, --subtract=N       Path to a filter whose k-mers are not added.\n"
            "  -h, --help             Display this dialog.\n";
    exit(EXIT_SUCCESS);
}

int main(int argc, char *argv[]) {
    string prefix, outDir, subtractFilter;
    double fpr = 0.0078125;
    unsigned hashNum = 0, k = 25;
    uint64_t entryNum = 0;
    bool die = false;
    static struct option lo[] = {{"file_prefix", required_argument, NULL, 'p'},
                                 {"output_dir", required_argument, NULL, 'o'},
                                 {"fal_pos_rate", required_argument, NULL, 'f'},
                                 {"hash_num", required_argument, NULL, 'g'},
                                 {"kmer_size", required_argument, NULL, 'k'},
                                 {"num_ele", required_argument, NULL, 'n'},
                                 {"help", no_argument, NULL, 'h'},
                                 {"progressive", required_argument, NULL, 'r'},
                                 {"subtract", required_argument, NULL, 's'},
                                 {NULL, 0, NULL, 0}};
    int c, idx = 0;
    while ((c = getopt_long(argc, argv, "p:o:f:g:k:n:hr:s:", lo, &idx)) != -1) {
        stringstream cv(optarg ? optarg : "");
        switch (c) {
        case 'p': prefix = optarg; break;
        case 'o':
            outDir = optarg;
            if (!outDir.empty() && outDir.back() != '/')
                outDir += "/";
            break;
        case 'f': cv >> fpr; break;
        case 'g': cv >> hashNum; break;
        case 'k': cv >> k; break;
        case 'n': cv >> entryNum; break;
        case 'h': usage(); break;
        case 's':
            if (!(cv >> subtractFilter)) {
                cerr << "Error - Invalid set of bloom filter parameters! s: " << optarg << endl;
                return 0;
            }
            break;
        case 'r':
            cerr << "Error: progressive filters are not available on this path" << endl;
            die = true;
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
    /* read every record; records joined by '\n' so no k-mer spans two */
    string all;
    uint64_t expected = 0;
    string sources;
    for (const string &f : inputs) {
        bbt_fe::SeqReader rd(f);
        if (!rd.ok()) {
            cerr << "file " << f << " cannot be opened" << endl;
            exit(1);
        }
        cerr << "Opening File " << f << endl;
        bbt_fe::SeqRecord rec;
        while (rd.read(rec) >= 0) {
            expected += uint64_t(rec.seq.size()) - k + 1; /* size_t arithmetic */
            all += rec.seq;
            all.push_back('\n');
        }
        sources += f + " ";
    }
    if (entryNum == 0)
        entryNum = expected;
    const uint64_t m = bbt_calc_optimal_size(entryNum, fpr, hashNum);
    cerr << "Allocating " << m
         << " bits of space for filter and will output filter this size (plus header)"
         << endl;
    bbt_filterset *fs = nullptr;
    uint32_t hh = hashNum, kk = k;
    uint64_t mm = m;
    if (bbt_filters_create(1, &mm, &hh, &kk, &fs) != BBT_OK) {
        cerr << bbt_last_error() << endl;
        exit(1);
    }
    void *d = nullptr;
    if (hipMalloc(&d, all.size() + 8) != hipSuccess ||
        hipMemcpy(d, all.data(), all.size(), hipMemcpyHostToDevice) != hipSuccess) {
        cerr << "device upload failed" << endl;
        exit(1);
    }
    uint64_t counts[3] = {0, 0, 0};
    bbt_status rc;
    if (subtractFilter.empty()) {
        rc = bbt_filter_insert(fs, 0, static_cast<const char *>(d), all.size(), counts, nullptr);
    } else {
        /* generate(filename, subtractFilter): the subtract filter with its .txt */
        bbt_filterset *sub = nullptr;
        const char *sp = subtractFilter.c_str();
        if (bbt_filters_load(&sp, 1, &sub) != BBT_OK) {
            cerr << bbt_last_error() << endl;
            exit(1);
        }
        rc = bbt_filter_insert_subtract(fs, 0, sub, 0, static_cast<const char *>(d), all.size(),
                                        counts, nullptr);
        bbt_filters_free(sub);
        if (rc == BBT_OK)
            cerr << "Total Number of K-mers not added: " << counts[2] << endl;
    }
    if (rc != BBT_OK || bbt_filter_refresh(fs, 0, nullptr) != BBT_OK ||
        bbt_filter_set_id(fs, 0, prefix.c_str())) {
        cerr << bbt_last_error() << endl;
        exit(1);
    }
    (void)hipFree(d);
    cerr << "Approximated (due to false positives) total unique k-mers in reference files "
         << counts[0] << endl;
    if (bbt_filter_store(fs, 0, (outDir + prefix).c_str(), fpr, entryNum, counts[0], counts[1],
                         sources.c_str()) != BBT_OK) {
        cerr << bbt_last_error() << endl;
        exit(1);
    }
    bbt_filters_free(fs);
    cerr << "Filter Creation Complete." << endl;
    return 0;
}
