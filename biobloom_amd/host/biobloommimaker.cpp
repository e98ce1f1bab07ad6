/*
 * biobloommimaker -- the reference's multi-index Bloom filter maker
 * (BioBloomMaker/BioBloomMIMaker.cpp:83-256, MIBFGen.hpp:44-309) over
 * libbbt_hip: the same options, IDs per FASTA header (or per file with -F),
 * entry count sum(len - k + 1), occupancy sizing, and the <prefix>.bf +
 * <prefix>_ids.txt outputs; the three construction passes run on the GPU
 * (bbt_mibf_build).
 *
 * Differences kept on purpose: a record shorter than the window adds 0 to the
 * entry count (the reference's size_t arithmetic wraps); IDs are 32-bit (the
 * reference's are 16-bit, MIBFGen.hpp:112-118), so up to 2^31 - 1 IDs.
 */
#include <getopt.h>
#include <hip/hip_runtime.h>

#include <cstring>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "front_end.h"

using namespace std;

static void usage() {
    cerr << "Usage: biobloommimaker -p [FILTERID] [OPTION]... [FILE]...\n"
            "Creates a multi-index Bloom Filter from a list of fasta files.\n"
            "\n"
            "  -p, --file_prefix=N    Filter prefix and filter ID. Required option.\n"
            "  -h, --help             Display this dialog.\n"
            "      --version          Display version information.\n"
            "  -v  --verbose          Display verbose output.\n"
            "  -t, --threads=N        The number of threads to use.\n"
            "  -b, --occupancy=N      Occupancy of Bloom filter.[0.5]\n"
            "  -n, --num_ele=N        Set the number of expected distinct k-mer frames.\n"
            "                         If set to 0 number is determined from sequences\n"
            "                         sizes within files. [0]\n"
            "  -S, --seed_str=N       Generate a miBF using multiple spaced seeds. Expects\n"
            "                         list of seed 1s & 0s separated by spaces.\n"
            "  -F, --by_file          Assign IDs by file rather than by fasta header\n"
            "k-mer mode options (disabled when using spaced seeds):\n"
            "  -g, --hash_num=N       Set number of hash functions when using k-mers.\n"
            "  -k, --kmer_size=N      K-mer size to use to create filter. [25]\n";
    exit(0);
}

static vector<string> split_ws(const string &s) {
    vector<string> out;
    stringstream cv(s);
    string t;
    while (cv >> t)
        out.push_back(t);
    return out;
}

int main(int argc, char *argv[]) {
    string prefix;
    double occupancy = 0.5;
    uint64_t entryNum = 0;
    unsigned hashNum = 0, k = 25;
    int verbose = 0;
    bool byFile = false, die = false;
    vector<string> seeds;
    enum { OPT_VERSION = 1000 };
    static struct option lo[] = {{"file_prefix", required_argument, NULL, 'p'},
                                 {"help", no_argument, NULL, 'h'},
                                 {"threads", required_argument, NULL, 't'},
                                 {"occupancy", required_argument, NULL, 'b'},
                                 {"seed_str", required_argument, NULL, 'S'},
                                 {"hash_num", required_argument, NULL, 'g'},
                                 {"kmer_size", required_argument, NULL, 'k'},
                                 {"num_ele", required_argument, NULL, 'n'},
                                 {"by_file", no_argument, NULL, 'F'},
                                 {"verbose", no_argument, NULL, 'v'},
                                 {"version", no_argument, NULL, OPT_VERSION},
                                 {NULL, 0, NULL, 0}};
    int c, idx = 0;
    while ((c = getopt_long(argc, argv, "p:ht:b:S:g:k:n:Fv", lo, &idx)) != -1) {
        stringstream cv(optarg ? optarg : "");
        switch (c) {
        case 'b':
            if (!(cv >> occupancy)) {
                cerr << "Error - Invalid set of bloom filter parameters! b: " << optarg << endl;
                return 0;
            }
            break;
        case 'p': prefix = optarg; break;
        case 't': break; /* host threads: the passes run on the GPU */
        case 'S':
            seeds = split_ws(optarg);
            if (!seeds.empty())
                k = unsigned(seeds[0].size());
            break;
        case 'F': byFile = true; break;
        case 'k':
            if (!(cv >> k)) {
                cerr << "Error - Invalid set of bloom filter parameters! k: " << optarg << endl;
                return 0;
            }
            break;
        case 'g':
            if (!(cv >> hashNum)) {
                cerr << "Error - Invalid set of bloom filter parameters! g: " << optarg << endl;
                return 0;
            }
            break;
        case 'n':
            if (!(cv >> entryNum)) {
                cerr << "Error - Invalid set of bloom filter parameters! n: " << optarg << endl;
                return 0;
            }
            break;
        case 'h': usage(); break;
        case 'v': ++verbose; break;
        case OPT_VERSION:
            cerr << "biobloommimaker (biobloom_amd)" << endl;
            return 0;
        default: die = true;
        }
    }
    vector<string> inputs;
    while (optind < argc)
        inputs.push_back(argv[optind++]);
    if (inputs.empty()) {
        cerr << "Need Input File" << endl;
        die = true;
    }
    if (prefix.empty()) {
        cerr << "Need Filter Prefix ID" << endl;
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
    if (!seeds.empty()) {
        hashNum = unsigned(seeds.size());
        if (verbose)
            cerr << "Spaced Seeds Detected" << endl;
    } else if (verbose) {
        cerr << "K-mer mode detected" << endl;
    }
    if (hashNum == 0) {
        cerr << "Please pick number of hash values (-g)\n";
        exit(1);
    }

    /* IDs (MIBFGen.hpp:44-129): one per file name (-F) or per header; a
     * name seen again maps to its latest ID */
    vector<string> names;
    map<string, uint32_t> nameToID;
    if (byFile)
        for (const string &f : inputs) {
            names.push_back(f.substr(f.find_last_of('/') + 1));
            nameToID[names.back()] = uint32_t(names.size());
        }
    string all;
    vector<uint64_t> rec_off{0};
    vector<string> rec_name;
    uint64_t counts = 0;
    for (const string &f : inputs) {
        /* records as the sequential kseq reader gives them, parsed on threads */
        bbt_fe::ParallelSeqReader rd(f);
        if (!rd.ok()) {
            cerr << "file " << f << " cannot be opened" << endl;
            exit(1);
        }
        if (verbose)
            cerr << "Opening " << f << endl;
        bbt_fe::SeqCursor cur(rd);
        const string fname = f.substr(f.find_last_of('/') + 1);
        while (const bbt_fe::SeqView *v = cur.next()) {
            const string name(v->name, v->name_len);
            if (!byFile) {
                names.push_back(name);
                nameToID[name] = uint32_t(names.size());
            }
            rec_name.push_back(byFile ? fname : name);
            counts += v->seq_len >= k ? v->seq_len - k + 1 : 0;
            all.append(v->seq, v->seq_len);
            all.push_back('\n'); /* the separator bbt_mibf_build requires */
            rec_off.push_back(all.size());
        }
    }
    if (rec_name.empty()) {
        cerr << "Error: no sequences in the input files" << endl;
        exit(1);
    }
    if (names.size() >= 0x7FFFFFFFu) {
        cerr << "Max Size: " << 0x7FFFFFFFu << " Required Size: " << names.size() + 1 << endl;
        exit(1);
    }
    vector<uint32_t> rec_id;
    for (const string &n : rec_name)
        rec_id.push_back(nameToID[n]);
    if (entryNum == 0)
        entryNum = counts;
    if (verbose)
        cerr << "Expected number of elements: " << entryNum << endl;

    char *dseq = nullptr;
    uint64_t *doff = nullptr;
    uint32_t *did = nullptr;
    if (hipMalloc(&dseq, all.size() + 8) != hipSuccess ||
        hipMalloc(&doff, rec_off.size() * 8) != hipSuccess ||
        hipMalloc(&did, rec_id.size() * 4) != hipSuccess ||
        hipMemcpy(dseq, all.data(), all.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(doff, rec_off.data(), rec_off.size() * 8, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(did, rec_id.data(), rec_id.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        cerr << "device upload failed" << endl;
        exit(1);
    }
    vector<const char *> np, sp;
    for (const string &n : names)
        np.push_back(n.c_str());
    for (const string &s : seeds)
        sp.push_back(s.c_str());
    bbt_mibf *mf = nullptr;
    if (bbt_mibf_build(dseq, doff, did, rec_off.data(), rec_off.size() - 1, np.data(),
                       uint32_t(names.size()), seeds.empty() ? nullptr : sp.data(), hashNum, k,
                       occupancy, entryNum, 0, &mf, nullptr) != BBT_OK) {
        cerr << "Error: " << bbt_last_error() << endl;
        exit(1);
    }
    (void)hipFree(dseq);
    (void)hipFree(doff);
    (void)hipFree(did);
    bbt_mibf_info_t info;
    bbt_mibf_info(mf, &info);
    cerr << "Outputting IDs file: " << prefix + "_ids.txt" << endl;
    cerr << "PopCount: " << info.pop << endl;
    cerr << "PopSaturated: " << info.pop_saturated << endl;
    cerr << "PopCount Ratio: " << double(info.pop) / double(info.m) << endl;
    cerr << "Storing filter" << endl;
    if (bbt_mibf_store(mf, prefix.c_str()) != BBT_OK) {
        cerr << "Error: " << bbt_last_error() << endl;
        exit(1);
    }
    bbt_mibf_free(mf);
    if (verbose)
        cerr << "Multi Index Bloom Filter Creation Complete." << endl;
    return 0;
}
