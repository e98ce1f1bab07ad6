/*
 * biobloomcategorizer -- the reference CLI (BioBloomCategorizer/
 * BioBloomCategorizer.cpp:145-533) over libbbt_hip: the same options,
 * threshold interpretation, checks and dispatch; classification runs on the
 * GPU, -D dust masking included (a device sdust pre-pass), and -e with one
 * input file pairs mates by name within it (smart pair).
 */
#include <getopt.h>
#include <sys/stat.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "front_end.h"

using namespace std;

static void printHelpDialog() {
    cerr << "Usage: biobloomcategorizer [OPTION]... -f \"[FILTER1]...\" [FILE]...\n"
            "biobloomcategorizer [OPTION]... -e -f \"[FILTER1]...\" [FILE1.fq] [FILE2.fq]\n"
            "Categorize Sequences. The input format may be FASTA, FASTQ, and compressed gz.\n"
            "\n"
            "  -p, --prefix=N         Output prefix to use.\n"
            "  -f, --filter_files=N   List of filter files to use. Required option.\n"
            "  -e, --paired_mode      Uses paired-end information.\n"
            "  -i, --inclusive        If one paired read matches, both reads will be included.\n"
            "  -s, --score=N          Score threshold (0..1) or minimum match length (>1).\n"
            "  -b, --best_hit         The best hit is used rather than the score threshold.\n"
            "  -w, --with_score       Output multimatches with scores in the order of filter.\n"
            "  -t, --threads=N        Threads that parse each input file. [all; at most 8\n"
            "                         for plain files, 16 for gzip]\n"
            "  -g, --gz_output        Outputs all output files in compressed gzip.\n"
            "      --fa               Output categorized reads in Fasta files.\n"
            "      --fq               Output categorized reads in Fastq files.\n"
            "  -l, --file_list=N      A file of list of file pairs.\n"
            "  -I, --interval         Interval to report file processing status [10000000]\n"
            "  -r, --streak=N         Hits tiling needed to jump tiles upon a miss. [3]\n"
            "  -c, --ordered          Use ordered filtering.\n"
            "  -d, --stdout_filter    Outputs all matching reads to stdout for the first filter.\n"
            "  -n, --inverse          Inverts the output of -d.\n"
            "  -S, --score_type=N     'harmonic' or 'binomial' scoring. [simple]\n"
            "      --status           Also write <prefix>_readStatus.tsv (read -> class).\n"
            "      --plain            Use the plain one-launch-per-batch kernel.\n"
            "      --ascii_reads      Send reads to the device as text (default: 2-bit\n"
            "                         codes and an N mask, 3/8 of the bytes).\n"
            "      --tune_file=FILE   Tuning table of the batched kernels: the knobs for\n"
            "                         this filter size and read length are taken from it.\n"
            "      --autotune         Time the batched kernels' knobs on the first batch\n"
            "                         and keep the fastest (saved to --tune_file).\n"
            "      --batch_reads=N    Reads (or pairs) per device batch. [262144]\n"
            "      --ingest_block=N   Bytes of input parsed per block. [16777216]\n"
            "  -h, --help             Display this dialog.\n";
    exit(EXIT_SUCCESS);
}

static vector<string> convertInputString(const string &s) {
    vector<string> out;
    string t;
    stringstream c(s);
    while (c >> t)
        out.push_back(t);
    return out;
}

static void folderCheck(const string &path) {
    struct stat sb;
    if (stat(path.c_str(), &sb) == 0) {
        if (!S_ISDIR(sb.st_mode)) {
            cerr << "Error: Output folder - file exists with this name. " << path << endl;
            exit(1);
        }
    } else {
        cerr << "Error: Output folder does not exist. " << path << endl;
        exit(1);
    }
}

static bool fexists(const string &f) {
    ifstream i(f.c_str());
    return i.good();
}

int main(int argc, char *argv[]) {
    bbt_fe::Options opt;
    bool die = false, paired = false, stdoutFilter = false, collab = false;
    int fastq = 0, fasta = 0, status = 0, plain = 0, ascii_reads = 0, autotune = 0;
    string filtersFile, prefix, postfix, outputType, fileListFilename;
    double binomialScore = 100;
    static struct option long_options[] = {
        {"prefix", required_argument, NULL, 'p'},
        {"filter_files", required_argument, NULL, 'f'},
        {"paired_mode", no_argument, NULL, 'e'},
        {"inclusive", no_argument, NULL, 'i'},
        {"score", required_argument, NULL, 's'},
        {"help", no_argument, NULL, 'h'},
        {"interval", required_argument, NULL, 'I'},
        {"threads", required_argument, NULL, 't'},
        {"gz_output", no_argument, NULL, 'g'},
        {"fq", no_argument, &fastq, 1},
        {"fa", no_argument, &fasta, 1},
        {"status", no_argument, &status, 1},
        {"plain", no_argument, &plain, 1},
        {"ascii_reads", no_argument, &ascii_reads, 1},
        {"autotune", no_argument, &autotune, 1},
        {"tune_file", required_argument, NULL, 1002},
        {"file_list", required_argument, NULL, 'l'},
        {"version", no_argument, NULL, 'v'},
        {"multi", required_argument, NULL, 'm'},
        {"streak", required_argument, NULL, 'r'},
        {"ordered", no_argument, NULL, 'c'},
        {"stdout_filter", no_argument, NULL, 'd'},
        {"inverse", no_argument, NULL, 'n'},
        {"best_hit", no_argument, NULL, 'b'},
        {"with_score", no_argument, NULL, 'w'},
        {"score_type", required_argument, NULL, 'S'},
        {"dust", no_argument, NULL, 'D'},
        {"T_dust", required_argument, NULL, 'T'},
        {"window_dust", required_argument, NULL, 'W'},
        {"batch_reads", required_argument, NULL, 1000},
        {"ingest_block", required_argument, NULL, 1001},
        {NULL, 0, NULL, 0}};
    int c, idx = 0;
    while ((c = getopt_long(argc, argv, "f:p:hegl:vs:r:t:cdiwI:nm:S:bDT:W:", long_options,
                            &idx)) != -1) {
        switch (c) {
        case 's': {
            stringstream cv(optarg);
            cv >> opt.score;
            binomialScore = opt.score;
            break;
        }
        case 'f': filtersFile = optarg; break;
        case 'p': prefix = optarg; break;
        case 'h': printHelpDialog(); break;
        case 'I': {
            stringstream cv(optarg);
            if (!(cv >> opt.file_interval)) {
                cerr << "Error - Invalid parameters! I: " << optarg << endl;
                return 0;
            }
            break;
        }
        case 1000: {
            stringstream cv(optarg);
            if (!(cv >> opt.batch_reads) || opt.batch_reads == 0) {
                cerr << "Error - Invalid parameters! batch_reads: " << optarg << endl;
                die = true;
            }
            break;
        }
        case 1001: {
            stringstream cv(optarg);
            if (!(cv >> opt.ingest_block) || opt.ingest_block == 0) {
                cerr << "Error - Invalid parameters! ingest_block: " << optarg << endl;
                die = true;
            }
            break;
        }
        case 1002: opt.tune_file = optarg; break;
        case 'e': paired = true; break;
        case 'i': opt.inclusive = true; break;
        case 't': {
            stringstream cv(optarg);
            if (!(cv >> opt.threads)) {
                cerr << "Error - Invalid parameter! t: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        }
        case 'g': postfix = ".gz"; break;
        case 'l': fileListFilename = optarg; break;
        case 'v': cerr << "biobloomcategorizer (BioBloom Tools, MI355X path)" << endl; exit(0);
        case 'r': {
            stringstream cv(optarg);
            if (!(cv >> opt.streak)) {
                cerr << "Error - Invalid parameter! r: " << optarg << endl;
                exit(EXIT_FAILURE);
            }
            break;
        }
        case 'c': collab = true; break;
        case 'd': stdoutFilter = true; break;
        case 'n': opt.inverse = true; break;
        case 'S':
            if (strcmp(optarg, "harmonic") == 0)
                opt.scorer = BBT_HARMONIC;
            else if (strcmp(optarg, "binomial") == 0)
                opt.scorer = BBT_BINOMIAL;
            else {
                cerr << "scoring method not recognized. Valid strings: harmonic or binomial"
                     << endl;
                die = true;
            }
            break;
        case 'm': break;
        case 'w':
            if (opt.mode == BBT_STD)
                opt.mode = BBT_SCORES;
            else {
                cerr << "Filter scoring modes cannot be mixed with score mode" << endl;
                die = true;
            }
            break;
        case 'b':
            if (opt.mode == BBT_STD)
                opt.mode = BBT_BESTHIT;
            else {
                cerr << "Filter scoring modes cannot be mixed with best hit mode" << endl;
                die = true;
            }
            break;
        case 'D':
            opt.dust = true;
            break;
        case 'T': {
            stringstream convert(optarg);
            convert >> opt.dust_t;
            break;
        }
        case 'W': {
            stringstream convert(optarg);
            convert >> opt.dust_w;
            break;
        }
        case '?': die = true; break;
        }
    }
    opt.read_status = status != 0;
    if (plain)
        opt.path = BBT_PATH_PLAIN;
    opt.pack2bit = ascii_reads == 0;
    opt.autotune = autotune != 0;
    vector<string> filterFilePaths = convertInputString(filtersFile);
    vector<string> inputFiles;
    while (optind < argc)
        inputFiles.push_back(argv[optind++]);
    /* BioBloomCategorizer.cpp:357-370: one input file with -e is a smart pair */
    bool smartPair = false;
    if (paired) {
        if (inputFiles.size() == 1) {
            smartPair = true;
        } else if (inputFiles.size() != 2 && fileListFilename.empty()) {
            cerr << "Usage of paired end mode:\n"
                 << "BioBloomCategorizer [OPTION]... -f \"[FILTER1]...\" [FILEPAIR1] [FILEPAIR2]\n"
                 << "or BioBloomCategorizer [OPTION]... -f \"[FILTER1]...\" [SMARTPAIR]\n"
                 << endl;
            die = true;
        }
    }
    if (!fileListFilename.empty() && !inputFiles.empty()) {
        cerr << "--file_list (-l) cannot be used with read files in specified in arguments"
             << endl;
        die = true;
    }
    if (inputFiles.empty() && fileListFilename.empty()) {
        cerr << "Error: Need Input File" << endl;
        die = true;
    }
    if (filterFilePaths.empty()) {
        cerr << "Error: Need Filter File (-f)" << endl;
        die = true;
    }
    if (die) {
        cerr << "Try '--help' for more information.\n";
        exit(EXIT_FAILURE);
    }
    if (prefix.find('/') != string::npos)
        folderCheck(prefix.substr(0, prefix.find_last_of("/")));
    if (fastq && fasta) {
        cerr << "Error: fasta (--fa) and fastq (--fq) outputs types cannot be both set" << endl;
        exit(1);
    } else if (fastq) {
        outputType = "fq";
    } else if (fasta) {
        outputType = "fa";
    }
    if (opt.mode == BBT_SCORES && outputType.empty()) {
        cerr << "Error: -w option cannot be used without output method" << endl;
        exit(EXIT_FAILURE);
    }
    /* threshold interpretation (BioBloomCategorizer.cpp:416-441) */
    if (opt.mode == BBT_BESTHIT) {
        cerr << "Best Hit mode (no minimum threshold)" << endl;
    } else if (opt.scorer == BBT_BINOMIAL) {
        opt.score = pow(10.0, -(binomialScore / 10.0));
        cerr << "FPR of a match: " << opt.score << endl;
    } else {
        unsigned matchLen = unsigned(opt.score);
        if (matchLen > 1) {
            opt.score = (unsigned)matchLen;
            opt.scorer = BBT_LENGTH;
            cerr << "Min match length threshold: " << matchLen << " bp" << endl;
        } else {
            cerr << "Min score threshold: " << opt.score << endl;
        }
        if (opt.score < 0) {
            cerr << "Error - s must be a positive integer or a floating "
                 << "point between 0 and 1." << endl;
            exit(EXIT_FAILURE);
        }
    }
    for (const string &f : filterFilePaths)
        if (!fexists(f)) {
            cerr << "Error: " + f + " File cannot be opened" << endl;
            exit(EXIT_FAILURE);
        }
    bbt_fe::BioBloomClassifier bbc(filterFilePaths, opt.score, prefix, postfix, opt);
    if (stdoutFilter)
        bbc.setStdout();
    if (collab)
        bbc.setOrderedFilter();
    vector<string> list1, list2;
    if (!fileListFilename.empty()) {
        ifstream lf(fileListFilename.c_str());
        string line;
        while (getline(lf, line)) {
            stringstream ss(line);
            string a, b;
            ss >> a;
            ss >> b;
            list1.push_back(a);
            list2.push_back(b);
        }
        cerr << "Using file list" << endl;
    }
    if (paired) {
        if (opt.inclusive)
            bbc.setInclusive();
        if (!outputType.empty() && smartPair)
            bbc.filterPairPrint(inputFiles[0], outputType);
        else if (!outputType.empty())
            bbc.filterPairPrint(inputFiles[0], inputFiles[1], outputType);
        else if (smartPair)
            bbc.filterPair(inputFiles[0]);
        else if (!fileListFilename.empty())
            bbc.filterPair(list1, list2);
        else
            bbc.filterPair(inputFiles[0], inputFiles[1]);
    } else {
        const vector<string> &files = fileListFilename.empty() ? inputFiles : list1;
        if (!outputType.empty())
            bbc.filterPrint(files, outputType);
        else
            bbc.filter(files);
    }
    cerr << "k-mer probes: " << bbc.probes() << "  bit lookups: " << bbc.lookups() << endl;
    return 0;
}
