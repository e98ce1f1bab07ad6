/*
 * bbt_host.cpp -- filter file formats, sizing, binomial thresholds and the
 * summary text for libbbt_hip.  Host code only.
 */
#include "bbt_host.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace bbt {

/* ---------------------------------------------------------------- .txt */

static std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos)
        return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

bool read_info_file(const std::string &path, InfoFile &out, std::string &err) {
    std::ifstream in(path.c_str());
    if (!in.good()) {
        err = path + ": cannot open filter info file";
        return false;
    }
    std::string line, section;
    bool have_id = false;
    while (std::getline(in, line)) {
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#')
            continue;
        if (t[0] == '[') {
            size_t e = t.find(']');
            if (e == std::string::npos) {
                err = path + ": malformed section line";
                return false;
            }
            section = trim(t.substr(1, e - 1));
            continue;
        }
        size_t eq = t.find('=');
        if (eq == std::string::npos) {
            err = path + ": malformed line '" + t + "'";
            return false;
        }
        std::string key = section + "." + trim(t.substr(0, eq));
        std::string val = trim(t.substr(eq + 1));
        std::istringstream cv(val);
        if (key == "user_input_options.filter_id") {
            out.filter_id = val;
            have_id = true;
        } else if (key == "user_input_options.kmer_size") {
            cv >> out.kmer_size;
        } else if (key == "user_input_options.desired_false_positve_rate") {
            cv >> out.desired_fpr;
        } else if (key == "user_input_options.number_of_hash_functions") {
            cv >> out.hash_num;
        } else if (key == "user_input_options.expected_num_entries") {
            cv >> out.expected_entries;
        } else if (key == "user_input_options.sequence_sources") {
            out.sequence_sources = val;
        } else if (key == "runtime_options.size") {
            cv >> out.size;
        } else if (key == "runtime_options.num_entries") {
            cv >> out.num_entries;
        } else if (key == "runtime_options.approximate_false_positive_rate") {
            cv >> out.approx_fpr;
        } else if (key == "runtime_options.redundant_sequences") {
            cv >> out.redundant_sequences;
        } else if (key == "runtime_options.redundant_fpr") {
            cv >> out.redundant_fpr;
        }
    }
    if (!have_id) {
        err = path + ": missing user_input_options.filter_id";
        return false;
    }
    return true;
}

bool write_info_file(const std::string &path, const InfoFile &in, std::string &err) {
    std::ofstream o(path.c_str(), std::ios::out);
    if (!o.good()) {
        err = path + ": cannot write filter info file";
        return false;
    }
    o << "[user_input_options]\nfilter_id=" << in.filter_id
      << "\nkmer_size=" << in.kmer_size
      << "\ndesired_false_positve_rate=" << in.desired_fpr
      << "\nnumber_of_hash_functions=" << in.hash_num
      << "\nexpected_num_entries=" << in.expected_entries << "\nsequence_sources=";
    std::istringstream src(in.sequence_sources);
    std::string s;
    while (src >> s)
        o << s << " ";
    o << "\n\n[runtime_options]\nsize=" << in.size << "\nnum_entries=" << in.num_entries
      << "\napproximate_false_positive_rate=" << in.approx_fpr
      << "\nredundant_sequences=" << in.redundant_sequences
      << "\nredundant_fpr=" << in.redundant_fpr << "\n";
    o.close();
    if (!o.good()) {
        err = path + ": write failed";
        return false;
    }
    return true;
}

/* ----------------------------------------------------------------- .bf */

static const char kBfMagic[8] = {'B', 'l', 'O', 'O', 'M', 'F', 'X', 'X'};

bool read_bf_file(const std::string &path, BfHeader &hdr, std::vector<uint8_t> &bits,
                  std::string &err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = path + ": File cannot be opened";
        return false;
    }
    bool ok = std::fread(&hdr, sizeof(hdr), 1, f) == 1;
    if (!ok || std::memcmp(hdr.magic, kBfMagic, 8) != 0 || hdr.hlen != sizeof(hdr)) {
        std::fclose(f);
        err = path + ": not a Bloom filter file (bad header)";
        return false;
    }
    if (hdr.size == 0 || hdr.size % 8 != 0 || hdr.nhash == 0 || hdr.nhash > 16 ||
        hdr.kmer == 0 || hdr.kmer > 64) {
        std::fclose(f);
        err = path + ": unsupported filter geometry in header";
        return false;
    }
    std::fseek(f, 0, SEEK_END);
    long long fsize = std::ftell(f);
    uint64_t nbytes = hdr.size / 8;
    if (fsize < 0 || uint64_t(fsize) != sizeof(hdr) + nbytes) {
        std::fclose(f);
        err = path + ": file size does not match the header's filter size";
        return false;
    }
    std::fseek(f, long(sizeof(hdr)), SEEK_SET);
    bits.resize(nbytes);
    ok = nbytes == 0 || std::fread(bits.data(), 1, nbytes, f) == nbytes;
    std::fclose(f);
    if (!ok) {
        err = path + ": short read of bit array";
        return false;
    }
    return true;
}

bool write_bf_file(const std::string &path, const BfHeader &hdr0, const uint8_t *bits,
                   uint64_t nbytes, std::string &err) {
    BfHeader hdr = hdr0;
    std::memcpy(hdr.magic, kBfMagic, 8);
    hdr.hlen = sizeof(hdr);
    hdr.pad = 0;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = path + ": cannot write filter file";
        return false;
    }
    bool ok = std::fwrite(&hdr, sizeof(hdr), 1, f) == 1 &&
              (nbytes == 0 || std::fwrite(bits, 1, nbytes, f) == nbytes);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok)
        err = path + ": write failed";
    return ok;
}

/* -------------------------------------------------------------- sizing */

uint64_t calc_optimal_size(uint64_t entries, double fpr, unsigned h) {
    /* BloomFilterInfo::calcOptimalSize (Common/BloomFilterInfo.h:57-65) */
    uint64_t v = uint64_t(-double(entries) * double(h) /
                          std::log(1.0 - std::pow(fpr, double(1 / double(h)))));
    return v + (64 - v % 64);
}

unsigned calc_optimal_hashnum(double fpr) {
    /* BloomFilterInfo::calcOptimalHashNum (Common/BloomFilterInfo.h:73-76) */
    return unsigned(-std::log(fpr) / std::log(2));
}

double calc_approx_fpr(uint64_t size, uint64_t entries, unsigned h) {
    /* BloomFilterInfo::calcApproxFPR (Common/BloomFilterInfo.cpp:172-178) */
    return std::pow(1.0 - std::pow(1.0 - 1.0 / double(size), double(entries) * h),
                    double(h));
}

double calc_redundancy_fpr(uint64_t size, uint64_t entries, unsigned h) {
    /* BloomFilterInfo::calcRedunancyFPR (Common/BloomFilterInfo.cpp:183-191) */
    if (entries == 0)
        return 0;
    double total = std::log(calc_approx_fpr(size, 1, h));
    for (uint64_t i = 2; i < entries; ++i)
        total = std::log(std::exp(total) + calc_approx_fpr(size, i, h));
    return std::exp(total) / double(entries);
}

/* ------------------------------------------------------------ binomial */
/* Binomial probabilities by Loader's saddle-point method (C. Loader, "Fast
 * and accurate computation of binomial probabilities", 2000). */

static double stirling_err(double n) {
    const double S0 = 1.0 / 12, S1 = 1.0 / 360, S2 = 1.0 / 1260, S3 = 1.0 / 1680,
                 S4 = 1.0 / 1188;
    if (n <= 15.0) {
        long double x = (long double)n;
        int sign = 0; /* the reentrant form: tables are computed on several threads */
        return double(lgammal_r(x + 1.0L, &sign) - (x + 0.5L) * logl(x) + x -
                      0.5L * logl(2.0L * 3.14159265358979323846264338327950288L));
    }
    double nn = n * n;
    if (n > 500)
        return (S0 - S1 / nn) / n;
    if (n > 80)
        return (S0 - (S1 - S2 / nn) / nn) / n;
    if (n > 35)
        return (S0 - (S1 - (S2 - S3 / nn) / nn) / nn) / n;
    return (S0 - (S1 - (S2 - (S3 - S4 / nn) / nn) / nn) / nn) / n;
}

static double deviance(double x, double np) {
    if (std::fabs(x - np) < 0.1 * (x + np)) {
        double v = (x - np) / (x + np);
        double s = (x - np) * v, ej = 2 * x * v;
        v *= v;
        for (int j = 1; j < 1000; ++j) {
            ej *= v;
            double s1 = s + ej / (2 * j + 1);
            if (s1 == s)
                return s1;
            s = s1;
        }
    }
    return x * std::log(x / np) + np - x;
}

static double binom_pmf(double x, double n, double p) {
    double q = 1.0 - p;
    if (p == 0)
        return x == 0 ? 1 : 0;
    if (q == 0)
        return x == n ? 1 : 0;
    if (x == 0) {
        if (n == 0)
            return 1;
        return std::exp(p < 0.1 ? -deviance(n, n * q) - n * p : n * std::log(q));
    }
    if (x == n)
        return std::exp(q < 0.1 ? -deviance(n, n * p) - n * q : n * std::log(p));
    if (x < 0 || x > n)
        return 0;
    double lc = stirling_err(n) - stirling_err(x) - stirling_err(n - x) -
                deviance(x, n * p) - deviance(n - x, n * q);
    double lf = std::log(2 * 3.14159265358979323846) + std::log(x) + std::log1p(-x / n);
    return std::exp(lc - 0.5 * lf);
}

double binom_sf(int64_t x, uint64_t n, double p) {
    if (x < 0)
        return 1.0;
    if (uint64_t(x) >= n)
        return 0.0;
    const double mean = double(n) * p;
    double sum = 0;
    if (double(x) + 1 >= mean) {
        for (uint64_t i = uint64_t(x) + 1; i <= n; ++i) {
            double t = binom_pmf(double(i), double(n), p);
            sum += t;
            if (t < sum * 1e-18 && double(i) > mean)
                break;
        }
        return sum;
    }
    for (int64_t i = x; i >= 0; --i) {
        double t = binom_pmf(double(i), double(n), p);
        sum += t;
        if (t < sum * 1e-18 && double(i) < mean)
            break;
    }
    return 1.0 - sum;
}

uint64_t binom_isf_up(double q, uint64_t n, double p) {
    if (q >= 1.0)
        return 0;
    const double mean = double(n) * p;
    const double sd = std::sqrt(mean * (1 - p));
    const double top = std::ceil(mean + 40 * sd + 40);
    uint64_t x = top >= double(n) ? n : uint64_t(top);
    double s = binom_sf(int64_t(x), n, p);
    while (x > 0) {
        double below = s + binom_pmf(double(x), double(n), p);
        if (below > q)
            return x;
        s = below;
        --x;
    }
    return 0;
}

uint64_t calc_min_count(uint64_t frame_len, double bf_fpr, double min_fpr) {
    /* SeqEval::calcMinCount (Common/SeqEval.h:199-207) */
    unsigned i = unsigned(binom_isf_up(min_fpr, frame_len, bf_fpr));
    return i > 1 ? i : 1;
}

/* ------------------------------------------------------------- summary */

std::string summary_text(const std::vector<std::string> &ids, const uint64_t *sm,
                         uint64_t nrec) {
    /* ResultsManager::getResultsSummary (ResultsManager.hpp:191-237) */
    const size_t nf = ids.size();
    const uint64_t *above = sm, *unique = sm + nf;
    const uint64_t multi = sm[2 * nf], none = sm[2 * nf + 1];
    std::stringstream o;
    o << "filter_id\thits\tmisses\tshared\trate_hit\trate_miss\trate_shared\n";
    for (size_t i = 0; i < nf; ++i) {
        o << ids[i] << "\t" << above[i] << "\t" << nrec - above[i] << "\t"
          << (above[i] - unique[i]) << "\t" << double(above[i]) / double(nrec) << "\t"
          << double(nrec - above[i]) / double(nrec) << "\t"
          << double(above[i] - unique[i]) / double(nrec) << "\n";
    }
    o << "multiMatch\t" << multi << "\t" << nrec - multi << "\t" << 0 << "\t"
      << double(multi) / double(nrec) << "\t" << double(nrec - multi) / double(nrec)
      << "\t" << 0.0 << "\n";
    o << "noMatch\t" << none << "\t" << nrec - none << "\t" << 0 << "\t"
      << double(none) / double(nrec) << "\t" << double(nrec - none) / double(nrec) << "\t"
      << 0.0 << "\n";
    return o.str();
}

static thread_local std::string g_last_error;

int32_t set_error(int32_t code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

const char *last_error() { return g_last_error.c_str(); }

} // namespace bbt
