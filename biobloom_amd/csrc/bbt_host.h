/*
 * bbt_host.h -- host-side pieces of libbbt_hip that touch no device:
 * filter file formats, sizing, binomial thresholds, summary text.
 */
#ifndef BBT_HOST_H
#define BBT_HOST_H

#include <stdint.h>

#include <string>
#include <vector>

namespace bbt {

/* .txt INI written by BloomFilterInfo::printInfoFile
 * (Common/BloomFilterInfo.cpp:81-116) and read by its file constructor
 * (:31-53).  Keys keep the reference spelling, including
 * "desired_false_positve_rate". */
struct InfoFile {
    std::string filter_id;
    unsigned kmer_size = 0;
    double desired_fpr = 0;
    unsigned hash_num = 0;
    uint64_t expected_entries = 0;
    std::string sequence_sources;
    uint64_t size = 0;
    uint64_t num_entries = 0;
    double approx_fpr = 0;
    uint64_t redundant_sequences = 0;
    double redundant_fpr = 0;
};

bool read_info_file(const std::string &path, InfoFile &out, std::string &err);
bool write_info_file(const std::string &path, const InfoFile &in, std::string &err);

/* .bf: fixed binary header followed by the m/8-byte bit array. */
struct BfHeader {
    char magic[8];     /* "BlOOMFXX" */
    uint32_t hlen;     /* sizeof(BfHeader) */
    uint32_t pad;
    uint64_t size;     /* m, bits */
    uint32_t nhash;
    uint32_t kmer;
    double dFPR;       /* desired FPR */
    uint64_t nEntry;   /* inserted k-mers */
    uint64_t tEntry;   /* expected k-mers */
};
static_assert(sizeof(BfHeader) == 56, "BfHeader layout");

bool read_bf_file(const std::string &path, BfHeader &hdr, std::vector<uint8_t> &bits,
                  std::string &err);
bool write_bf_file(const std::string &path, const BfHeader &hdr, const uint8_t *bits,
                   uint64_t nbytes, std::string &err);

uint64_t calc_optimal_size(uint64_t entries, double fpr, unsigned h);
unsigned calc_optimal_hashnum(double fpr);
double calc_approx_fpr(uint64_t size, uint64_t entries, unsigned h);
double calc_redundancy_fpr(uint64_t size, uint64_t entries, unsigned h);

/* Binomial(n, p): P(X > x), and the smallest x with P(X > x) <= q. */
double binom_sf(int64_t x, uint64_t n, double p);
uint64_t binom_isf_up(double q, uint64_t n, double p);
uint64_t calc_min_count(uint64_t frame_len, double bf_fpr, double min_fpr);

std::string summary_text(const std::vector<std::string> &ids, const uint64_t *summary,
                         uint64_t nrec);

/* Message of the last failure on the calling thread (bbt_last_error). */
int32_t set_error(int32_t code, const std::string &msg);
const char *last_error();

} // namespace bbt

#endif
