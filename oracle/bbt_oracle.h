/*
 * bbt_oracle.h -- CPU restatement of the BioBloom Tools read-classification
 * path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the
 * CPU-baseline leg of the benchmark load this library as the checker.  No
 * product code path links, loads or calls it.
 *
 * What it restates (reference files, relative to the BioBloom source tree):
 *   ntHash / ntHashIterator / stHashIterator   btl_bloomfilter (absent
 *       submodule, Makefile.am:11-19); restated from the published ntHash
 *       algorithm (Mohamadi et al., Bioinformatics 32(22), 2016), call sites
 *       Common/SeqEval.h:50, BioBloomCategorizer/MIBFClassifier.hpp:788-793
 *   BloomFilter::contains / insert              btl_bloomfilter (absent);
 *       call sites Common/SeqEval.h:54,79; BloomFilterGenerator.h:170-176
 *   SeqEval scorers                             Common/SeqEval.h:37-564
 *   evaluateRead* modes                         BioBloomClassifier.h:531-571,
 *                                               BioBloomClassifier.cpp:1145-1237
 *   ResultsManager                              ResultsManager.hpp:40-237
 *   filter sizing                               Common/BloomFilterInfo.h:57-76
 *   binomial thresholds                         Common/SeqEval.h:199-216
 *   dust masking (-D)                           Common/sdust.c, Common/SDust.hpp,
 *                                               SeqEval.h:496-529
 *   kseq_read (sequences only)                  Common/kseq.h:93-219 (bbt_kseq.cpp)
 *
 * Parity status of each function is recorded beside it in bbt_oracle.cpp.
 */
#ifndef BBT_ORACLE_H
#define BBT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scorer ids follow opt::ScoringMethod (Common/Options.h:35) */
enum { ORC_SIMPLE = 0, ORC_LENGTH = 1, ORC_HARMONIC = 2, ORC_BINOMIAL = 3 };
/* mode ids follow opt::FilteringMode (BioBloomCategorizer/Options.h:118) */
enum { ORC_STD = 0, ORC_ORDERED = 1, ORC_BESTHIT = 2, ORC_SCORES = 3 };

/* ---- hashing ---------------------------------------------------------- */
/* Rolling iterator over the N-free windows of seq: writes window starts to
 * pos_out and h hashes per window to hash_out (row-major).  Returns the
 * number of windows (or -1 if cap is too small). */
int64_t orc_nthash(const char *seq, uint64_t len, uint32_t k, uint32_t h,
                   uint64_t *pos_out, uint64_t *hash_out, uint64_t cap);
/* Closed form of one window: forward and reverse-complement strand values. */
void orc_nthash_window(const char *win, uint32_t k, uint64_t *fwd, uint64_t *rc);
/* Multi-hash expansion of canonical value h0 into out[0..h). */
void orc_multihash(uint64_t h0, uint32_t k, uint32_t h, uint64_t *out);
/* Spaced-seed iterator; seeds is nseeds strings of '0'/'1' of length k,
 * concatenated.  One canonical hash per seed per window. */
int64_t orc_sthash(const char *seq, uint64_t len, const char *seeds,
                   uint32_t nseeds, uint32_t k, uint64_t *pos_out,
                   uint64_t *hash_out, uint64_t cap);

/* ---- Bloom filter ---------------------------------------------------- */
int orc_bf_contains(const uint8_t *bits, uint64_t m, uint32_t h,
                    const uint64_t *hashes, uint32_t *lookups);
/* Returns 1 if every bit was already set (insertAndCheck). */
int orc_bf_insert(uint8_t *bits, uint64_t m, uint32_t h, const uint64_t *hashes);
/* Inserts every k-mer of one sequence; accumulates new/redundant counts. */
void orc_bf_insert_seq(uint8_t *bits, uint64_t m, uint32_t h, uint32_t k,
                       const char *seq, uint64_t len, uint64_t *n_new,
                       uint64_t *n_redundant);
/* loadFilterSubtract: as orc_bf_insert_seq, skipping (and counting in
 * *n_removed) every k-mer the filter sub_bits/sub_m (same h, k) contains */
void orc_bf_insert_seq_subtract(uint8_t *bits, uint64_t m, uint32_t h, uint32_t k,
                                const uint8_t *sub_bits, uint64_t sub_m, const char *seq,
                                uint64_t len, uint64_t *n_new, uint64_t *n_redundant,
                                uint64_t *n_removed);
uint64_t orc_popcount(const uint8_t *bits, uint64_t nbytes);
uint64_t orc_calc_optimal_size(uint64_t entries, double fpr, uint32_t h);
uint32_t orc_calc_optimal_hashnum(double fpr);

/* ---- statistics -------------------------------------------------------- */
double orc_binom_pmf(uint64_t x, uint64_t n, double p);
double orc_binom_sf(int64_t x, uint64_t n, double p);      /* P(X > x) */
uint64_t orc_binom_isf_up(double q, uint64_t n, double p); /* min x: sf(x)<=q */
uint64_t orc_calc_min_count(uint64_t frame_len, double bf_fpr, double min_fpr);

/* ---- single (read, filter) evaluation --------------------------------- */
typedef struct {
    const uint8_t *bits;
    uint64_t m;
    uint32_t h;
    uint32_t k;
    double fpr; /* getFPRPrecompute(): (popcount/m)^h */
} orc_filter;

typedef struct {
    uint32_t scorer;   /* ORC_SIMPLE .. */
    uint32_t streak;   /* opt::streakThreshold */
    double threshold;  /* opt::score after CLI interpretation */
    uint32_t dust;     /* opt::dust (-D) */
    uint32_t dust_t;   /* opt::dustT (-T) */
    uint32_t dust_w;   /* opt::dustWindow (-W) */
    uint32_t pad;
} orc_eval_params;

/* evalRead with a subtract filter (SeqEval.h:37-114, 295-332): a hit that sub
 * also contains scores nothing; SIMPLE and LENGTH scorers. */
int orc_eval_read_sub(const orc_filter *f, const orc_filter *sub, const char *seq,
                      uint64_t len, const orc_eval_params *p);

/* Progressive filter growth over read pairs (reads 2i, 2i+1 of seq/off), or
 * over single reads (npairs = reads) when p->single. */
typedef struct {
    double score;        /* -r: the SIMPLE score, or the minimum match length */
    double bait;         /* -b: the bait filter's threshold */
    uint32_t scorer;     /* ORC_SIMPLE or ORC_LENGTH */
    uint32_t inclusive;  /* -i: PROG_INC (either read), else PROG_STD (both) */
    uint32_t streak;     /* opt::streakThreshold */
    uint32_t iterations; /* -e */
    uint32_t single;     /* reads one by one (the single-end file-list builder) */
    uint32_t pad;
    uint64_t expected;   /* the k-mers after which growth stops */
} orc_prog_params;
int orc_progressive(uint8_t *bits, uint64_t m, uint32_t h, uint32_t k, const orc_filter *bait,
                    const orc_filter *sub, const char *seq, const uint64_t *off,
                    uint64_t npairs, const orc_prog_params *p, uint64_t *total_entries,
                    uint64_t *total_reads, uint64_t *tagged);

int orc_eval_read(const orc_filter *f, const char *seq, uint64_t len,
                  const orc_eval_params *p, uint64_t *probes, uint64_t *lookups);
double orc_eval_score(const orc_filter *f, const char *seq, uint64_t len,
                      const orc_eval_params *p, uint64_t *probes,
                      uint64_t *lookups);

/* ---- batch classification (the CPU baseline) ------------------------- */
/* Reads are seq[off[i]..off[i+1]).  With paired != 0, reads 2j and 2j+1 are
 * mates and one record is one pair.  Outputs (any may be NULL):
 *   hits[read*nf + f]   1 if filter f is in the read's hit list
 *   scores[read*nf + f] evalScore value (BESTHIT / SCORES modes)
 *   cls[record]         ResultsManager class index (n = noMatch, n+1 = multi)
 *   counters[0..1]      contains() calls, bit lookups
 * Returns 0 on success. */
int orc_classify(const orc_filter *filters, uint32_t nf, const char *seq,
                 const uint64_t *off, uint64_t nreads, int paired, int inclusive,
                 uint32_t mode, const orc_eval_params *p, int nthreads,
                 uint8_t *hits, double *scores, uint32_t *cls,
                 uint64_t *counters);

/* ResultsManager summary text for a classified batch; writes at most cap
 * bytes (NUL-terminated) and returns the full length. */
uint64_t orc_summary(const char *const *filter_ids, uint32_t nf,
                     const uint8_t *hits, uint64_t nreads, int paired,
                     int inclusive, char *out, uint64_t cap);

/* ---- dust (bbt_dust.cpp) ------------------------------------------------ */
/* sdust(seq, len, T, W): merged low-complexity base intervals [start, finish);
 * returns their number, or -1 if cap is too small. */
int64_t orc_sdust(const char *seq, uint64_t len, int T, int W, uint32_t *starts,
                  uint32_t *finishes, uint64_t cap);

/* ---- kseq_read (bbt_kseq.cpp) ------------------------------------------- */
/* Every record's sequence of the FASTA/FASTQ text buf[0..len) appended to
 * seq_out (cap bytes) with its end in off_out[1..] (off_out[0] = 0, at most
 * cap_recs - 1 records); *status = kseq_read's return after the last record
 * (-1 end of input, -2 bad quality).  Returns the records, or -1 if a
 * capacity is too small. */
int64_t orc_kseq_seqs(const char *buf, uint64_t len, char *seq_out, uint64_t cap,
                      uint64_t *off_out, uint64_t cap_recs, int64_t *status);

/* ---- multi-index Bloom filter (bbt_mibf.cpp) ------------------------- */
/* seeds: nseeds concatenated '0'/'1' strings of length span (spaced-seed
 * mode), or NULL for k-mer mode (nseeds ntHash values per span-mer). */
typedef struct {
    uint32_t id, count, nonsat, total, total_nonsat, solid, nonsat_frame;
} orc_mibf_hit;

uint64_t orc_mibf_calc_size(uint64_t entries, uint32_t nseeds, double occupancy);
void orc_mibf_set_bits(const char *seq, uint64_t len, const char *seeds, uint32_t nseeds,
                       uint32_t span, uint64_t m, uint64_t *bv);
uint64_t orc_mibf_popcount(const uint64_t *bv, uint64_t m);
void orc_mibf_assign(const char *seq, const uint64_t *rec_off, const uint32_t *rec_id,
                     uint64_t nrec, const char *seeds, uint32_t nseeds, uint32_t span,
                     uint64_t m, const uint64_t *bv, uint32_t *data);
void orc_mibf_frame_probs(const uint32_t *data, uint64_t pop, uint64_t m, uint32_t nseeds,
                          uint32_t allowed_miss, uint32_t nids, double *probs);
int orc_mibf_query(const uint64_t *bv, const uint32_t *data, uint64_t m, const char *seeds,
                   uint32_t nseeds, uint32_t span, const double *probs, uint32_t nids,
                   double score, uint32_t allowed_miss, double multi_thresh,
                   uint32_t min_nonsat, int paired, int best_agree, const char *seq,
                   const uint64_t *off,
                   uint64_t nreads, uint32_t max_hits, orc_mibf_hit *hits,
                   uint32_t *nhits, uint32_t *satcount, uint64_t *counters);

#ifdef __cplusplus
}
#endif
#endif
