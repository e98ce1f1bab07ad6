/*
 * bbt_hip.h -- C ABI of the MI355X read-classification library (libbbt_hip).
 *
 * Plain C types and pointers only.  Device buffers are HIP device pointers;
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 * Every entry point returns a bbt_status and never exits or throws; the
 * message of the last failure on the calling thread is bbt_last_error().
 *
 * Reference interfaces each entry point replaces (BioBloom Tools 2.3.4):
 *   bbt_filters_load   BioBloomClassifier::loadFilters
 *                      (BioBloomCategorizer/BioBloomClassifier.cpp:950-971):
 *                      <x>.txt via BloomFilterInfo(Common/BloomFilterInfo.cpp:31-53)
 *                      and new BloomFilter(<x>.bf) per path, in order
 *   bbt_eval_batch     per read, BioBloomClassifier::evaluateRead /
 *                      evaluateReadPair (BioBloomClassifier.h:531-571), which
 *                      call SeqEval::evalRead / evalScore (Common/SeqEval.h:
 *                      493-564) per filter, followed by
 *                      ResultsManager::updateSummaryData
 *                      (BioBloomCategorizer/ResultsManager.hpp:66-189)
 *   bbt_summary_text   ResultsManager::getResultsSummary (ResultsManager.hpp:
 *                      191-237), the text of <prefix>_summary.tsv
 *   bbt_filter_insert  BloomFilterGenerator::loadFilter
 *                      (BioBloomMaker/BloomFilterGenerator.h:143-188)
 *   bbt_eval_subtract  SeqEval::evalRead with a subtract filter (SeqEval.h:
 *                      37-114, 295-332), the per-read test of the progressive
 *                      builders (BloomFilterGenerator.cpp:85-1170)
 *   bbt_filter_insert_subtract
 *                      BloomFilterGenerator::loadFilterSubtract
 *                      (BloomFilterGenerator.h:217-284), the maker's -s mode
 *                      (BloomFilterGenerator.cpp:1175-1199)
 *   bbt_filter_store   BloomFilter::storeFilter + BloomFilterInfo::printInfoFile
 *                      (BloomFilterGenerator.cpp:71, BloomFilterInfo.cpp:81-116)
 *   bbt_calc_optimal_size / bbt_calc_optimal_hashnum
 *                      BloomFilterInfo::calcOptimalSize / calcOptimalHashNum
 *                      (Common/BloomFilterInfo.h:57-76)
 *   bbt_mibf_build     MIBFGen::generate (BioBloomMaker/MIBFGen.hpp:131-309):
 *                      bit-vector pass, ID pass, saturation pass
 *   bbt_mibf_store / bbt_mibf_load
 *                      MIBloomFilter::store + MIBFGen::writeIDs
 *                      (MIBFGen.hpp:284-292, 424-430); MIBFClassifier
 *                      constructor (BioBloomCategorizer/MIBFClassifier.hpp:
 *                      49-95), which reads <prefix>.bf and <prefix>_ids.txt
 *   bbt_mibf_frame_probs
 *                      MIBloomFilter::calcFrameProbs (MIBFClassifier.hpp:86-87)
 *   bbt_mibf_min_count MIBFClassifier::getMinCount (MIBFClassifier.hpp:930-939)
 *   bbt_mibf_query_batch
 *                      per read, MIBFClassifier::classify (MIBFClassifier.hpp:
 *                      773-796): MIBFQuerySupport::query over the read's
 *                      spaced-seed frames, best hit first
 */
#ifndef BBT_HIP_H
#define BBT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t bbt_status;
#define BBT_OK 0
#define BBT_ERR_ARG 1         /* invalid argument */
#define BBT_ERR_IO 2          /* file cannot be opened / read / written */
#define BBT_ERR_FORMAT 3      /* malformed .bf / .txt */
#define BBT_ERR_HIP 4         /* HIP runtime error */
#define BBT_ERR_NOMEM 5       /* device allocation failed */
#define BBT_ERR_UNSUPPORTED 6 /* combination not available on the device path */

/* scorer ids: opt::ScoringMethod (Common/Options.h:35) */
#define BBT_SIMPLE 0
#define BBT_LENGTH 1
#define BBT_HARMONIC 2
#define BBT_BINOMIAL 3
/* mode ids: opt::FilteringMode (BioBloomCategorizer/Options.h:118-120) */
#define BBT_STD 0
#define BBT_ORDERED 1
#define BBT_BESTHIT 2
#define BBT_SCORES 3
/* execution paths */
#define BBT_PATH_PLAIN 0   /* one kernel launch per batch, one lane per walk */
#define BBT_PATH_BATCHED 1 /* persistent work-queue kernel, graph-replayed */

/* flags */
#define BBT_FLAG_TIME_WALK 1u /* time the walk kernel(s) with HIP events on the
                                 stream; the call then waits for them and
                                 bbt_last_walk_ms() returns the duration */
#define BBT_FLAG_GRAPH 2u     /* batch as one captured hipGraph: captured on the
                                 first call with a given set of buffers and
                                 parameters, replayed by later identical calls */

typedef struct bbt_filterset bbt_filterset;

/* Scheduling knobs of the batched path; 0 = library default.  They change
 * only how work is spread over the GPU, never a result (nor the reference
 * lookup counters; probe_ahead adds to the issued-lookup counter only). */
typedef struct {
    uint32_t wave_min_len;  /* reads per wavefront: reads this long are walked
                               by a whole wavefront, shorter ones 64 per wave */
    uint32_t refill_min;    /* idle lanes a wave gathers before refilling */
    uint32_t blocks_per_cu; /* persistent blocks (256 threads) per CU */
    uint32_t probe_ahead;   /* load scheduling of the lane walks, four bits:
                               bit 0 (probe unroll): a contains() call's
                               lookups are issued 4 at a time (independent
                               loads; those past the first clear bit are
                               wasted), else one at a time; bit 1
                               (lookahead): the next window's first lookups
                               are issued before the pending call's bits are
                               taken (wasted when that window is not probed);
                               bit 2 (interleave): lane walks go read by
                               read, and each group of 2 to 64 filters of one
                               m, h and k is read through a copy of its
                               filters interleaved byte by byte, so the lanes
                               of one read fetch one line per lookup for all
                               its filters of a group (the copies are made per
                               batch; bbt_eval_batch only, ignored by ORDERED
                               mode and for sets without such a group); bit 3
                               with bit 2 (read per lane): a group of 3 to 16
                               filters of one geometry (the set's only
                               group), SIMPLE evalRead in STD mode, lane-walked
                               reads shorter than 32768: each lane hashes one
                               read once and reads all the group's filters
                               with one load per position; with reads of
                               wave_min_len or more (at most 32768) and a group
                               of at most 8, each long read is walked by one
                               wave against all of them; the set's other
                               filters are walked over their own arrays.
                               Values above 15 are rejected.  Without knobs
                               of its own (or in the tuning table) a set whose
                               filters all share one geometry takes 7, and 15
                               for a batch of 65536 reads or more where bit 3
                               applies; a set with one such group of 3 to 16
                               filters and others takes 15 for such a batch
                               (the group read per lane, the others over their
                               own arrays). */
} bbt_tuning;

typedef struct {
    uint64_t m;         /* filter bits */
    uint32_t h;         /* hash functions */
    uint32_t k;         /* k-mer size */
    double fpr;         /* (popcount / m)^h  (getFPRPrecompute) */
    uint64_t popcount;  /* set bits */
    double desired_fpr; /* .txt desired_false_positve_rate (0 if unknown) */
} bbt_filter_info;

typedef struct {
    uint32_t mode;      /* BBT_STD .. BBT_SCORES */
    uint32_t scorer;    /* BBT_SIMPLE .. BBT_BINOMIAL */
    double threshold;   /* opt::score after the categorizer's interpretation
                           (BioBloomCategorizer.cpp:416-441) */
    uint32_t streak;    /* opt::streakThreshold (-r), default 3 */
    uint32_t paired;    /* reads 2j and 2j+1 are mates; one record per pair */
    uint32_t inclusive; /* -i: pair hit if either mate hits */
    uint32_t path;      /* BBT_PATH_PLAIN or BBT_PATH_BATCHED */
    uint32_t flags;     /* BBT_FLAG_* */
    uint32_t wave_min_len; /* batched path: reads at least this long are walked
                              by a whole wavefront, 64 windows per step
                              (0 = library default) */
    uint32_t dust;      /* -D: windows starting in an sdust low-complexity interval
                           are misses without a contains() call (SeqEval.h:496) */
    uint32_t dust_t;    /* -T, sdust threshold [20] */
    uint32_t dust_w;    /* -W, sdust window [64], 3..4096 */
    uint32_t reserved;
} bbt_eval_params;

/* read formats (bbt_reads.format) */
#define BBT_READS_ASCII 0 /* one byte per base, as read from FASTA/FASTQ */
#define BBT_READS_2BIT 1  /* 2-bit codes and an N mask, 3/8 byte per base:
                             base j of the batch is bits 2(j%4)..2(j%4)+1 of
                             byte j/4 (A 0, C 1, T 2, G 3, i.e. (ASCII >> 1) & 3),
                             and bit j%8 of byte seq_bytes + j/8 is set when
                             base j is not A/C/G/T in either case (its code is
                             then ignored).  Case is not kept: no kernel
                             depends on it.  Unpacked on the device. */

typedef struct {
    const char *seq;       /* device, 8-byte aligned: bases of all reads back to
                              back, in format */
    uint64_t seq_bytes;    /* ASCII: readable bytes at seq, >= off[nreads]
                              rounded up to a multiple of 8.  2BIT: bytes of
                              codes, a multiple of 8 with 4 x seq_bytes >=
                              off[nreads]; the mask (seq_bytes / 2 bytes)
                              follows them at seq + seq_bytes */
    const uint64_t *off;   /* device: nreads + 1 offsets (in bases) */
    uint64_t nreads;
    uint32_t max_len;      /* upper bound on read length */
    uint32_t format;       /* BBT_READS_ASCII or BBT_READS_2BIT */
    const uint32_t *order; /* device, optional: processing order of reads
                              (a permutation of 0..nreads-1), or NULL */
} bbt_reads;

typedef struct {
    uint8_t *hits;      /* device: nreads * nf bytes, 1 = filter in hit list */
    double *scores;     /* device: nreads * nf evalScore values, or NULL;
                           required for BBT_BESTHIT and BBT_SCORES */
    uint32_t *cls;      /* device: per record class index (filter id,
                           nf = noMatch, nf+1 = multiMatch), or NULL */
    uint64_t *summary;  /* device: 2*nf + 2 counters ADDED TO: hits per filter,
                           unique hits per filter, multiMatch, noMatch; or NULL */
    uint64_t *counters; /* device: 4 counters ADDED TO: contains() calls and bit
                           lookups the reference walk makes, bit lookups
                           issued (>= the former when whole waves probe ahead),
                           error flag (1: a BINOMIAL read longer than
                           max_len); or NULL */
} bbt_eval_out;

const char *bbt_last_error(void);

/* ---- filters ------------------------------------------------------------ */
bbt_status bbt_filters_load(const char *const *bf_paths, uint32_t n,
                            bbt_filterset **out);
bbt_status bbt_filters_create(uint32_t n, const uint64_t *m, const uint32_t *h,
                              const uint32_t *k, bbt_filterset **out);
bbt_status bbt_filters_free(bbt_filterset *fs);
uint32_t bbt_filterset_size(const bbt_filterset *fs);
bbt_status bbt_filter_info_get(const bbt_filterset *fs, uint32_t idx,
                               bbt_filter_info *out);
const char *bbt_filter_id(const bbt_filterset *fs, uint32_t idx);
bbt_status bbt_filter_set_id(bbt_filterset *fs, uint32_t idx, const char *id);
/* device bit array of filter idx (read-only use by callers) */
const void *bbt_filter_device_bits(const bbt_filterset *fs, uint32_t idx);
bbt_status bbt_filter_upload(bbt_filterset *fs, uint32_t idx, const void *host_bits,
                             uint64_t nbytes);
bbt_status bbt_filter_download(const bbt_filterset *fs, uint32_t idx, void *host_bits,
                               uint64_t nbytes);
/* Insert every k-mer of the device sequence buffer dseq[0..nbytes) into filter
 * idx on the device; records are separated by any non-ACGT byte.  counts
 * (host, optional) receives {k-mers newly set, k-mers already present}. */
bbt_status bbt_filter_insert(bbt_filterset *fs, uint32_t idx, const char *dseq,
                             uint64_t nbytes, uint64_t *counts, void *stream);
/* BloomFilterGenerator::loadFilter(bf, str) (BloomFilterGenerator.h:193-202)
 * for one read: as bbt_filter_insert, with the k-mers taken in sequence order
 * by one device thread, so the {new, present} split is exactly the
 * sequential count the progressive builders add up. */
bbt_status bbt_filter_insert_ordered(bbt_filterset *fs, uint32_t idx, const char *dseq,
                                     uint64_t nbytes, uint64_t *counts, void *stream);
/* As bbt_filter_insert, but a k-mer that filter sub_idx of sub contains is
 * not inserted.  The two filters must share h and k (BBT_ERR_ARG with the
 * reference's message otherwise).  counts (host, optional) receives {k-mers
 * newly set, k-mers already present, k-mers removed by the subtract filter}. */
/* pass[r] = SeqEval::evalRead(read r, filter idx, threshold, sub) for every
 * read of a device batch, with scorer BBT_SIMPLE (threshold = the score, streak
 * = opt::streakThreshold) or BBT_LENGTH (threshold rounded = the minimum match
 * length).  A k-mer both filters contain is a hit that scores nothing.  sub may
 * be null; otherwise it must share h and k.  Asynchronous on stream. */
bbt_status bbt_eval_subtract(const bbt_filterset *fs, uint32_t idx, const bbt_filterset *sub,
                             uint32_t sub_idx, const bbt_reads *in, uint32_t scorer,
                             double threshold, uint32_t streak, uint8_t *pass, void *stream);
bbt_status bbt_filter_insert_subtract(bbt_filterset *fs, uint32_t idx,
                                      const bbt_filterset *sub, uint32_t sub_idx,
                                      const char *dseq, uint64_t nbytes, uint64_t *counts,
                                      void *stream);
/* Recomputes popcount and the precomputed FPR after uploads / inserts. */
bbt_status bbt_filter_refresh(bbt_filterset *fs, uint32_t idx, void *stream);
/* Writes <path_prefix>.bf and <path_prefix>.txt. */
bbt_status bbt_filter_store(const bbt_filterset *fs, uint32_t idx,
                            const char *path_prefix, double desired_fpr,
                            uint64_t expected_entries, uint64_t num_entries,
                            uint64_t redundant, const char *seq_sources);
uint64_t bbt_calc_optimal_size(uint64_t entries, double fpr, uint32_t h);
uint32_t bbt_calc_optimal_hashnum(double fpr);

/* ---- classification ----------------------------------------------------- */
bbt_status bbt_eval_batch(bbt_filterset *fs, const bbt_reads *in,
                          const bbt_eval_params *p, bbt_eval_out *out,
                          void *stream);
/* Work queues and replay state for one stream of batches.  bbt_eval_batch
 * uses the filter set's own; callers running batches on several streams at
 * once give each stream a workspace of its own (no counterpart in the
 * reference, whose filters are shared read-only by OpenMP threads,
 * BioBloomClassifier.cpp:59). */
typedef struct bbt_workspace bbt_workspace;
bbt_status bbt_workspace_create(bbt_workspace **out);
bbt_status bbt_workspace_free(bbt_workspace *ws);
bbt_status bbt_eval_batch_ws(bbt_filterset *fs, bbt_workspace *ws, const bbt_reads *in,
                             const bbt_eval_params *p, bbt_eval_out *out, void *stream);
/* bbt_eval_batch without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * outputs and results as bbt_eval_batch; the batch's control words and
 * read lists live in a library-owned device array of 8 MiB + 1 KiB per GPU,
 * held until the process ends, instead of the filter set's workspace.
 * Limits: at most 1048576 reads per call (BBT_ERR_ARG beyond); -D (dust),
 * BBT_FLAG_GRAPH and BBT_FLAG_TIME_WALK are rejected (BBT_ERR_UNSUPPORTED);
 * BINOMIAL scoring needs its tables built by an earlier bbt_eval_batch with
 * the same threshold and no longer reads (BBT_ERR_UNSUPPORTED otherwise).
 * Every call on one GPU shares that GPU's array: calls whose kernels
 * overlap on one GPU (other streams, other threads, other filter sets) race
 * and give wrong results; order them (one stream, or events). */
bbt_status bbt_eval_batch_noalloc(bbt_filterset *fs, const bbt_reads *in,
                                  const bbt_eval_params *p, bbt_eval_out *out, void *stream);
/* Scheduling knobs of a filter set (no counterpart in the reference). */
bbt_status bbt_set_tuning(bbt_filterset *fs, const bbt_tuning *t);
bbt_status bbt_get_tuning(const bbt_filterset *fs, bbt_tuning *t);
/* Times candidate knob settings of the batched path on this batch (each
 * candidate runs the full batch; outputs are overwritten, counters added
 * to), keeps the fastest in fs and returns it in *best.  The choice is also
 * kept in the process's tuning table under the batch's geometry: filter-set
 * size (log2 of its bytes), filter count, k, h, longest read (log2), mode,
 * scorer and pairing.  A batch on a filter set without knobs of its own
 * (bbt_set_tuning, bbt_autotune) takes the table's entry for its geometry. */
bbt_status bbt_autotune(bbt_filterset *fs, const bbt_reads *in, const bbt_eval_params *p,
                        bbt_eval_out *out, void *stream, bbt_tuning *best);
/* The tuning table as text (one geometry and its knobs per line), so a
 * later process reuses the tuned knobs: save writes every entry, load adds
 * the file's entries (BBT_ERR_IO if it cannot be read, BBT_ERR_FORMAT for a
 * malformed line).  No counterpart in the reference. */
bbt_status bbt_tuning_save(const char *path);
bbt_status bbt_tuning_load(const char *path);
/* Knobs the table holds for this batch's geometry on fs: BBT_OK and *t, or
 * BBT_ERR_ARG when there is no entry. */
bbt_status bbt_tuning_lookup(const bbt_filterset *fs, const bbt_reads *in,
                             const bbt_eval_params *p, bbt_tuning *t);
/* Measurement utility (no counterpart in the reference): rate of
 * independent random byte loads over a fresh nbytes device buffer (power of
 * two), 8 in flight per lane, rounds x 8 loads per lane; each load touches
 * one random cache line. */
bbt_status bbt_measure_gather(uint64_t nbytes, uint32_t rounds, void *stream,
                              double *lines_per_s, double *ms);
/* Duration of the walk kernel(s) of the last bbt_eval_batch on fs that was
 * called with BBT_FLAG_TIME_WALK, in milliseconds (-1 if none). */
double bbt_last_walk_ms(const bbt_filterset *fs);
/* Measurement utility (no counterpart in the reference): name of the walk
 * kernel(s) the last bbt_eval_batch or bbt_eval_batch_noalloc on fs ran, in
 * launch order joined by '+' ("k_walk_persistent", "k_walk_rows",
 * "k_walk_group_long+k_walk_rows", "k_walk_plain", "k_ordered", ...; "" before
 * the first batch), for profiles that name the kernels they time. */
const char *bbt_last_walk_kernel(const bbt_filterset *fs);
/* summary.tsv text from the host copy of bbt_eval_out.summary counters;
 * writes at most cap bytes (NUL-terminated), *needed = full length. */
bbt_status bbt_summary_text(const bbt_filterset *fs, const uint64_t *summary,
                            uint64_t nrec, char *buf, uint64_t cap,
                            uint64_t *needed);
/* Binomial helpers used by the BINOMIAL scorer (SeqEval.h:199-216). */
uint64_t bbt_calc_min_count(uint64_t frame_len, double bf_fpr, double min_fpr);
double bbt_binomial_score(uint64_t count, uint64_t frame_len, double bf_fpr);

/* ---- serving ------------------------------------------------------------ */
/* A request loop over a resident filter set (no counterpart in the reference,
 * which classifies one input stream per process: BioBloomClassifier::filter,
 * BioBloomClassifier.cpp:33-90).  Requests (host reads) are queued FIFO and
 * admitted into launches while a slot is free and their reads and outputs fit
 * in a device arena of device_bytes; requests arriving while launches run
 * join the next launch (continuous batching), up to `slots` launches in
 * flight on their own streams.  BBT_SERVE_PLAIN: one request per launch with
 * the plain kernel, in order (the per-request baseline).  Results are the
 * per-read hits and per-record classes of bbt_eval_batch. */
typedef struct bbt_server bbt_server;
enum { BBT_SERVE_BATCHED = 0, BBT_SERVE_PLAIN = 1 };

typedef struct {
    uint32_t slots;          /* launches in flight (1..8; 1 in plain mode) */
    uint32_t mode;           /* BBT_SERVE_* */
    uint64_t device_bytes;   /* arena for the reads and outputs in flight */
    uint64_t max_wave_reads; /* cap on reads merged into one launch */
    uint64_t max_wave_bases; /* cap on bases per launch and per request */
    uint64_t host_bytes;     /* pinned host arena holding the bases of queued and
                                in-flight requests; submit waits while it is full
                                (0 = (slots + 2) x max_wave_bases) */
} bbt_server_config;

typedef struct {
    uint64_t nreads;
    uint64_t wave;           /* launch that carried the request (1-based) */
    double submit_us, admit_us, done_us; /* since server creation */
} bbt_request_stats;

typedef struct {
    uint64_t waves, requests_done, reads_done, max_waves_in_flight;
    uint64_t arena_bytes, arena_high_water;
    double mean_wave_reads, mean_wave_requests;
    double submit_ms;  /* host time spent copying requests in bbt_server_submit */
    double pack_ms;    /* host time packing launches into pinned staging */
    double device_ms;  /* device time of the launches: copies in, kernels, copies out */
    uint64_t probes;   /* contains() calls of the reference walk, over all launches */
    uint64_t lookups;  /* their bit lookups (one random 64 B line each) */
    uint64_t lookups_issued; /* bit lookups the kernels made (>= lookups) */
} bbt_server_stats;

/* p: the classification parameters of every request, any mode (path and
 * flags are set by the server; BBT_BESTHIT and BBT_SCORES also return each
 * read's evalScore values, bbt_server_scores).  fs must outlive the server. */
bbt_status bbt_server_create(bbt_filterset *fs, const bbt_eval_params *p,
                             const bbt_server_config *cfg, bbt_server **out);
/* Copies the request (host seq, nreads + 1 offsets into it) and queues it. */
bbt_status bbt_server_submit(bbt_server *s, const char *seq, const uint64_t *off,
                             uint64_t nreads, uint64_t *req_id);
/* Blocks until the request is done (timeout_s < 0: no limit). */
bbt_status bbt_server_wait(bbt_server *s, uint64_t req_id, double timeout_s);
/* Blocks until every queued request is done. */
bbt_status bbt_server_drain(bbt_server *s, double timeout_s);
/* hits: nreads * nf bytes, cls: per record; either may be NULL. */
bbt_status bbt_server_result(bbt_server *s, uint64_t req_id, uint8_t *hits, uint32_t *cls,
                             bbt_request_stats *stats);
/* BBT_BESTHIT / BBT_SCORES servers: the request's nreads * nf evalScore
 * values (the scores output of bbt_eval_batch). */
bbt_status bbt_server_scores(bbt_server *s, uint64_t req_id, double *scores);
bbt_status bbt_server_release(bbt_server *s, uint64_t req_id);
bbt_status bbt_server_get_stats(bbt_server *s, bbt_server_stats *out);
/* Stops admitting, finishes the launches in flight, frees everything. */
bbt_status bbt_server_destroy(bbt_server *s);

/* ---- multi-index Bloom filter ------------------------------------------- */
/* One bit vector of m bits shared by all references; each set bit owns one
 * 32-bit slot (rank order) holding a reference ID (bits 0..30, IDs 1..n) and
 * a saturation flag (bit 31).  Hashes are spaced-seed ntHash values, one per
 * seed per window; a seed is a string of '0'/'1' ('1' = position used).
 * Device layout: the bit vector in 64-byte lines (a rank sample and seven
 * words) and the rank-ordered ID array; and, when m x 4 bytes fit in half of
 * the free device memory (BBT_MIBF_DIRECT=0 in the environment turns it off),
 * a direct-mapped table of one 32-bit word per bit position (0 = clear bit,
 * else the slot's ID word), which makes each query lookup one random load
 * instead of a line load and a dependent ID load.  Results are the same. */
typedef struct bbt_mibf bbt_mibf;

typedef struct {
    uint64_t m;             /* bits */
    uint64_t pop;           /* set bits = ID slots */
    uint64_t pop_saturated; /* slots with the saturation flag */
    uint32_t nids;
    uint32_t nseeds;
    uint32_t span;          /* seed length / k-mer size (window size) */
    uint32_t kmer_mode;     /* 1: nseeds ntHash values per k-mer, no seeds */
    uint32_t direct;        /* 1: queries read the direct-mapped table (above) */
    uint32_t reserved;
} bbt_mibf_info_t;

#define BBT_MIBF_PAIRED 1     /* reads 2j, 2j+1 are mates: one record per pair */
#define BBT_MIBF_BEST_AGREE 2 /* -b: drop results whose best hit is ambiguous */
#define BBT_MIBF_LINES 4      /* look up through the lines and the ID array even
                                 when the direct-mapped table exists (same
                                 results; for comparison) */

typedef struct {
    double score;           /* opt::score: significance spread over the IDs */
    double multi_thresh;    /* report candidates with nonSat >= this x best's */
    uint32_t allowed_miss;  /* seeds of a frame that may miss (S - frameMatches) */
    uint32_t min_nonsat;    /* opt::minCountNonSatCount */
    uint32_t max_hits;      /* hit records per record in bbt_mibf_out.hits (1..8192) */
    uint32_t flags;         /* BBT_MIBF_* */
} bbt_mibf_params;

typedef struct {
    uint32_t *nhits;    /* device, per record (read or pair): reported candidates
                           (may exceed max_hits) */
    uint32_t *hits;     /* device, per record max_hits records of 7 words in rank
                           order: id, count, nonSat, total, totalNonSat, solid,
                           nonSatFrame */
    uint32_t *sat;      /* device, per record: frames whose hit slots are all
                           saturated (satCount) */
    uint64_t *counters; /* device, 4 counters ADDED TO: bit lookups, ID slot
                           reads, candidates beyond max_hits, records with
                           more than 64 candidates (ranked by a second pass
                           that holds every candidate) */
} bbt_mibf_out;

/* Builds on the device from reference records packed back to back in dseq,
 * hashing with nseeds spaced seeds, or, with seeds = NULL, in k-mer mode with
 * nseeds ntHash values per kmer_size-mer (biobloommimaker without -S);
 * records: record r at [rec_off[r], rec_off[r+1]) ENDING WITH one non-ACGT separator
 * byte (rejected otherwise).  drec_off / drec_id are device copies of the
 * nrec + 1 offsets and of the record IDs (1..nids: one per record, or one per
 * file); rec_off is the host copy.  names: nids strings (ID i = names[i-1]).
 * entries = 0 counts len - span + 1 per record as MIBFGen does; m_override = 0
 * sizes the vector for the occupancy. */
bbt_status bbt_mibf_build(const char *dseq, const uint64_t *drec_off, const uint32_t *drec_id,
                          const uint64_t *rec_off, uint64_t nrec, const char *const *names,
                          uint32_t nids, const char *const *seeds, uint32_t nseeds,
                          uint32_t kmer_size, double occupancy, uint64_t entries,
                          uint64_t m_override, bbt_mibf **out, void *stream);
/* Writes <prefix>.bf and <prefix>_ids.txt. */
bbt_status bbt_mibf_store(const bbt_mibf *mf, const char *prefix);
/* Reads <x>.bf and the ID file <x>_ids.txt beside it. */
bbt_status bbt_mibf_load(const char *path, bbt_mibf **out, void *stream);
bbt_status bbt_mibf_free(bbt_mibf *mf);
bbt_status bbt_mibf_info(const bbt_mibf *mf, bbt_mibf_info_t *info);
/* name of ID id (0 = "repeat"), or NULL */
const char *bbt_mibf_name(const bbt_mibf *mf, uint32_t id);
/* spaced seed i, or NULL (k-mer mode, i out of range) */
const char *bbt_mibf_seed(const bbt_mibf *mf, uint32_t i);
/* host copies: bv = m/64 words of the bit vector, data = pop slots */
bbt_status bbt_mibf_download(const bbt_mibf *mf, uint64_t *bv, uint32_t *data);
uint64_t bbt_mibf_calc_size(uint64_t entries, uint32_t nseeds, double occupancy);
/* <prefix>_summary.tsv text of biobloommicategorizer (ResultsManager<ID>
 * over "repeat" and the ID names, ResultsManager.hpp:40-64, 191-237) from
 * host counters: summary[0..nids] hits per ID, [nids+1..2nids+1] unique hits
 * per ID, [2nids+2] multiMatch, [2nids+3] noMatch. */
bbt_status bbt_mibf_summary_text(const bbt_mibf *mf, const uint64_t *summary, uint64_t nrec,
                                 char *buf, uint64_t cap, uint64_t *needed);
/* probs[0..nids]: chance that a random frame reports each ID */
bbt_status bbt_mibf_frame_probs(const bbt_mibf *mf, uint32_t allowed_miss, double *probs);
uint32_t bbt_mibf_min_count(uint64_t frames, double prob, double score, uint32_t nids);
/* Classifies a batch of reads or pairs (bbt_reads as for bbt_eval_batch;
 * order is ignored); one wavefront per record.  Records of up to 8192
 * frames x hashes (both mates of a pair together).  The work queue and the
 * wide-pass lists belong to the filter: calls on one filter run one after
 * another on the device (each call's kernels wait for the previous call's,
 * whichever stream either was given), so calls from several streams are
 * correct but do not overlap. */
bbt_status bbt_mibf_query_batch(bbt_mibf *mf, const bbt_reads *in, const bbt_mibf_params *p,
                                bbt_mibf_out *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
