/*
 * front_end.h -- host side of biobloomcategorizer over libbbt_hip.
 *
 * Mirrors the reference's BioBloomClassifier (BioBloomCategorizer/
 * BioBloomClassifier.{h,cpp}) : same constructor arguments, same filter(),
 * filterPair(), filterPrint() entry points and outputs (<prefix>_summary.tsv,
 * --fa/--fq partition files, -d/-n stdout), with the per-read evaluation
 * replaced by device batches through bbt_eval_batch.
 */
#ifndef BBT_FRONT_END_H
#define BBT_FRONT_END_H

#include <stdint.h>
#include <zlib.h>

#include <string>
#include <vector>

#include "bbt_hip.h"

namespace bbt_fe {

/* Options the reference keeps in the opt:: namespace
 * (Common/Options.h:17-54, BioBloomCategorizer/Options.h:18-49). */
struct Options {
    uint32_t mode = BBT_STD;         /* opt::mode */
    uint32_t scorer = BBT_SIMPLE;    /* opt::scoringMethod */
    double score = 0.15;             /* opt::score */
    uint32_t streak = 3;             /* opt::streakThreshold */
    bool inclusive = false;          /* opt::inclusive */
    bool inverse = false;            /* opt::inverse */
    uint64_t file_interval = 10000000; /* opt::fileInterval */
    uint32_t path = BBT_PATH_BATCHED;
    uint64_t batch_reads = 1u << 18; /* reads per device batch */
    bool read_status = false;        /* write <prefix>_readStatus.tsv */
    bool dust = false;               /* opt::dust (-D) */
    uint32_t dust_t = 20;            /* opt::dustT (-T) */
    uint32_t dust_w = 64;            /* opt::dustWindow (-W) */
};

/* One FASTA/FASTQ record as kseq_read delivers it (Common/kseq.h:177-219). */
struct SeqRecord {
    std::string name, comment, seq, qual;
    bool fastq = false;
};

/* FASTA/FASTQ(.gz) reader with kseq_read's contract: name up to the first
 * whitespace, comment = rest of the header line, sequence lines joined,
 * FASTQ quality read until it is as long as the sequence.  read() returns
 * the sequence length, -1 at end of file, -2 for a truncated quality, -3 on
 * a stream error. */
class SeqReader {
  public:
    explicit SeqReader(const std::string &path);
    ~SeqReader();
    bool ok() const { return m_fp != nullptr; }
    int64_t read(SeqRecord &rec);

  private:
    int getc_();
    bool getline_(std::string &out, bool stop_at_space, int *delim);
    gzFile m_fp = nullptr;
    std::vector<char> m_buf;
    int m_begin = 0, m_end = 0;
    bool m_eof = false, m_err = false;
    int m_last = 0; /* header char already consumed */
};

class BioBloomClassifier {
  public:
    BioBloomClassifier(const std::vector<std::string> &filterFilePaths,
                       double scoreThreshold, const std::string &outputPrefix,
                       const std::string &outputPostFix, const Options &opt);
    ~BioBloomClassifier();
    void filter(const std::vector<std::string> &inputFiles);
    void filterPrint(const std::vector<std::string> &inputFiles,
                     const std::string &outputType);
    void filterPair(const std::string &file1, const std::string &file2);
    void filterPairPrint(const std::string &file1, const std::string &file2,
                         const std::string &outputType);
    /* one file holding both mates ("smart pair") */
    void filterPair(const std::string &file);
    void filterPairPrint(const std::string &file, const std::string &outputType);
    void filterPair(const std::vector<std::string> &files1,
                    const std::vector<std::string> &files2);
    void setOrderedFilter();
    void setInclusive() { m_opt.inclusive = true; }
    void setStdout() { m_stdout = true; }
    uint64_t probes() const { return m_probes; }
    uint64_t lookups() const { return m_lookups; }

  private:
    struct Batch;
    void run_batch(std::vector<SeqRecord> &recs, bool paired);
    void emit_batch(const std::vector<SeqRecord> &recs, bool paired,
                    const std::string &outputType);
    void close_outputs(const std::string &outputType);
    void write_summary(uint64_t totalReads);
    void progress(uint64_t &total, uint64_t add);
    void pair_loop(const std::string &f1, const std::string &f2, const std::string &type);
    void smart_pair_loop(const std::string &file, const std::string &type);
    void pair_run(const std::string &file1, const std::string *file2, const std::string &type);

    bbt_filterset *m_fs = nullptr;
    uint32_t m_nf = 0;
    std::vector<std::string> m_ids;
    double m_threshold;
    std::string m_prefix, m_postfix;
    Options m_opt;
    bool m_stdout = false;
    Batch *m_b = nullptr;
    uint64_t m_probes = 0, m_lookups = 0;
    uint64_t m_pairs = 0;
    std::vector<void *> m_outs; /* gzFile or FILE* per class */
    bool m_gz = false;
    void *m_status = nullptr;
};

} // namespace bbt_fe

#endif
