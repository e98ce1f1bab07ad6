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

#include <chrono>
#include <deque>
#include <functional>
#include <string>
#include <vector>

#include "bbt_hip.h"
#include "seq_ingest.h"

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
    unsigned threads = 0;            /* opt::threads (-t): parse threads per file, 0 = the reader's default */
    uint64_t ingest_block = 16u << 20; /* bytes of input parsed per block */
    bool pack2bit = true;            /* send reads 2-bit packed (false: ASCII) */
    bool autotune = false;           /* tune the batched path's knobs on the first batch */
    std::string tune_file;           /* tuning table: loaded at start, saved after tuning */
};

/* The records of one device batch: views into parsed blocks, the blocks
 * they point into, and records copied out of blocks (smart-pair mates). */
struct RecBatch {
    std::vector<SeqView> v;
    std::vector<ParallelSeqReader::BlockRef> holds;
    std::deque<SeqRecord> owned;
    void clear() {
        v.clear();
        holds.clear();
        owned.clear();
    }
    /* keeps blk alive with the batch (two readers alternate in pair input) */
    void hold(const ParallelSeqReader::BlockRef &blk) {
        const size_t n = holds.size();
        if ((n && holds[n - 1] == blk) || (n > 1 && holds[n - 2] == blk))
            return;
        holds.push_back(blk);
    }
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
    struct Writer;
    void pack(int slot, const RecBatch &recs, bool paired, uint64_t scale);
    void submit(int slot);
    void finish(int slot);
    void pipeline(const std::function<bool(RecBatch &)> &next, bool paired,
                  const std::string &type);
    void open_outputs(const std::string &type, bool paired);
    void emit_batch(int slot, const RecBatch &recs, bool paired, const std::string &outputType);
    void close_outputs(const std::string &outputType);
    void finish_writes();
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
    Batch *m_slot[2] = {nullptr, nullptr};
    uint64_t m_batch_limit = 1; /* reads (or pairs) the batch being gathered may take */
    uint64_t *m_dsummary = nullptr, *m_dctr = nullptr;
    uint64_t m_probes = 0, m_lookups = 0;
    uint64_t m_pairs = 0;
    std::vector<void *> m_outs; /* gzFile or FILE* per class */
    bool m_gz = false;
    void *m_status = nullptr;
    Writer *m_writer = nullptr; /* writes formatted batches in order, off this thread */
    std::chrono::steady_clock::time_point m_t0; /* "Filtering Start" */
    /* s: gather, pack + launch, device wait, output; staging allocation (part
     * of pack + launch); device time of the batches' kernels */
    double m_phase[6] = {0, 0, 0, 0, 0, 0};
    bool m_autotuned = false; /* --autotune ran on a batch */
};

} // namespace bbt_fe

#endif
