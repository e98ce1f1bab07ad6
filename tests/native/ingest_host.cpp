/*
 * ingest_host.cpp -- C entry points over the front ends' FASTA/FASTQ readers
 * (biobloom_amd/host/seq_ingest.cpp) for the CPU tests: every record of a
 * file, serialized, from the sequential or the block-parallel reader.
 */
#include <cstring>
#include <string>

#include "read_pack.h"
#include "seq_ingest.h"

extern "C" {

/* Reads the file once and serializes its records as name 0x1f comment 0x1f
 * seq 0x1f qual 0x1f (q|a) 0x1e; *out / *len receive the text (valid until
 * the next call), *status the first negative return.  mode 0: SeqReader,
 * 1: ParallelSeqReader::read, 2: its blocks of views
 * (ParallelSeqReader::next_block).  Returns the record count, or -1 when the
 * file cannot be opened. */
static std::string g_text;

long ingest_dump(const char *path, int mode, unsigned threads, unsigned long block,
                 const char **out, unsigned long *len, long *status, unsigned long *reparsed) {
    std::string &s = g_text;
    s.clear();
    bbt_fe::SeqRecord r;
    long n = 0, st = 0;
    auto put = [&] {
        s += r.name + '\x1f' + r.comment + '\x1f' + r.seq + '\x1f' + r.qual + '\x1f' +
             (r.fastq ? 'q' : 'a') + '\x1e';
        ++n;
    };
    if (mode == 0) {
        bbt_fe::SeqReader rd(path);
        if (!rd.ok())
            return -1;
        while ((st = long(rd.read(r))) >= 0)
            put();
    } else if (mode == 1) {
        bbt_fe::ParallelSeqReader rd(path, threads, block);
        if (!rd.ok())
            return -1;
        while ((st = long(rd.read(r))) >= 0)
            put();
        if (reparsed)
            *reparsed = rd.reparsed();
    } else {
        /* the views of whole blocks, as the classifier front end takes them */
        bbt_fe::ParallelSeqReader rd(path, threads, block);
        if (!rd.ok())
            return -1;
        while (auto blk = rd.next_block()) {
            for (size_t p = 0; p < blk->parts.size(); ++p)
                for (size_t i = 0; i < blk->count[p]; ++i) {
                    blk->parts[p][i].copy_to(r);
                    put();
                }
            st = long(blk->status);
        }
        if (reparsed)
            *reparsed = rd.reparsed();
    }
    *out = s.data();
    *len = s.size();
    *status = st;
    return n;
}

/* The 2-bit packer of the front ends (read_pack.h) over reads given back to
 * back in seq with nreads + 1 offsets: writes packed_bytes(off[nreads])
 * bytes into out and returns the code bytes. */
unsigned long pack_2bit_reads(const char *seq, const unsigned long *off, unsigned long nreads,
                              unsigned char *out, unsigned threads) {
    const uint64_t *o = reinterpret_cast<const uint64_t *>(off);
    const auto get = [&](uint64_t r) { return seq + o[r]; };
    return bbt_host::pack_2bit(get, o, nreads, out, threads);
}

unsigned long packed_bytes_for(unsigned long nbases) { return bbt_host::packed_bytes(nbases); }

/* the packer's form: -1 automatic, 0 SWAR, 1 AVX2; returns whether AVX2 runs here */
int pack_set_simd(int mode) {
    bbt_host::pack_simd_mode() = mode;
    return bbt_host::pack_avx2_available() ? 1 : 0;
}

} // extern "C"
