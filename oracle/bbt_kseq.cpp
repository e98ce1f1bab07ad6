/*
 * bbt_kseq.cpp -- kseq_read over a buffer, for the CPU end-to-end baseline.
 * TEST INFRASTRUCTURE ONLY (see bbt_oracle.h).
 *
 * Restates Common/kseq.h:93-219 (ks_getuntil2 :93-143, kseq_read :177-219)
 * one byte at a time, keeping only the sequences:
 *   - a record starts at '>' or '@' (bytes before the first one are skipped);
 *   - the name ends at the first whitespace byte, and the rest of a header line
 *     that goes on is the comment;
 *   - sequence lines follow up to a line that starts with '>', '+' or '@'; empty
 *     lines are skipped; a line read as a whole loses a final '\r' when the
 *     text read so far is longer than one byte;
 *   - after a '+' line, quality lines are read until the quality is as long as the
 *     sequence. A missing quality, or one of another length, ends the input
 *     with -2.
 * Parity: pinned against tests/kseq_model.py (an independent byte-by-byte
 * restatement) and the hand-written cases of tests/test_ingest_host.py, in
 * tests/test_oracle_kseq.py.
 */
#include <stdint.h>

#include <string>

#include "bbt_oracle.h"

namespace {

struct Reader {
    const char *d;
    uint64_t n, p = 0;
    int last = 0;

    int getc() { return p < n ? (unsigned char)d[p++] : -1; }
    static bool space(int c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

    /* ks_getuntil2: to the end of the line, or to whitespace; returns the
     * delimiter (or -1 at the end of input) */
    int until(bool line, std::string &s, bool append, bool &eof) {
        if (!append)
            s.clear();
        eof = p >= n;
        if (eof)
            return -1;
        int dret = 0;
        while (p < n) {
            const int c = (unsigned char)d[p++];
            if (line ? c == '\n' : space(c)) {
                dret = c;
                break;
            }
            s.push_back(char(c));
        }
        if (line && s.size() > 1 && s.back() == '\r')
            s.pop_back();
        return dret;
    }

    /* kseq_read: the sequence length, or -1 / -2 */
    int64_t read(std::string &seq) {
        std::string name, comment, qual;
        bool eof = false;
        int c = 0;
        if (last == 0) {
            while ((c = getc()) >= 0 && c != '>' && c != '@') {
            }
            if (c < 0)
                return -1;
            last = c;
        }
        seq.clear();
        const int dl = until(false, name, false, eof);
        if (eof)
            return -1;
        if (dl != '\n')
            until(true, comment, false, eof);
        while ((c = getc()) >= 0 && c != '>' && c != '+' && c != '@') {
            if (c == '\n')
                continue;
            seq.push_back(char(c));
            until(true, seq, true, eof);
        }
        if (c == '>' || c == '@')
            last = c;
        if (c != '+')
            return int64_t(seq.size());
        while ((c = getc()) >= 0 && c != '\n') {
        }
        if (c == -1)
            return -2;
        do {
            until(true, qual, true, eof);
        } while (!eof && qual.size() < seq.size());
        last = 0;
        if (qual.size() != seq.size())
            return -2;
        return int64_t(seq.size());
    }
};

} // namespace

extern "C" int64_t orc_kseq_seqs(const char *buf, uint64_t len, char *seq_out, uint64_t cap,
                                 uint64_t *off_out, uint64_t cap_recs, int64_t *status) {
    Reader r{buf, len};
    std::string seq;
    uint64_t o = 0, nrec = 0;
    off_out[0] = 0;
    int64_t st;
    while ((st = r.read(seq)) >= 0) {
        if (nrec + 1 >= cap_recs || o + seq.size() > cap)
            return -1;
        seq.copy(seq_out + o, seq.size());
        o += seq.size();
        off_out[++nrec] = o;
    }
    if (status)
        *status = st;
    return int64_t(nrec);
}
