"""kseq_read (Common/kseq.h:93-219) in plain Python, byte by byte, for the
ingest tests: a second, independent restatement that the C++ readers are
compared with on generated files (the hand-written cases in
test_ingest_host.py pin both)."""

_SPACE = frozenset(b" \t\n\v\f\r")


class Kseq:
    def __init__(self, data):
        self.d = bytes(data)
        self.p = 0
        self.last = 0

    def getc(self):
        if self.p >= len(self.d):
            return -1
        c = self.d[self.p]
        self.p += 1
        return c

    def getuntil(self, line, s, append):
        """ks_getuntil2: line=True splits at '\\n', else at isspace."""
        if not append:
            del s[:]
        if self.p >= len(self.d):
            return -1, 0
        dret = 0
        while self.p < len(self.d):
            c = self.d[self.p]
            self.p += 1
            if (c == 10) if line else (c in _SPACE):
                dret = c
                break
            s.append(c)
        if line and len(s) > 1 and s[-1] == 13:
            s.pop()
        return len(s), dret

    def read(self):
        """-> (status, (name, comment, seq, qual, fastq) or None)"""
        if self.last == 0:
            while True:
                c = self.getc()
                if c < 0 or c in (62, 64):
                    break
            if c < 0:
                return c, None
            self.last = c
        name, comment, seq, qual = bytearray(), bytearray(), bytearray(), bytearray()
        r, c = self.getuntil(False, name, False)
        if r < 0:
            return r, None
        if c != 10:
            self.getuntil(True, comment, False)
        while True:
            c = self.getc()
            if c < 0 or c in (62, 43, 64):
                break
            if c == 10:
                continue
            seq.append(c)
            self.getuntil(True, seq, True)
        if c in (62, 64):
            self.last = c
        if c != 43:
            return len(seq), (bytes(name), bytes(comment), bytes(seq), b"", False)
        while True:
            c = self.getc()
            if c < 0 or c == 10:
                break
        if c == -1:
            return -2, None
        while True:
            r, _ = self.getuntil(True, qual, True)
            if not (r >= 0 and len(qual) < len(seq)):
                break
        self.last = 0
        if len(seq) != len(qual):
            return -2, None
        return len(seq), (bytes(name), bytes(comment), bytes(seq), bytes(qual), True)


def read_all(data):
    k = Kseq(data)
    out = []
    while True:
        st, rec = k.read()
        if st < 0:
            return out, st
        out.append(rec)
