"""The oracle's kseq_read restatement (oracle/bbt_kseq.cpp), which the CPU
end-to-end baseline parses with, against the hand-derived kseq_read cases
(Common/kseq.h:93-219) and the independent byte-by-byte model in
kseq_model.py on generated files of every record shape."""
import numpy as np
import pytest

import oracle_lib as O
from kseq_model import read_all
from test_ingest_host import CASES, random_file


def seqs(data):
    buf, off, st = O.kseq_seqs(data)
    return [bytes(buf[int(off[i]):int(off[i + 1])]) for i in range(len(off) - 1)], st


@pytest.mark.parametrize("case", range(len(CASES)))
def test_hand_derived_cases(case):
    data, want, status = CASES[case]
    got, st = seqs(data)
    assert got == [w[2] for w in want]
    assert st == status


@pytest.mark.parametrize("seed", range(6))
def test_generated_files_equal_model(seed):
    rng = np.random.default_rng(100 + seed)
    data = random_file(rng, 800, strict=seed % 3 == 0)
    want, wst = read_all(data)
    got, st = seqs(data)
    assert got == [r[2] for r in want]
    assert st == wst


def test_truncated_inputs_equal_model():
    """Every prefix of a small mixed file: the records and status the model
    gives (a record cut inside its quality ends the input with -2)."""
    rng = np.random.default_rng(7)
    data = random_file(rng, 12)
    for cut in range(0, len(data), 7):
        want, wst = read_all(data[:cut])
        got, st = seqs(data[:cut])
        assert got == [r[2] for r in want] and st == wst, cut
