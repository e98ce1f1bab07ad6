"""The known answers of the reference's own unit tests (Tests/SeqEvalTests.cpp,
Tests/BloomFilterTests.cpp, Tests/ntHashTests.cpp), kept as data in
tests/golden/reference_kat.json, checked against the oracle on the CPU and
against the device path through the C ABI on the GPU."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "reference_kat.json")))


def test_oracle_seqeval_length():
    c = KAT["seqeval_length"]
    f = O.Filter(c["filter"]["m"], c["filter"]["h"], c["filter"]["k"])
    for s in c["insert"]:
        f.insert_seq(s)
    for q in c["queries"]:
        got = O.eval_read(f, q["read"], O.params(O.LENGTH, q["min_match_len"]))[0]
        assert got is q["expect"], q


def test_oracle_bloom_membership():
    c = KAT["bloom_membership"]
    m, h, k = c["filter"]["m"], c["filter"]["h"], c["filter"]["k"]
    f = O.Filter(m, h, k)
    for s in c["insert"]:
        f.insert_hashes(O.nthash(s, k, h)[1][0])
    for s in c["present"]:
        assert f.contains_hashes(O.nthash(s, k, h)[1][0])[0]
    for s in c["absent"]:
        assert not f.contains_hashes(O.nthash(s, k, h)[1][0])[0]


def test_oracle_spaced_seed_symmetry():
    c = KAT["spaced_seed_symmetry"]
    span = len(c["seeds"][0])
    _, a = O.sthash(c["forward"], c["seeds"], span)
    _, b = O.sthash(c["reverse"], c["seeds"], span)
    assert bool(a[0, 0] != b[0, 1]) == c["first_window"]["fw0_ne_rv1"]
    assert bool(a[0, 1] != b[0, 0]) == c["first_window"]["fw1_ne_rv0"]
    assert bool(a[1, 0] == b[0, 1]) == c["after_one_step_of_forward"]["fw0_eq_rv1"]
    assert bool(a[1, 1] == b[0, 0]) == c["after_one_step_of_forward"]["fw1_eq_rv0"]


def _device():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    B.load()
    return torch, D, B


def _device_filter(torch, D, B, m, h, k, records):
    fs = B.FilterSet.create([m], [h], [k])
    seq = np.frombuffer(("\n".join(records) + "\n").encode(), np.uint8)
    dseq = torch.from_numpy(seq.copy()).cuda()
    fs.insert(0, dseq.data_ptr(), dseq.numel())
    fs.refresh(0)
    torch.cuda.synchronize()
    return fs


@pytest.mark.gpu
def test_device_seqeval_length():
    torch, D, B = _device()
    c = KAT["seqeval_length"]
    fs = _device_filter(torch, D, B, c["filter"]["m"], c["filter"]["h"], c["filter"]["k"],
                        c["insert"])
    for q in c["queries"]:
        buf, off = O.pack_reads([q["read"]])
        for path in (B.PATH_BATCHED, B.PATH_PLAIN):
            res = D.classify(fs, D.DeviceReads(buf, off),
                             B.params(scorer=B.LENGTH, threshold=q["min_match_len"], path=path))
            assert bool(res["hits"][0, 0]) is q["expect"], (q, path)


@pytest.mark.gpu
def test_device_bloom_membership():
    """A read of exactly k bases is one contains() call; LENGTH with
    min_match_len = k accepts it exactly when the k-mer is present."""
    torch, D, B = _device()
    c = KAT["bloom_membership"]
    m, h, k = c["filter"]["m"], c["filter"]["h"], c["filter"]["k"]
    fs = _device_filter(torch, D, B, m, h, k, c["insert"])
    reads = c["present"] + c["absent"]
    buf, off = O.pack_reads(reads)
    res = D.classify(fs, D.DeviceReads(buf, off), B.params(scorer=B.LENGTH, threshold=k))
    want = [1] * len(c["present"]) + [0] * len(c["absent"])
    assert res["hits"][:, 0].tolist() == want
    assert fs.info(0).m % c["stored_bits_multiple_of"] == 0
