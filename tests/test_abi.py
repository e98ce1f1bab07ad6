"""The C ABI without a GPU: the library loads, exports every entry point that
include/bbt_hip.h declares, and reports file errors (missing .txt, bad or
wrong-sized .bf) with status codes instead of exiting."""
import ctypes as C
import os
import re
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "bbt_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bbt_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from biobloom_amd import lib as B
    L = B.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(B._SIGS), "binding and header disagree"


def write_txt(path, fid="myfilter", k=25, h=3, size=64 * 100):
    with open(path, "w") as f:
        f.write(f"[user_input_options]\nfilter_id={fid}\nkmer_size={k}\n"
                f"desired_false_positve_rate=0.075\nnumber_of_hash_functions={h}\n"
                f"expected_num_entries=10\nsequence_sources=a.fa \n\n[runtime_options]\n"
                f"size={size}\nnum_entries=10\napproximate_false_positive_rate=0.01\n"
                f"redundant_sequences=0\nredundant_fpr=0\n")


def write_bf(path, m, h=3, k=25, nbytes=None, magic=b"BlOOMFXX"):
    hdr = struct.pack("<8sIIQIIdQQ", magic, 56, 0, m, h, k, 0.075, 10, 10)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(b"\0" * (m // 8 if nbytes is None else nbytes))


def status_of(paths):
    from biobloom_amd import lib as B
    L = B.load()
    arr = (C.c_char_p * len(paths))(*[p.encode() for p in paths])
    h = C.c_void_p()
    st = L.bbt_filters_load(arr, len(paths), C.byref(h))
    return st, L.bbt_last_error().decode()


def test_missing_info_file(tmp_path):
    bf = str(tmp_path / "a.bf")
    write_bf(bf, 6400)
    st, msg = status_of([bf])
    assert st == 2 and "a.txt" in msg


def test_missing_bf(tmp_path):
    write_txt(str(tmp_path / "b.txt"))
    st, msg = status_of([str(tmp_path / "b.bf")])
    assert st == 3 and "cannot be opened" in msg


def test_wrong_size_bf(tmp_path):
    write_txt(str(tmp_path / "c.txt"))
    write_bf(str(tmp_path / "c.bf"), 6400, nbytes=100)
    st, msg = status_of([str(tmp_path / "c.bf")])
    assert st == 3 and "size" in msg


def test_bad_magic(tmp_path):
    write_txt(str(tmp_path / "d.txt"))
    write_bf(str(tmp_path / "d.bf"), 6400, magic=b"NOTABLOO")
    st, msg = status_of([str(tmp_path / "d.bf")])
    assert st == 3 and "header" in msg


def test_host_helpers_match_reference_formulas():
    from biobloom_amd import lib as B
    import oracle_lib as O
    L = B.load()
    assert L.bbt_calc_optimal_hashnum(0.075) == 3
    assert L.bbt_calc_optimal_size(5_000_000 - 24, 0.075, 3) == 27_387_648
    for n, p, q in [(126, 0.075, 1e-10), (1976, 0.0237, 1e-5), (30, 0.3, 0.01)]:
        assert L.bbt_calc_min_count(n, p, q) == O.lib().orc_calc_min_count(n, p, q)


def test_tuning_table_save_and_load(tmp_path):
    """The tuning table's text form: loaded entries are saved back line for
    line (sorted by geometry); a malformed line is a format error; a missing
    file an I/O error."""
    from biobloom_amd import lib as B
    L = B.load()
    lines = ["bytes_log2=26 nf=10 k=25 h=3 len_log2=7 mode=0 scorer=0 paired=1\t"
             "2147483647 0 4 0",
             "bytes_log2=28 nf=4 k=25 h=4 len_log2=15 mode=0 scorer=0 paired=0\t256 16 0 1",
             "bytes_log2=22 nf=1 k=25 h=3 len_log2=7 mode=0 scorer=0 paired=0\t"
             "2147483647 32 3 3"]
    src = tmp_path / "tune.tsv"
    src.write_text("\n".join(lines) + "\n")
    assert L.bbt_tuning_load(str(src).encode()) == 0
    out = tmp_path / "saved.tsv"
    assert L.bbt_tuning_save(str(out).encode()) == 0
    saved = out.read_text().splitlines()
    for ln in lines:
        assert ln in saved
    bad = tmp_path / "bad.tsv"
    bad.write_text("bytes_log2=1 nf=1\t1 2 x\n")
    assert L.bbt_tuning_load(str(bad).encode()) == 3
    assert "malformed line 1" in L.bbt_last_error().decode()
    bad.write_text("bytes_log2=1 nf=1\t256 16 0 16\n")  # probe_ahead has four bits
    assert L.bbt_tuning_load(str(bad).encode()) == 3
    assert L.bbt_tuning_load(str(tmp_path / "missing.tsv").encode()) == 2
