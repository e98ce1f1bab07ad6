"""Randomized parity: filter sets (k 13..32, h 1..5, one to four filters,
sometimes of different k), read sets (empty to long reads, N runs, lower case,
other bytes) and evaluation settings (mode, scorer, threshold, streak, pairing,
-D, whole-wave length) drawn from fixed seeds, each checked on both paths
against the oracle: hits, classes, scores, contains() calls, bit lookups and
summary.tsv text."""
import numpy as np
import pytest

import oracle_lib as O
from test_gpu_parity import check_batch, oracle_filters_like

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from biobloom_amd import device as D
    from biobloom_amd import lib as B
    B.load()
    return D, B


def random_reads(rng, genomes, n_short, n_long):
    out = []
    for i in range(n_short + n_long):
        L = int(rng.integers(0, 300)) if i < n_short else int(rng.integers(500, 5000))
        g = genomes[int(rng.integers(len(genomes)))]
        if rng.random() < 0.6 and L < len(g):
            a = int(rng.integers(0, len(g) - L))
            r = g[a:a + L].copy()
            sub = rng.random(L) < rng.choice([0.0, 0.01, 0.05])
            r[sub] = rng.choice(np.frombuffer(b"ACGT", np.uint8), int(sub.sum()))
        else:
            r = rng.choice(np.frombuffer(b"ACGT", np.uint8), L)
        nrate = rng.choice([0.0, 0.001, 0.03])
        r[rng.random(L) < nrate] = ord("N")
        if rng.random() < 0.1:
            r[rng.random(L) < 0.3] |= 0x20  # lower case
        if rng.random() < 0.05 and L:
            r[int(rng.integers(L))] = rng.choice(np.frombuffer(b"RY-.", np.uint8))
        if rng.random() < 0.05 and L > 40:
            a = int(rng.integers(0, L - 30))
            r[a:a + 30] = np.frombuffer(b"AC" * 15, np.uint8)  # low complexity
        out.append(bytes(r))
    return out


@pytest.mark.parametrize("seed", range(8))
def test_random_settings_match_oracle(dev, seed):
    D, B = dev
    rng = np.random.default_rng(1000 + seed)
    L = B.load()
    nf = int(rng.integers(1, 5))
    k0 = int(rng.integers(13, 33))
    ks = [k0 if rng.random() < 0.7 else int(rng.integers(13, 33)) for _ in range(nf)]
    h = int(rng.integers(1, 6))
    fpr = float(rng.choice([0.01, 0.05, 0.2]))
    genomes = [rng.choice(np.frombuffer(b"ACGT", np.uint8), int(rng.integers(20_000, 80_000)))
               for _ in range(nf)]
    ms = [L.bbt_calc_optimal_size(len(g) - k + 1, fpr, h) for g, k in zip(genomes, ks)]
    fs = B.FilterSet.create(ms, [h] * nf, ks)
    for i, g in enumerate(genomes):
        dg = torch.from_numpy(np.ascontiguousarray(g)).to("cuda")
        fs.insert(i, dg.data_ptr(), dg.numel())
        fs.refresh(i)
        fs.set_id(i, f"r{i}")
        del dg
    ofs = oracle_filters_like(fs, B)
    reads = random_reads(rng, genomes, 240, 8)
    if len(reads) % 2:
        reads.append(b"")
    buf, off = O.pack_reads(reads)
    print(f"[gpu] seed={seed} nf={nf} k={ks} h={h} fpr={fpr} m={ms}", flush=True)
    for _ in range(4):
        mode = int(rng.choice([O.STD, O.STD, O.ORDERED, O.BESTHIT, O.SCORES]))
        scorer = int(rng.choice([O.SIMPLE, O.SIMPLE, O.HARMONIC, O.LENGTH, O.BINOMIAL]))
        thr = {O.SIMPLE: float(rng.choice([0.05, 0.15, 0.5, 0.9])),
               O.HARMONIC: float(rng.choice([0.1, 0.3])),
               O.LENGTH: float(rng.choice([30.0, 60.0])),
               O.BINOMIAL: float(rng.choice([1e-3, 1e-10]))}[scorer]
        streak = int(rng.choice([0, 1, 2, 3, 5]))
        paired = bool(rng.random() < 0.4)
        inclusive = bool(paired and rng.random() < 0.5)
        dust = (int(rng.choice([10, 20])), int(rng.choice([16, 32, 64, 100]))) \
            if rng.random() < 0.3 else None
        wave_min = int(rng.choice([0, 26, 100, 1000]))
        check_batch(D, B, fs, ofs, buf, off, paired=paired, inclusive=inclusive, mode=mode,
                    scorer=scorer, thr=thr, streak=streak, wave_min=wave_min, dust=dust)
