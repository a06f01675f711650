"""Parity at the benchmark configurations' full sizes, through properties that hold for every
entry of the matrix, plus whole rows against the oracle.

* Gotoh `dist_matrix` on 20,000 CDR3-like strings (BLOSUM62, 11/1, both modes): the full
  score matrix and the full distance matrix of one set against itself. Every entry of the
  distance matrix equals s(a,a) + s(b,b) - 2 s(a,b) read from the score matrix (a
  different kernel output), both are symmetric (BLOSUM62 is), the distance diagonal is zero,
  and sampled rows equal the oracle's int64 DP.
* Hamming on 50,000 random ACGT UMIs of length 12 against themselves (2.5e9 pairs, a 10 GB
  matrix): zero diagonal, row sums equal column sums (int64), 2e6 random (i, j) entries equal
  their transposes, and sampled rows equal the oracle's.
* Levenshtein on the 20,000 ACGT reads of the benchmark (10,000 of length 40-60, 10,000
  derived by 1-3 edits, shuffled) against themselves: the same symmetry properties, the
  triangle inequality on random triples, and sampled rows equal the oracle's.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sym_checks(D: np.ndarray, rng: np.random.Generator, samples: int) -> None:
    n = D.shape[0]
    assert not np.diagonal(D).any()
    rows = D.sum(axis=1, dtype=np.int64)
    cols = np.zeros(n, np.int64)
    for i0 in range(0, n, 2048):  # column sums in row blocks (bounded temporaries)
        cols += D[i0:i0 + 2048].sum(axis=0, dtype=np.int64)
    assert np.array_equal(rows, cols)
    i = rng.integers(0, n, samples)
    j = rng.integers(0, n, samples)
    assert np.array_equal(D[i, j], D[j, i])


@pytest.mark.parametrize("mode", ["global", "local"])
def test_gotoh_c2_full_size(gpu, oracle_lib, mode):
    st = gpu
    from seqtree_amd import synthetic as syn

    rng = np.random.default_rng(2 if mode == "global" else 3)
    codes, off, _ = syn.cdr3_csr(rng, 20_000, 8, 25)
    seqs = syn.decode(codes, off)
    m = st.SubstitutionMatrix.blosum62()
    S = np.asarray(st.pairwise.score_matrix(seqs, seqs, m, mode=mode))
    D = np.asarray(st.pairwise.dist_matrix(seqs, seqs, m, mode=mode))
    assert S.shape == D.shape == (20_000, 20_000)
    d = np.diagonal(S).astype(np.int64)
    for i0 in range(0, 20_000, 2000):  # every entry: D = s_aa + s_bb - 2 s_ab
        blk = d[i0:i0 + 2000, None] + d[None, :] - 2 * S[i0:i0 + 2000].astype(np.int64)
        assert np.array_equal(D[i0:i0 + 2000], blk), i0
    for i0 in range(0, 20_000, 2000):
        assert np.array_equal(S[i0:i0 + 2000], S[:, i0:i0 + 2000].T), i0
    _sym_checks(D, rng, 1_000_000)
    syms = st.amino_acids()
    sim = np.array([[m.similarity(a, b) for b in syms] for a in syms], np.int32)
    rows = [0, 7_777, 19_999]
    code, want = oracle_lib.align_matrix([seqs[r] for r in rows], seqs, sim, mode == "local",
                                         True, 11, 1, threads=16)
    assert code == 0
    assert np.array_equal(D[rows], np.asarray(want).reshape(len(rows), -1))


def test_hamming_c3_full_size(gpu, oracle_lib):
    st = gpu
    rng = np.random.default_rng(5)
    umis = ["".join(u) for u in np.array(list("ACGT"))[rng.integers(0, 4, (50_000, 12))]]
    D = np.asarray(st.distance.hamming_matrix(umis, umis))
    assert D.shape == (50_000, 50_000)
    assert D.max() <= 12 and D.min() >= 0
    _sym_checks(D, rng, 2_000_000)
    rows = [1, 24_999, 49_998]
    code, want = oracle_lib.distance_matrix("hamming", [umis[r] for r in rows], umis,
                                            threads=16)
    assert code == 0 and np.array_equal(D[rows], want)


def _edit(rng: np.random.Generator, s: str) -> str:
    out = list(s)
    for _ in range(int(rng.integers(1, 4))):
        op, p = int(rng.integers(0, 3)), int(rng.integers(0, len(out)))
        c = "ACGT"[int(rng.integers(0, 4))]
        if op == 0:
            out[p] = c
        elif op == 1:
            out.insert(p, c)
        elif len(out) > 1:
            del out[p]
    return "".join(out)


def test_levenshtein_c3_full_size(gpu, oracle_lib):
    st = gpu
    rng = np.random.default_rng(6)
    base = ["".join(np.array(list("ACGT"))[rng.integers(0, 4, int(rng.integers(40, 61)))])
            for _ in range(10_000)]
    reads = base + [_edit(rng, s) for s in base]
    reads = [reads[i] for i in rng.permutation(len(reads))]
    D = np.asarray(st.distance.levenshtein_matrix(reads, reads))
    assert D.shape == (20_000, 20_000)
    _sym_checks(D, rng, 1_000_000)
    a, b, c = (rng.integers(0, 20_000, 200_000) for _ in range(3))
    assert np.all(D[a, c] <= D[a, b].astype(np.int64) + D[b, c])
    rows = [3, 10_001, 19_996]
    code, want = oracle_lib.distance_matrix("levenshtein", [reads[r] for r in rows], reads,
                                            threads=16)
    assert code == 0 and np.array_equal(D[rows], want)
