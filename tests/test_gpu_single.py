"""The scalar entry points' single-pair kernels (one wave per call, the pair in the kernel
arguments): pairwise.score and pairwise.align (reference src/pairwise.cpp:131-267),
Index.align (src/index.cpp:105-179), distance.hamming and distance.levenshtein
(src/distance.cpp:68-95). Every result equals the oracle's (itself pinned to the compiled
reference), including tie-heavy three-letter pairs, the longest pairs the kernels take
(query 64, reference 256) and the first lengths past them, which take the batch kernels."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import launch_labels, rand_seq, seed_of

pytestmark = pytest.mark.gpu


def _tables(st, m):
    syms = st.amino_acids()
    sim = np.array([[m.similarity(a, b) for b in syms] for a in syms], np.int32)
    pen = np.array([[m.penalty(a, b) for b in syms] for a in syms], np.int32)
    return sim, pen


def _pairs(rng):
    cases = [(rand_seq(rng, 1, 30), rand_seq(rng, 1, 30)) for _ in range(40)]
    cases += [(rand_seq(rng, 1, 14, "AAC"), rand_seq(rng, 1, 14, "AAC")) for _ in range(40)]
    cases += [(rand_seq(rng, 60, 64), rand_seq(rng, 200, 256)) for _ in range(4)]
    cases += [(rand_seq(rng, 64, 64), rand_seq(rng, 256, 256)), ("W", "W"), ("A", "C")]
    return cases


@pytest.mark.parametrize("mode", ["global", "local"])
@pytest.mark.parametrize("go,ge", [(11, 1), (5, 5), (1, 3), (0, 0)])
def test_score_and_align(gpu, oracle_lib, mode, go, ge):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    sim, _ = _tables(st, m)
    rng = random.Random(seed_of("single", mode, go, ge))
    for q, r in _pairs(rng):
        s, labels = launch_labels(lambda: st.pairwise.score(q, r, m, mode=mode, gap_open=go,
                                                            gap_extend=ge))
        assert f"single_gotoh_{mode}" in labels, labels
        assert s == oracle_lib.gotoh(q, r, sim, mode == "local", go, ge), (q, r)
        a, labels = launch_labels(lambda: st.pairwise.align(q, r, m, mode=mode, gap_open=go,
                                                            gap_extend=ge))
        assert f"single_traceback_{mode}" in labels, labels
        s_, ops, qb, rb = oracle_lib.align_pair(q, r, sim, mode == "local", go, ge)
        assert (a.score, a.ops) == (s_, ops), (q, r)


def test_index_align(gpu, oracle_lib):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    _, pen = _tables(st, m)
    rng = random.Random(17)
    for go, ge in ((11, 1), (3, 1), (2, 5)):
        for q, r in _pairs(rng)[:50]:
            idx = st.Index.build([r], "aa")
            for mat, table in ((m, pen), ("", None)):
                p = st.SearchParams(matrix=mat, gap_open=go, gap_extend=ge)
                a, labels = launch_labels(lambda: idx.align(0, q, p))
                assert "single_traceback_penalty" in labels, labels
                assert (a.score, a.ops) == oracle_lib.index_align(q, r, table, go, ge), (q, r)


def test_distances(gpu, oracle_lib):
    st = gpu
    rng = random.Random(23)
    for n in list(range(0, 70)) + [128, 255, 256]:
        a, b = rand_seq(rng, n, n, "ACGT"), rand_seq(rng, n, n, "ACGT")
        h, labels = launch_labels(lambda: st.distance.hamming(a, b))
        assert "single_hamming" in labels, labels
        assert h == oracle_lib.hamming(a, b)
    for _ in range(60):
        a, b = rand_seq(rng, 1, 64, "ACGT"), rand_seq(rng, 1, 256, "ACGT")
        if rng.random() < 0.5:
            a, b = b, a
        d, labels = launch_labels(lambda: st.distance.levenshtein(a, b))
        assert "single_levenshtein" in labels, labels
        assert d == oracle_lib.levenshtein(a, b), (a, b)


def test_past_the_limits_take_the_batch_kernels(gpu, oracle_lib):
    """Queries of 65 residues or references of 257, empty sides, and gap costs past the int32
    range bound go through the batch kernels and still equal the oracle."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    sim, pen = _tables(st, m)
    rng = random.Random(29)
    cases = [(rand_seq(rng, 65, 65), rand_seq(rng, 10, 20)),
             (rand_seq(rng, 10, 20), rand_seq(rng, 257, 257)), ("", "ACD"), ("ACD", "")]
    for q, r in cases:
        for mode in ("global", "local"):
            s, labels = launch_labels(lambda: st.pairwise.score(q, r, m, mode=mode))
            assert not any(lab.startswith("single_") for lab in labels), labels
            assert s == oracle_lib.gotoh(q, r, sim, mode == "local", 11, 1)
            a = st.pairwise.align(q, r, m, mode=mode)
            s_, ops, _, _ = oracle_lib.align_pair(q, r, sim, mode == "local", 11, 1)
            assert (a.score, a.ops) == (s_, ops)
    s, labels = launch_labels(lambda: st.pairwise.score("CASSL", "CASSF", m, gap_open=10**8,
                                                        gap_extend=1))
    assert not any(lab.startswith("single_") for lab in labels), labels
    assert s == oracle_lib.gotoh("CASSL", "CASSF", sim, False, 10**8, 1)
    long_a, long_b = rand_seq(rng, 257, 257, "ACGT"), rand_seq(rng, 257, 257, "ACGT")
    assert st.distance.hamming(long_a, long_b) == oracle_lib.hamming(long_a, long_b)
    x, y = rand_seq(rng, 65, 80, "ACGT"), rand_seq(rng, 65, 80, "ACGT")
    assert st.distance.levenshtein(x, y) == oracle_lib.levenshtein(x, y)
    assert st.distance.levenshtein("", "ACG") == 3 == oracle_lib.levenshtein("", "ACG")
