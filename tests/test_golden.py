"""Golden vectors written by the reference's own C++ (tests/golden/make_golden.py).

The CPU test pins the oracle to them on any host (the reference tree need not be present);
the GPU test checks the product's kernels against the same vectors."""
from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "golden.json"


@lru_cache(maxsize=1)
def cases():
    return json.loads(GOLDEN.read_text())["cases"]


def _prior(name):
    from golden.make_golden import PRIORS

    return PRIORS[name]


def _tables(st, matrix, alphabet="aa"):
    syms = st.alphabet_symbols(alphabet)
    if matrix is None:
        A = len(syms)
        return np.eye(A, dtype=np.int32), (1 - np.eye(A, dtype=np.int32))
    m = getattr(st.SubstitutionMatrix, matrix)()
    sim = np.array([[m.similarity(a, b) for b in syms] for a in syms], np.int32)
    pen = np.array([[m.penalty(a, b) for b in syms] for a in syms], np.int32)
    return sim, pen


def test_golden_file_covers_every_kernel():
    kinds = {c["kind"] for c in cases()}
    assert kinds == {"gapblock", "gotoh", "hamming", "levenshtein"}
    assert any(c.get("error") for c in cases())  # an overflow case is recorded


def test_oracle_reproduces_golden(oracle_lib, st):
    from seqtree_amd.gapblock import _prior_cube

    for c in cases():
        q, r = c["queries"], c["refs"]
        if c["kind"] == "gapblock":
            _, pen = _tables(st, c["matrix"], c["alphabet"])
            prior = _prior(c["prior"])
            w = max(len(s) for s in q + r)
            got = oracle_lib.gapblock_matrix(q, r, pen, c["gap_open"], c["gap_extend"],
                                             _prior_cube(prior, w) if prior else None, w,
                                             alphabet=c["alphabet"])
        elif c["kind"] == "gotoh":
            sim, _ = _tables(st, c["matrix"])
            code, got = oracle_lib.align_matrix(q, r, sim, c["mode"] == "local", c["dist"],
                                                c["gap_open"], c["gap_extend"])
            if c["error"]:
                assert code == 2
                continue
        else:
            code, got = oracle_lib.distance_matrix(c["kind"], q, r)
        assert got.tolist() == c["expected"], c["kind"]


@pytest.mark.gpu
def test_gpu_reproduces_golden(gpu):
    st = gpu
    for c in cases():
        q, r = c["queries"], c["refs"]
        if c["kind"] == "gapblock":
            m = getattr(st.SubstitutionMatrix, c["matrix"])() if c["matrix"] else None
            got = st.score_matrix(q, r, m, c["gap_open"], c["gap_extend"], _prior(c["prior"]),
                                  alphabet=c["alphabet"])
        elif c["kind"] == "gotoh":
            m = getattr(st.SubstitutionMatrix, c["matrix"])()
            fn = st.pairwise.dist_matrix if c["dist"] else st.pairwise.score_matrix
            if c["error"]:
                with pytest.raises(OverflowError, match="distance" if c["dist"] else "score"):
                    fn(q, r, m, mode=c["mode"], gap_open=c["gap_open"],
                       gap_extend=c["gap_extend"])
                continue
            got = fn(q, r, m, mode=c["mode"], gap_open=c["gap_open"], gap_extend=c["gap_extend"])
        elif c["kind"] == "hamming":
            got = st.distance.hamming_matrix(q, r)
        else:
            got = st.distance.levenshtein_matrix(q, r)
        assert np.asarray(got).tolist() == c["expected"], c["kind"]
