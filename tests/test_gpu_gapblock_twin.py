"""GPU parity for the twin short launch: two queries of one length per workgroup, packed into
the 16-bit halves of the LDS tables (reference: src/gapblock.cpp:39-62 `cell`). Every
result must equal the int32 short launch (single-query calls) and the CPU oracle, including
at the edge of the int16 range proof, for odd counts per length and for empty queries."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import cdr3_like, launch_labels, rand_seq

pytestmark = pytest.mark.gpu


def _pen(st, matrix):
    syms = st.alphabet_symbols("aa")
    return np.array([[matrix.penalty(a, b) for b in syms] for a in syms], np.int32)


def _oracle_rows(st, oracle_lib, q, r, m, go, ge, prior):
    from seqtree_amd.gapblock import _prior_cube

    w = max((len(s) for s in q + r), default=0)
    cube = _prior_cube(prior, w) if prior is not None else None
    return oracle_lib.gapblock_matrix(q, r, _pen(st, m), go, ge, cube, w, threads=8)


@pytest.mark.parametrize("prior", ["central", None])
def test_twin_rows_equal_single_rows_and_oracle(gpu, oracle_lib, prior):
    """Lengths with odd and even counts, singletons, empty queries and the longest short
    queries: the twin launch's rows equal single-query (int32) calls and the oracle."""
    st = gpu
    rng = random.Random(301 if prior else 302)
    m = st.SubstitutionMatrix.blosum62()
    gp = st.central_prior(21) if prior else None
    special = ([rand_seq(rng, 5, 5) for _ in range(3)] + [""] + [rand_seq(rng, 32, 32)
               for _ in range(2)] + [rand_seq(rng, 1, 1)] + [rand_seq(rng, 17, 17)])
    q = special + [rand_seq(rng, 0, 32) for _ in range(220)]
    rng.shuffle(q)
    r = [rand_seq(rng, 0, 64) for _ in range(3000)] + q[:5]
    full, labels = launch_labels(lambda: np.asarray(st.score_matrix(q, r, m, 28, 1, gp)))
    assert "gapblock_twin" in labels, labels
    special_rows = [i for i, s in enumerate(q) if s in special]
    rows = sorted(set(special_rows) | set(rng.sample(range(len(q)), 30)))
    for i in rows:
        one = np.asarray(st.score_matrix(q[i:i + 1], r, m, 28, 1, gp))
        assert np.array_equal(full[i:i + 1], one), f"row {i} (length {len(q[i])})"
    sample = sorted(rng.sample(range(len(q)), 12)) + special_rows[:4]
    want = _oracle_rows(st, oracle_lib, [q[i] for i in sample], r, m, 28, 1, gp)
    assert np.array_equal(full[sample], want)


def test_all_lengths_distinct(gpu, oracle_lib):
    """Every length once: every twin is a query paired with itself (one row written)."""
    st = gpu
    rng = random.Random(303)
    m = st.SubstitutionMatrix.pam250()
    q = [rand_seq(rng, n, n) for n in range(33)]
    rng.shuffle(q)
    r = [rand_seq(rng, 0, 40) for _ in range(1500)]
    full, labels = launch_labels(lambda: np.asarray(st.score_matrix(q, r, m, 17, 2)))
    assert "gapblock_twin" in labels
    assert np.array_equal(full, _oracle_rows(st, oracle_lib, q, r, m, 17, 2, None))


@pytest.mark.parametrize("over", [0, 1])
def test_int16_bound_edge(gpu, oracle_lib, over):
    """A constant prior R with 6 T + R = 2^15 - 2 (twin launch, the tightest admitted case)
    and 2^15 - 1 (the int32 short launch): both equal the oracle."""
    st = gpu
    rng = random.Random(304 + over)
    m = st.SubstitutionMatrix.blosum62()
    q = [rand_seq(rng, 3, 12) for _ in range(40)] + [rand_seq(rng, 12, 12)]
    r = [rand_seq(rng, 0, 24) for _ in range(1200)]
    T = 12 * int(np.abs(_pen(st, m)).max())
    R = 32766 - 6 * T + over
    assert R > 0
    gp = lambda i, d, n: R  # noqa: E731
    full, labels = launch_labels(lambda: np.asarray(st.score_matrix(q, r, m, 28, 1, gp)))
    assert ("gapblock_twin" in labels) == (over == 0), labels
    assert np.array_equal(full, _oracle_rows(st, oracle_lib, q, r, m, 28, 1, gp))


def test_twin_threshold_equals_dense(gpu):
    """Thresholded output from the twin launch: both queries of a twin append their hits."""
    st = gpu
    rng = random.Random(305)
    m = st.SubstitutionMatrix.blosum62()
    q = cdr3_like(rng, 301, 8, 20)
    r = cdr3_like(rng, 4000, 8, 20)
    gp = st.central_prior(21)
    (hq, hr, hs), labels = launch_labels(
        lambda: st.gapblock.score_threshold(q, r, 70, m, 28, 1, gp))
    assert "gapblock_threshold_twin" in labels, labels
    dense = np.concatenate([np.asarray(st.score_matrix(q[i:i + 1], r, m, 28, 1, gp))
                            for i in range(len(q))])
    want = np.argwhere(dense <= 70)
    assert np.array_equal(np.stack([hq, hr], 1), want)
    assert np.array_equal(hs, dense[dense <= 70])
