"""GPU parity for the gap-block score matrix (reference: src/gapblock.cpp, the matrix contract
of tests/python/test_gapblock.py:479-557). The checker is the CPU oracle."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import AA, cdr3_like, rand_seq, seed_of

pytestmark = pytest.mark.gpu


def _pen(st, matrix, alphabet="aa"):
    """Penalty grid in codec order; unit matrices (any alphabet) are the 0/1 identity cost."""
    A = len(st.alphabet_symbols(alphabet))
    if matrix is None or alphabet != "aa":
        assert matrix is None or matrix.size() == A
        return (np.arange(A)[:, None] != np.arange(A)[None, :]).astype(np.int32)
    syms = st.alphabet_symbols(alphabet)
    return np.array([[matrix.penalty(a, b) for b in syms] for a in syms], np.int32)


def _check(st, oracle_lib, q, r, matrix=None, go=28, ge=1, prior=None, alphabet="aa"):
    from seqtree_amd.gapblock import _prior_cube

    got = np.asarray(st.score_matrix(q, r, matrix, go, ge, prior, alphabet=alphabet))
    width = max((len(s) for s in q + r), default=0)
    cube = _prior_cube(prior, width) if prior is not None else None
    want = oracle_lib.gapblock_matrix(q, r, _pen(st, matrix, alphabet), go, ge, cube, width,
                                      alphabet=alphabet, threads=8)
    assert got.shape == want.shape
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} cells differ, first {bad[:3].tolist()}"
    return got


def _priors(st):
    g = st.gapblock
    return {
        "none": None,
        "central": g.central_prior(21),
        "frame": g.frame_prior(21, 5),
        "positions": g.positions_prior((3, 4, -4, -3)),
        "profile": g.profile_prior(21, lambda j, m: 1.0 - abs(2 * j + 1 - m) / m),
    }


@pytest.mark.parametrize("name", ["none", "central", "frame", "positions", "profile"])
def test_matrix_equals_scalar_scorer_and_oracle(gpu, oracle_lib, name):
    st = gpu
    prior = _priors(st)[name]
    rng = random.Random(7)
    seqs = [rand_seq(rng, 4, 20) for _ in range(40)]
    m = st.SubstitutionMatrix.blosum62()
    got = _check(st, oracle_lib, seqs, seqs, m, 28, 1, prior)
    pen = {(a, b): m.penalty(a, b) for a in AA for b in AA}
    for i, q in enumerate(seqs):
        for k, r in enumerate(seqs):
            assert got[i, k] == st.gapblock_score(q, r, m, 28, 1, prior, _pen=pen)[0]
    # a valid ball: zero diagonal, non-negative, symmetric
    assert (np.diag(got) == 0).all() and (got >= 0).all() and (got == got.T).all()


@pytest.mark.parametrize("matrix", ["blosum62", "blosum45", "blosum80", "pam250", "pam100",
                                    "structural", None])
@pytest.mark.parametrize("go,ge", [(1, 1), (28, 1), (0, 3), (2, 5), (0, 0)])
def test_random_ragged_sets(gpu, oracle_lib, matrix, go, ge):
    st = gpu
    rng = random.Random(seed_of(matrix, go, ge))
    q = [rand_seq(rng, 0, 33) for _ in range(37)]
    r = [rand_seq(rng, 0, 33) for _ in range(2100)]  # three chunks, the last one partial
    mat = getattr(st.SubstitutionMatrix, matrix)() if matrix else None
    _check(st, oracle_lib, q, r, mat, go, ge, st.gapblock.central_prior(9))
    _check(st, oracle_lib, q, r, mat, go, ge, None)


def test_unit_cost_and_empty_inputs(gpu, oracle_lib):
    st = gpu
    sm = st.score_matrix(["ACGT", "ACTT"], ["ACGT"], None, gap_open=1)
    assert sm[0, 0] == 0 and sm[1, 0] == 1
    assert st.score_matrix([], ["ACGT"]).shape == (0, 1)
    assert st.score_matrix(["ACGT"], []).shape == (1, 0)
    _check(st, oracle_lib, ["", "A", "AC", ""], ["", "ACGT", "T"], None, 1, 1, alphabet="nt")


def test_nucleotide_alphabets(gpu, oracle_lib):
    st = gpu
    rng = random.Random(3)
    q = [rand_seq(rng, 5, 30, "ACGT") for _ in range(50)]
    r = [rand_seq(rng, 5, 30, "acgt") for _ in range(300)]
    _check(st, oracle_lib, q, r, None, 3, 1, st.gapblock.central_prior(2), alphabet="nt")
    qi = [rand_seq(rng, 5, 30, "ACGTRYSWKMBDHVN") for _ in range(40)]
    _check(st, oracle_lib, qi, qi, st.SubstitutionMatrix.unit(15), 2, 1, alphabet="iupac")


def test_wrapping_prior_takes_the_exact_path(gpu, oracle_lib):
    """A prior near INT32_MAX makes the int32 candidate sums wrap; the kernel must then pick
    the same minimum as the reference's int32 evaluation order."""
    st = gpu
    rng = random.Random(5)
    seqs = [rand_seq(rng, 3, 14) for _ in range(60)]
    prior = lambda i, d, m: (2**31 - 1) - 7 * i  # noqa: E731
    _check(st, oracle_lib, seqs, seqs, st.SubstitutionMatrix.blosum62(), 28, 1, prior)


def test_long_queries_use_the_table_path(gpu, oracle_lib):
    """Queries past 512 residues no longer fit an LDS profile; the A x A table path runs."""
    st = gpu
    rng = random.Random(9)
    q = [rand_seq(rng, 500, 700) for _ in range(6)]
    r = [rand_seq(rng, 480, 720) for _ in range(90)] + q[:2]
    _check(st, oracle_lib, q, r, st.SubstitutionMatrix.blosum62(), 28, 1, None)


def test_custom_asymmetric_matrix_operand_order(gpu, oracle_lib):
    """pen is indexed [query code][ref code]; an asymmetric table exposes a swapped operand."""
    st = gpu
    rng = np.random.default_rng(11)
    grid = rng.integers(-6, 12, size=(24, 24)).tolist()
    m = st.SubstitutionMatrix.from_similarity(grid)
    rr = random.Random(11)
    q = [rand_seq(rr, 2, 25) for _ in range(30)]
    r = [rand_seq(rr, 2, 25) for _ in range(500)]
    _check(st, oracle_lib, q, r, m, 17, 2, st.gapblock.frame_prior(3, 4))


def test_cdr3_workload_shape(gpu, oracle_lib):
    """1,000 x 10,000 CDR3-like strings, BLOSUM62, gap_open = 2 * scale, central prior."""
    st = gpu
    rng = random.Random(2024)
    q = cdr3_like(rng, 1000, 10, 20)
    r = cdr3_like(rng, 10000, 10, 20)
    m = st.SubstitutionMatrix.blosum62()
    _check(st, oracle_lib, q, r, m, 2 * m.scale(), 1, st.gapblock.central_prior(21))


def test_errors_and_buffer(gpu):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    sm = st.score_matrix(["CASSL", "CASL"], ["CASSL"], m, 28)
    mv = memoryview(sm)
    assert mv.format == "i" and mv.shape == (2, 1) and mv.c_contiguous
    assert mv[0, 0] == 0 and mv[1, 0] == sm[1, 0] > 0
    with pytest.raises(ValueError, match="negative"):
        st.score_matrix(["CAS"], ["CA"], m, 28, gap_prior=lambda i, d, n: -1)


# ---------------------------------------------------------------------------- sparse output

@pytest.mark.parametrize("cutoff", [-1, 0, 40, 90, 10**9])
def test_threshold_equals_the_dense_matrix(gpu, oracle_lib, cutoff):
    st = gpu
    from seqtree_amd.gapblock import _prior_cube

    rng = random.Random(cutoff & 0xFFFF)
    q = cdr3_like(rng, 60, 6, 30)
    r = cdr3_like(rng, 2500, 6, 30)
    m = st.SubstitutionMatrix.blosum62()
    prior = st.central_prior(21)
    qi, ri, sc = st.gapblock.score_threshold(q, r, cutoff, m, 28, 1, prior)
    w = max(len(s) for s in q + r)
    dense = oracle_lib.gapblock_matrix(q, r, _pen(st, m), 28, 1, _prior_cube(prior, w), w,
                                       threads=8)
    want = np.argwhere(dense <= cutoff)
    assert qi.dtype == np.uint32 and sc.dtype == np.int32
    assert np.array_equal(np.stack([qi, ri], 1), want)  # sorted row-major
    assert np.array_equal(sc, dense[want[:, 0], want[:, 1]])


def test_threshold_grows_past_the_first_capacity(gpu):
    """More hits than the first pass keeps (4 Mi): the count drives a second, complete pass."""
    st = gpu
    rng = random.Random(1)
    q = cdr3_like(rng, 2100, 8, 12)
    r = cdr3_like(rng, 2100, 8, 12)
    qi, ri, sc = st.gapblock.score_threshold(q, r, 10**9, None, 1, 1)
    assert len(qi) == 2100 * 2100
    dense = np.asarray(st.score_matrix(q, r, None, 1, 1))
    assert np.array_equal(sc, dense.ravel())


def test_pairs_equal_the_dense_matrix(gpu, oracle_lib):
    st = gpu
    from seqtree_amd.gapblock import _prior_cube

    rng = random.Random(4)
    q = [rand_seq(rng, 0, 40) for _ in range(50)]
    r = [rand_seq(rng, 0, 40) for _ in range(70)]
    m = st.SubstitutionMatrix.pam250()
    w = max(len(s) for s in q + r)
    npr = np.random.default_rng(4)
    pq = npr.integers(0, len(q), 5000).astype(np.uint32)
    pr = npr.integers(0, len(r), 5000).astype(np.uint32)
    for prior in (None, st.central_prior(9), lambda i, d, n: (2**31 - 1) - 3 * i):
        got = st.gapblock.score_pairs(q, r, pq, pr, m, 17, 2, prior)
        cube = _prior_cube(prior, w) if prior else None
        dense = oracle_lib.gapblock_matrix(q, r, _pen(st, m), 17, 2, cube, w, threads=8)
        assert np.array_equal(got, dense[pq, pr])


@pytest.mark.parametrize("prior", ["central", None])
def test_short_launch_mixed_waves_and_chunk_groups(gpu, oracle_lib, prior):
    """Queries up to 32 and references up to 64 residues (the short launch: mixed-length
    waves, zero rows and prior sentinels) over enough rows that workgroups take several
    reference chunks each; every row is checked against the oracle on a sample of rows and
    against per-row recomputation for the rest."""
    st = gpu
    rng = random.Random(77 if prior else 78)
    m = st.SubstitutionMatrix.blosum62()
    q = [rand_seq(rng, 0, 32) for _ in range(2000)]
    r = [rand_seq(rng, 0, 64) for _ in range(5000)]
    gp = st.central_prior(21) if prior else None
    full = np.asarray(st.score_matrix(q, r, m, 28, 1, gp))
    rows = sorted(rng.sample(range(len(q)), 24))
    sub = _check(st, oracle_lib, [q[i] for i in rows], r, m, 28, 1, gp)
    assert np.array_equal(full[rows], sub)
    # small row blocks run with one chunk per workgroup: same values
    for i0 in range(0, len(q), 400):
        blk = np.asarray(st.score_matrix(q[i0:i0 + 7], r, m, 28, 1, gp))
        assert np.array_equal(full[i0:i0 + 7], blk)


def test_many_reference_chunks_in_groups(gpu, oracle_lib):
    """Enough references (40 chunks) that workgroups take chunk groups in group-major
    order, checked against the oracle on sampled rows and against single-row calls."""
    st = gpu
    rng = random.Random(91)
    m = st.SubstitutionMatrix.blosum62()
    q = cdr3_like(rng, 300, 6, 20)
    r = cdr3_like(rng, 40_000, 6, 20)
    gp = st.central_prior(21)
    full = np.asarray(st.score_matrix(q, r, m, 28, 1, gp))
    rows = sorted(rng.sample(range(len(q)), 5))
    sub = _check(st, oracle_lib, [q[i] for i in rows], r, m, 28, 1, gp)
    assert np.array_equal(full[rows], sub)
    for i in rng.sample(range(len(q)), 3):
        assert np.array_equal(full[i:i + 1], np.asarray(st.score_matrix(q[i:i + 1], r, m, 28, 1, gp)))
    hq, hr, hs = st.gapblock.score_threshold(q, r, 60, m, 28, 1, gp)
    want = np.argwhere(full <= 60)
    got = np.stack([hq, hr], 1)
    assert np.array_equal(got, want) and np.array_equal(hs, full[full <= 60])
