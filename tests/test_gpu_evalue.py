"""GPU parity for the score histograms behind control-calibrated E-values, and the reference's
E-value behaviour tests (tests/python/test_evalue.py of seqtree 0.6.0) over exhaustive
gap-block balls. Every histogram must equal the bincount of the oracle's dense gap-block
matrix; counts and cutoffs must equal what the sparse threshold path's hit lists give."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import cdr3_like, launch_labels, rand_seq, seed_of

pytestmark = pytest.mark.gpu

AA = "ACDEFGHIKLMNPQRSTVWY"


def _pen(st, matrix):
    syms = st.alphabet_symbols("aa")
    if matrix is None:
        return np.array([[0 if a == b else 1 for b in syms] for a in syms], np.int32)
    return np.array([[matrix.penalty(a, b) for b in syms] for a in syms], np.int32)


def _oracle_matrix(st, oracle_lib, q, r, m, go, ge, prior):
    from seqtree_amd.gapblock import _prior_cube

    w = max((len(s) for s in q + r), default=0)
    cube = _prior_cube(prior, w) if prior is not None else None
    return np.asarray(oracle_lib.gapblock_matrix(q, r, _pen(st, m), go, ge, cube, w, threads=8),
                      np.int64).reshape(len(q), len(r))


def _bincount_rows(mat, max_score):
    out = np.zeros((mat.shape[0], max_score + 2), np.uint64)
    for i, row in enumerate(mat):
        inside = row[row <= max_score]
        out[i, :max_score + 1] = np.bincount(inside, minlength=max_score + 1)
        out[i, max_score + 1] = np.count_nonzero(row > max_score)
    return out


@pytest.mark.parametrize("case", ["blosum_central", "unit", "blosum_wide_bins"])
def test_histogram_equals_oracle_bincount(gpu, oracle_lib, case):
    st = gpu
    from seqtree_amd.evalue import score_histogram

    rng = random.Random({"blosum_central": 1, "unit": 2, "blosum_wide_bins": 3}[case])
    q = cdr3_like(rng, 150, 6, 30) + [rand_seq(rng, 0, 40) for _ in range(30)] + [""]
    r = cdr3_like(rng, 900, 6, 30) + [rand_seq(rng, 0, 40) for _ in range(100)]
    if case == "unit":
        m, go, prior, max_score = None, 1, None, 12
    elif case == "blosum_central":
        m, go, prior, max_score = st.SubstitutionMatrix.blosum62(), 11, st.central_prior(3), 60
    else:
        m, go, prior, max_score = st.SubstitutionMatrix.blosum62(), 11, None, 5000
    got = score_histogram(q, r, max_score, matrix=m, gap_open=go, gap_prior=prior)
    want = _bincount_rows(_oracle_matrix(st, oracle_lib, q, r, m, go, 1, prior), max_score)
    assert got.dtype == np.uint64 and got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    assert (got.sum(axis=1) == len(r)).all()


def _variant_case(st, variant):
    rng = random.Random(seed_of(variant))
    blosum = st.SubstitutionMatrix.blosum62()
    if variant == "twin":          # short queries, 16-bit range proof holds
        q, r = cdr3_like(rng, 301, 6, 20), cdr3_like(rng, 3000, 6, 20)
        return q, r, blosum, 28, 1, st.central_prior(21), 90
    if variant == "short":         # a large prior breaks the 16-bit proof, not the 32-bit one
        q, r = cdr3_like(rng, 200, 6, 20), cdr3_like(rng, 3000, 6, 20)
        return q, r, blosum, 28, 1, st.central_prior(4000), 400
    if variant == "profile":       # queries past 32 residues
        q = [rand_seq(rng, 20, 60) for _ in range(120)]
        r = [rand_seq(rng, 20, 90) for _ in range(2500)]
        return q, r, blosum, 28, 1, None, 1000
    if variant == "table":         # queries too long for an LDS profile
        q = [rand_seq(rng, 500, 700) for _ in range(5)]
        r = [rand_seq(rng, 480, 720) for _ in range(150)] + q[:2]
        return q, r, blosum, 28, 1, None, 1000
    # exact: int32 sums can wrap (a prior near INT32_MAX); wrapped scores land below zero
    seqs = [rand_seq(rng, 3, 14) for _ in range(70)]
    return seqs, seqs, blosum, 28, 1, (lambda i, d, m: (2**31 - 1) - 7 * i), 1000


@pytest.mark.parametrize("variant", ["twin", "short", "profile", "table", "exact"])
def test_fused_histogram_in_every_kernel_variant(gpu, oracle_lib, variant):
    """Each scoring kernel adds its scores to LDS histograms (no score matrix written); the
    raw device layout, below-zero bin included, equals the oracle matrix's counts."""
    st = gpu
    from seqtree_amd import _core
    from seqtree_amd.gapblock import _prior_cube

    q, r, m, go, ge, prior, H = _variant_case(st, variant)
    w = max(len(x) for x in q + r)
    cube = _prior_cube(prior, w) if prior is not None else []
    got, labels = launch_labels(
        lambda: _core.gapblock_histogram(q, r, H, "aa", m, go, ge, cube, w))
    label = {"twin": "gapblock_histogram_twin", "short": "gapblock_histogram_short",
             "profile": "gapblock_histogram_profile", "table": "gapblock_histogram_table",
             "exact": "gapblock_histogram_profile_exact"}[variant]
    assert label in labels, labels
    assert "score_histogram" not in labels and "gapblock" not in labels
    mat = _oracle_matrix(st, oracle_lib, q, r, m, go, ge, prior)
    want = np.zeros((len(q), H + 2), np.uint64)
    for i, row in enumerate(mat):
        want[i, 0] = np.count_nonzero(row < 0)
        inside = row[(row >= 0) & (row < H)]
        want[i, 1:H + 1] = np.bincount(inside, minlength=H)
        want[i, H + 1] = np.count_nonzero(row >= H)
    np.testing.assert_array_equal(got, want)
    if variant == "exact":
        assert want[:, 0].any(), "the case must produce wrapped (negative) scores"


def test_wrapped_scores_raise_in_the_python_api(gpu):
    from seqtree_amd.evalue import score_histogram

    rng = random.Random(71)
    seqs = [rand_seq(rng, 3, 14) for _ in range(40)]
    with pytest.raises(OverflowError, match="below zero"):
        score_histogram(seqs, seqs, 50, matrix=None, gap_open=1,
                        gap_prior=lambda i, d, m: (2**31 - 1) - 7 * i)


def test_histogram_across_row_blocks(gpu, oracle_lib):
    """More than 65535 queries in one call (the fused kernels index rows in 64 bits; the
    reduction pass for wide histograms splits them into blocks)."""
    st = gpu
    from seqtree_amd.evalue import score_histogram

    rng = random.Random(9)
    q = cdr3_like(rng, 70_000, 8, 20)
    r = cdr3_like(rng, 40, 8, 20)
    mat = _oracle_matrix(st, oracle_lib, q, r, None, 1, 1, None)
    for ceiling in (30, 2000):  # fused, then dense row blocks and the reduction pass
        got = score_histogram(q, r, ceiling, matrix=None, gap_open=1)
        np.testing.assert_array_equal(got, _bincount_rows(mat, ceiling))


def test_empty_sets(gpu):
    from seqtree_amd.evalue import score_histogram

    assert score_histogram([], ["CASS"], 4).shape == (0, 6)
    h = score_histogram(["CASS", ""], [], 4)
    assert h.shape == (2, 6) and not h.any()


def test_counts_and_cutoffs_equal_the_threshold_hit_lists(gpu):
    """Counts and cutoffs from histograms equal those from the sparse threshold path's hits."""
    st = gpu
    from seqtree_amd import gapblock
    from seqtree_amd.evalue import (
        GapblockBall, neighbour_counts, thetas_from_scores, threshold_for_evalue,
    )

    rng = random.Random(17)
    m = st.SubstitutionMatrix.blosum62()
    control = cdr3_like(rng, 4000, 8, 20)
    queries = control[:40] + cdr3_like(rng, 40, 8, 20)
    ball = GapblockBall(max_penalty=40, matrix=m, gap_open=11)
    qi, _, sc = gapblock.score_threshold(queries, control, 40, matrix=m, gap_open=11)
    per_q = [sc[qi == i].tolist() for i in range(len(queries))]
    for ex in (False, True):
        got = neighbour_counts(queries, control, ball, exclude_exact=ex)
        want = [sum(1 for s in p if not (ex and s == 0)) for p in per_q]
        assert got.tolist() == want
        for e in (0.5, 2.0, 8.0, 60.0):
            th = threshold_for_evalue(400, control, queries, ball, e, exclude_exact=ex)
            assert th == thetas_from_scores(per_q, 400, len(control), e, 40, exclude_exact=ex)


# ---- the reference's behaviour tests, over the exhaustive unit-cost (Hamming) ball ----------

def _rand(n, rng, length=14):
    return ["".join(rng.choice(AA) for _ in range(length)) for _ in range(n)]


def _mutate(seq, k, rng):
    s = list(seq)
    for pos in rng.sample(range(len(seq)), k):
        s[pos] = rng.choice(AA)
    return "".join(s)


def _clustered(seeds, per_seed, rng, max_subs=4):
    return [_mutate(s, rng.randint(1, max_subs), rng) for s in seeds for _ in range(per_seed)]


def _ball(theta):
    from seqtree_amd.evalue import GapblockBall

    return GapblockBall(max_penalty=theta, matrix=None)


def test_planted_cluster_is_significant_background_is_not(gpu):
    from seqtree_amd.evalue import evalues

    rng = random.Random(0)
    control = _rand(20000, rng)
    q = "".join(rng.choice(AA) for _ in range(14))
    cluster = []
    for _ in range(50):
        s = list(q)
        s[rng.randrange(14)] = rng.choice(AA)
        cluster.append("".join(s))
    target = _rand(20000, rng) + cluster
    r = evalues(target, control, [q], _ball(1))[0]
    assert r["n_target"] >= 40
    assert r["p_enrichment"] < 1e-3
    qb = "".join(rng.choice(AA) for _ in range(14))
    assert evalues(target, control, [qb], _ball(1))[0]["p_enrichment"] > 0.01


def test_exclude_exact_drops_self_hit(gpu):
    from seqtree_amd.evalue import evalues

    rng = random.Random(7)
    q = "CASSPGTEAFF"
    mut = [_mutate(q, 1, rng) for _ in range(2)]
    target = list(dict.fromkeys([q] + mut))
    control = _rand(5000, rng)
    full = evalues(target, control, [q], _ball(1), exclude_exact=False)[0]
    punc = evalues(target, control, [q], _ball(1), exclude_exact=True)[0]
    assert punc["n_target"] == full["n_target"] - 1


def test_evalue_monotone_in_scope(gpu):
    from seqtree_amd.evalue import evalues

    rng = random.Random(1)
    db = _rand(10000, rng)
    control = _rand(10000, rng)
    e1 = evalues(db, control, [db[0]], _ball(1))[0]["E"]
    e2 = evalues(db, control, [db[0]], _ball(3))[0]["E"]
    assert e2 >= e1


def test_threshold_round_trips_through_evalues(gpu):
    """The returned cutoff is the largest with E <= e_target: one more overshoots."""
    from seqtree_amd.evalue import evalues, threshold_for_evalue

    rng = random.Random(3)
    seeds = _rand(40, rng)
    control = _clustered(seeds, 100, rng)
    target = _clustered(seeds, 10, rng)
    queries = [_mutate(s, 1, rng) for s in seeds[:12]]
    theta_max, e_target = 8, 1.0
    thetas = threshold_for_evalue(target, control, queries, _ball(theta_max), e_target)
    checked = 0
    for q, th in zip(queries, thetas):
        if th < 1:
            continue
        assert evalues(target, control, [q], _ball(th))[0]["E"] <= e_target
        if th < theta_max:
            assert evalues(target, control, [q], _ball(th + 1))[0]["E"] > e_target
        checked += 1
    assert checked >= 5


def test_calibrated_cutoff_varies_across_queries(gpu):
    from seqtree_amd.evalue import threshold_for_evalue

    rng = random.Random(5)
    hub = "CASSLGQAYEQYF"
    dense = [_mutate(hub, rng.randint(1, 2), rng) for _ in range(2000)]
    control = dense + _rand(2000, rng, len(hub))
    target = _rand(500, rng, len(hub))
    rare = _rand(1, rng, len(hub))[0]
    th_hub, th_rare = threshold_for_evalue(target, control, [hub, rare], _ball(6), e_target=1.0)
    assert th_hub < th_rare
