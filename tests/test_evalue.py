"""Host-side E-value math (reference: python/seqtree/evalue.py). The known answers are the
reference's own (tests/python/test_evalue.py); the histogram inversion is checked against the
score-list inversion and against a brute-force reading of its definition on random inputs;
the Poisson tail against scipy's."""
from __future__ import annotations

import math
import random

import numpy as np
import pytest
from scipy import stats

from seqtree_amd.evalue import (
    GapblockBall, _evalue_rows, _poisson_sf, evalues, score_histogram, thetas_from_histograms,
    thetas_from_scores, threshold_for_evalue,
)


def test_poisson_sf_basics():
    assert _poisson_sf(0, 5.0) == 1.0
    assert _poisson_sf(1, 0.0) == 0.0
    assert abs(_poisson_sf(1, 1.0) - (1 - math.e ** -1)) < 1e-6


@pytest.mark.parametrize("lam", [0.05, 0.3, 1.0, 2.5, 7.0, 20.0])
def test_poisson_sf_is_the_upper_tail(lam):
    """P(X >= k) = sf(k - 1) for the Poisson law; the complementary CDF loses absolute
    precision only near 1e-16, so compare absolutely."""
    for k in range(0, 40):
        want = 1.0 if k <= 0 else float(stats.poisson.sf(k - 1, lam))
        assert abs(_poisson_sf(k, lam) - want) < 1e-12, (k, lam)


def test_thetas_from_scores_known_answers():
    assert thetas_from_scores([[2, 5, 9]], 100, 100, 1.0, 10) == [4]
    assert thetas_from_scores([[2, 5, 9]], 100, 100, 2.0, 10) == [8]
    assert thetas_from_scores([[2, 5, 9]], 100, 100, 10.0, 10) == [10]
    assert thetas_from_scores([[]], n_target=100, m_control=1000, e_target=0.2,
                              theta_max=9) == [-1]
    assert thetas_from_scores([[]], n_target=100, m_control=1000, e_target=0.3,
                              theta_max=9) == [9]
    assert thetas_from_scores([[5, 5, 5]], n_target=100, m_control=500, e_target=0.2,
                              theta_max=9) == [-1]
    assert thetas_from_scores([[0, 0, 4]], n_target=100, m_control=100, e_target=1.0,
                              theta_max=9) == [-1]
    assert thetas_from_scores([[7]], n_target=100, m_control=100, e_target=0.5,
                              theta_max=9) == [-1]


def test_theta_is_monotone_in_e_target():
    prev = None
    for e in (8.0, 4.0, 2.0, 1.0, 0.5):
        (th,) = thetas_from_scores([[1, 3, 3, 7, 12, 20]], 10, 100, e, 30)
        assert prev is None or th <= prev
        prev = th


@pytest.mark.parametrize("fn", [thetas_from_scores, thetas_from_histograms])
def test_inversion_rejects_bad_arguments(fn):
    arg = [[1]] if fn is thetas_from_scores else np.zeros((1, 8), np.uint64)
    with pytest.raises(ValueError, match="e_target"):
        fn(arg, 10, 10, 0.0, 5)
    with pytest.raises(ValueError, match="n_target and m_control"):
        fn(arg, 0, 10, 1.0, 5)
    with pytest.raises(ValueError, match="theta_max"):
        fn(arg, 10, 10, 1.0, -1)


def _hist(scores_per_query, width):
    h = np.zeros((len(scores_per_query), width), np.uint64)
    for i, sc in enumerate(scores_per_query):
        for s in sc:
            if s < width:
                h[i, s] += 1
    return h


def _brute_theta(scores, N, M, e, theta_max, exclude_exact):
    """The definition read literally: the largest t in [0, theta_max] with E(t) <= e, where
    E(t) = (N/M) * count(<= t), or 3N/M when that count is 0; -1 if there is none."""
    s = [x for x in scores if not (exclude_exact and x == 0)]
    best = -1
    for t in range(theta_max + 1):
        c = sum(1 for x in s if x <= t)
        # k = floor(e * M / N) bounds the count; the empty ball needs c_max >= 3
        c_max = e * M / N
        if c <= int(c_max) and (c > 0 or c_max >= 3.0) and int(c_max) >= 1:
            best = t
        if c > int(c_max):
            break
    return best


def test_histogram_inversion_equals_the_score_list_inversion():
    rng = random.Random(5)
    for trial in range(400):
        theta_max = rng.randint(0, 25)
        N, M = rng.randint(1, 500), rng.randint(1, 3000)
        e = rng.choice([0.05, 0.2, 0.5, 1.0, 2.0, 3.5, 10.0, 40.0])
        scores = [[rng.randint(0, theta_max + 3) for _ in range(rng.randint(0, 30))]
                  for _ in range(6)] + [[0, 0, 0], [], [theta_max] * 5]
        for ex in (False, True):
            want = thetas_from_scores(scores, N, M, e, theta_max, exclude_exact=ex)
            got = thetas_from_histograms(_hist(scores, theta_max + 1), N, M, e, theta_max,
                                         exclude_exact=ex)
            assert list(got) == want, (trial, ex)
            assert want == [_brute_theta(s, N, M, e, theta_max, ex) for s in scores], trial


def test_score_list_inversion_with_negative_and_out_of_range_scores():
    """Scores below zero count in every cumulative count; scores above theta_max never do;
    exclude_exact drops exact zeros only. Checked against the literal definition."""
    rng = random.Random(11)
    for trial in range(300):
        theta_max = rng.randint(0, 12)
        N, M = rng.randint(1, 200), rng.randint(1, 2000)
        e = rng.choice([0.1, 0.5, 1.0, 3.0, 8.0, 30.0])
        scores = [[rng.randint(-3, theta_max + 4) for _ in range(rng.randint(0, 20))]
                  for _ in range(5)]
        for ex in (False, True):
            want = [_brute_theta(s, N, M, e, theta_max, ex) for s in scores]
            assert thetas_from_scores(scores, N, M, e, theta_max, exclude_exact=ex) == want, trial


def test_poisson_sf_point_masses_against_closed_form():
    """Small k: the lower tail is exp(-lam) * sum_{i<k} lam^i / i!, evaluated here in exact
    rational arithmetic scaled by exp(-lam)."""
    from fractions import Fraction

    for lam in (0.1, 0.9, 3.0, 12.5):
        for k in range(1, 15):
            lam_q = Fraction(lam)
            lower = sum(lam_q ** i / math.factorial(i) for i in range(k))
            want = 1.0 - math.exp(-lam) * float(lower)
            assert abs(_poisson_sf(k, lam) - max(0.0, want)) < 1e-13, (k, lam)


def test_evalue_rows():
    rows = _evalue_rows([5, 0], [0, 4], N=100, M=1000)
    assert rows[0]["rule_of_three"] and rows[0]["E"] == pytest.approx(0.3)
    assert rows[0]["p_any"] == pytest.approx(1 - math.exp(-0.3))
    assert rows[0]["p_enrichment"] == pytest.approx(float(stats.poisson.sf(4, 0.3)), abs=1e-15)
    assert not rows[1]["rule_of_three"] and rows[1]["E"] == pytest.approx(0.4)
    assert rows[1]["p_enrichment"] == 1.0 and rows[1]["n_control"] == 4


def test_argument_errors_raise_before_any_device_work():
    ball = GapblockBall(max_penalty=1)
    with pytest.raises(ValueError, match="control index is empty"):
        evalues(["CASS"], [], ["CASSPGTEAFF"], ball)
    with pytest.raises(ValueError, match="max_penalty"):
        threshold_for_evalue(["CASS"], ["CASS"], ["CASS"], GapblockBall(max_penalty=0), 1.0)
    with pytest.raises(ValueError, match="max_score"):
        score_histogram(["CASS"], ["CASS"], 20000)
