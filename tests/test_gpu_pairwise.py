"""GPU parity for Gotoh NW/SW score and distance matrices (reference: src/pairwise.cpp and
tests/python/test_pairwise.py). The checker is the CPU oracle; known answers come from the
reference's tests and docstrings."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import cdr3_like, launch_labels, rand_seq, seed_of

pytestmark = pytest.mark.gpu

SETTINGS = [
    ("global", 11, 1), ("global", 12, 1), ("global", 5, 5), ("global", 1, 1),
    ("global", 0, 0), ("global", 100, 5), ("global", 1, 3), ("global", 0, 5),
    ("local", 1, 10), ("global", 10**8, 1), ("global", 2**30, 1),
    ("local", 11, 1), ("local", 5, 5), ("local", 1, 1), ("local", 12, 1),
    # past the int16 range proof but inside int32: the one-query int32 register kernel
    ("global", 1000, 3), ("local", 2000, 7),
]


def _sim(st, m, alphabet="aa"):
    syms = st.alphabet_symbols(alphabet)
    if m.size() != len(syms):
        raise ValueError
    return np.array([[m.similarity(a, b) for b in syms] for a in syms], np.int32)


def _check(st, oracle_lib, q, r, m, mode, go, ge, dist=False, alphabet="aa", sim=None):
    fn = st.pairwise.dist_matrix if dist else st.pairwise.score_matrix
    code, want = oracle_lib.align_matrix(q, r, _sim(st, m, alphabet) if sim is None else sim,
                                         mode == "local", dist, go, ge, alphabet=alphabet,
                                         threads=8)
    if code == 2:
        with pytest.raises(OverflowError, match="distance" if dist else "score"):
            fn(q, r, m, mode=mode, gap_open=go, gap_extend=ge, alphabet=alphabet)
        return None
    got = np.asarray(fn(q, r, m, mode=mode, gap_open=go, gap_extend=ge, alphabet=alphabet))
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} cells differ, first {bad[:3].tolist()}"
    return got


@pytest.mark.parametrize("mode,go,ge", SETTINGS)
@pytest.mark.parametrize("name", ["blosum62", "blosum45", "blosum80"])
def test_settings_grid(gpu, oracle_lib, name, mode, go, ge):
    st = gpu
    m = getattr(st.SubstitutionMatrix, name)()
    rng = random.Random(seed_of(name, mode, go, ge))
    q = [rand_seq(rng, 0, 26) for _ in range(23)] + ["A", "W" * 3]
    r = [rand_seq(rng, 0, 26) for _ in range(1100)] + ["", "KKAAA"]
    _check(st, oracle_lib, q, r, m, mode, go, ge)


@pytest.mark.parametrize("mode", ["global", "local"])
@pytest.mark.parametrize("name", ["blosum62", "pam250", "pam100", "structural"])
def test_distance_matrices(gpu, oracle_lib, name, mode):
    st = gpu
    m = getattr(st.SubstitutionMatrix, name)()
    rng = random.Random(seed_of(name, mode))
    seqs = cdr3_like(rng, 40, 8, 25)
    d = _check(st, oracle_lib, seqs, seqs, m, mode, 11, 1, dist=True)
    assert (np.diag(d) == 0).all() and (d >= 0).all() and (d == d.T).all()
    q = [rand_seq(rng, 0, 30) for _ in range(17)]
    r = [rand_seq(rng, 0, 30) for _ in range(1500)]
    _check(st, oracle_lib, q, r, m, mode, 11, 1, dist=True)


@pytest.mark.parametrize("longest", [64, 160])
def test_distance_self_scores_in_lds(gpu, oracle_lib, longest):
    """Sets of at most 160 residues take the LDS self-score kernel (rows sized to the longest
    string): lengths 0 and up to the limit, an asymmetric matrix and gap_extend > gap_open
    included."""
    st = gpu
    rng = random.Random(23 + longest)
    grid = np.random.default_rng(23).integers(-7, 9, size=(24, 24)).astype(np.int32)
    for m, sim in ((st.SubstitutionMatrix.blosum62(), None),
                   (st.SubstitutionMatrix.from_similarity(grid.tolist()), grid)):
        q = [rand_seq(rng, longest - 14, longest) for _ in range(40)] + [""]
        r = [rand_seq(rng, 0, longest) for _ in range(300)] + ["A" * longest]
        for mode in ("global", "local"):
            for go, ge in ((11, 1), (2, 5)):
                _check(st, oracle_lib, q, r, m, mode, go, ge, dist=True, sim=sim)


def test_known_answers(gpu):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    score = st.pairwise.score
    for L in (1, 2, 3, 5):
        assert score("AAA", "AAA" + "A" * L, m, gap_open=11, gap_extend=1) == 12 - (11 + L - 1)
        if L <= 3:
            assert score("AAA", "AAA" + "A" * L, m, gap_open=5, gap_extend=5) == 12 - 5 * L
    assert score("AAA", "KKAAA", m, gap_open=11, gap_extend=1) == 0
    assert score("WWW", "KKK", m, mode="local") == 0
    assert score("WWWAAAWWW", "KKKAAAKKK", m, mode="local") == 12
    assert score("WWWAAAWWW", "KKKAAAKKK", m, mode="global") < 0
    s = "CASSLGQAYEQYF"
    assert score(s, s, m, mode="local") == score(s, s, m, mode="global")
    assert score("CASSLGQAYEQYF", "CASSPGQAYEQF", m) == 45
    assert score("CASSLGQAYEQYF", "CASSPGQAYEQF", m, gap_open=12) == 44
    assert score("AAA", "AAA", m) == 12 and score("AAA", "AAAAA", m) == 0
    assert score("", "", m) == 0
    assert score("", "AAA", m) == -(11 + 2)
    assert score("", "AAA", m, mode="local") == 0


def test_huge_gap_costs_take_the_int64_path(gpu):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    q, r = "CASSLGQAYEQYF", "CASSPGQAYEQF"
    assert st.pairwise.score(q, r, m, gap_open=10**8) == -99_999_944
    assert st.pairwise.score(q, r, m, gap_open=2**30) == -1_073_741_768
    assert st.pairwise.score(q, r, m, gap_open=2**31 - 1) == -2_147_483_591
    with pytest.raises(OverflowError, match="32 bits"):
        st.pairwise.score("A" * 400, "W" * 600, m, gap_open=2**31 - 1, gap_extend=2**31 - 1)


def test_overflow_raises_with_the_right_quantity(gpu):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    with pytest.raises(OverflowError):
        st.pairwise.score_matrix(["A" * 100], ["C"], m, gap_open=10**8, gap_extend=10**8)
    with pytest.raises(OverflowError):
        st.pairwise.dist_matrix(["A" * 100], ["C"], m, gap_open=10**8, gap_extend=10**8)
    with pytest.raises(OverflowError, match="distance"):
        st.pairwise.dist_matrix(["AAA"], ["C"], m, gap_open=2**31 - 1, gap_extend=2**31 - 1)


def test_long_queries_and_references(gpu, oracle_lib):
    """Queries past 64 residues run the multi-stripe register kernel; long references stay on
    the register kernels."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    rng = random.Random(17)
    q = [rand_seq(rng, 0, 64) for _ in range(10)] + [rand_seq(rng, 60, 130) for _ in range(6)]
    r = [rand_seq(rng, 0, 200) for _ in range(300)]
    for mode in ("global", "local"):
        _check(st, oracle_lib, q, r, m, mode, 12, 1)
        _check(st, oracle_lib, q, r, m, mode, 12, 1, dist=True)


STRIPE_CASES = [
    # (gap_open, gap_extend, label): int16 pairs in the H form, int32 in the H form, int32 in
    # the U/V form (gap_open < gap_extend), and int64 past the int32 proof
    (11, 1, "gotoh_i16x2_stripes"), (1000, 3, "gotoh_i32_stripes"),
    (1, 3, "gotoh_i32_stripes"), (10**7, 1, "gotoh_generic"),
]


@pytest.mark.parametrize("mode", ["global", "local"])
@pytest.mark.parametrize("go,ge,label", STRIPE_CASES)
def test_long_queries_run_in_stripes(gpu, oracle_lib, mode, go, ge, label):
    """Queries past one register stripe (32 rows): lengths on and around the stripe edges,
    germline-V-like lengths (~90) and a few long ones, mixed with short queries in one call,
    against references up to 260 residues; scores and distances equal the oracle's."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    rng = random.Random(seed_of(mode, go, ge))
    edges = [31, 32, 33, 63, 64, 65, 95, 96, 97, 128, 129]
    q = [rand_seq(rng, L, L) for L in edges] + [rand_seq(rng, 82, 98) for _ in range(24)]
    q += [rand_seq(rng, 150, 250) for _ in range(3)] + [rand_seq(rng, 0, 30) for _ in range(8)]
    r = [rand_seq(rng, 0, 120) for _ in range(500)] + [rand_seq(rng, 200, 260) for _ in range(8)]
    r += q[:6]
    # the 33 long queries meet the reference set 18 times over, so their pairs pass 2^18 and the
    # lane-per-pair kernel runs (fewer go to the one-wave-per-pair kernel): the oracle checks
    # the distinct columns, and every repeat must equal its original
    reps = 18
    for dist in (False, True):
        want = _check(st, oracle_lib, q, r, m, mode, go, ge, dist=dist)
        if want is None:  # the overflow error, checked by _check
            continue
        fn = st.pairwise.dist_matrix if dist else st.pairwise.score_matrix
        got, labels = launch_labels(lambda: np.asarray(fn(q, r * reps, m, mode=mode,
                                                          gap_open=go, gap_extend=ge)))
        if not dist:
            assert any(x.startswith(label) for x in labels), labels
        got = got.reshape(len(q), reps, len(r))
        assert all(np.array_equal(got[:, k], want) for k in range(reps))


def test_germline_v_all_vs_all(gpu, oracle_lib):
    """The reference benchmark's germline shape (bench/bench_pairwise.py: lengths 90 +- 8,
    BLOSUM62 11/1), all against all, global and local, scores and distances."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    rng = random.Random(90)
    xs = [rand_seq(rng, 82, 98) for _ in range(200)]
    for mode in ("global", "local"):
        got = _check(st, oracle_lib, xs, xs, m, mode, 11, 1)
        d = _check(st, oracle_lib, xs, xs, m, mode, 11, 1, dist=True)
        assert (np.diag(d) == 0).all() and (got == got.T).all()


def test_asymmetric_custom_matrix(gpu, oracle_lib):
    """sim is indexed [query code][ref code]; the transposed int64 path must use sim^T."""
    st = gpu
    rng = np.random.default_rng(5)
    grid = rng.integers(-7, 9, size=(24, 24)).astype(np.int32)
    m = st.SubstitutionMatrix.from_similarity(grid.tolist())
    rr = random.Random(5)
    q = [rand_seq(rr, 0, 80) for _ in range(20)]
    r = [rand_seq(rr, 0, 40) for _ in range(200)] + [rand_seq(rr, 70, 90) for _ in range(20)]
    for mode in ("global", "local"):
        for go, ge in ((11, 1), (3, 7), (10**7, 3)):
            _check(st, oracle_lib, q, r, m, mode, go, ge, sim=grid)
            _check(st, oracle_lib, q, r, m, mode, go, ge, dist=True, sim=grid)


def test_nucleotide_unit_matrix(gpu, oracle_lib):
    st = gpu
    m = st.SubstitutionMatrix.unit(4)
    rng = random.Random(21)
    q = [rand_seq(rng, 0, 40, "ACGT") for _ in range(30)]
    r = [rand_seq(rng, 0, 40, "ACGT") for _ in range(400)]
    sim = np.eye(4, dtype=np.int32)
    for mode in ("global", "local"):
        _check(st, oracle_lib, q, r, m, mode, 2, 1, alphabet="nt", sim=sim)


def test_cdr3_workload_slice(gpu, oracle_lib):
    """A 200-row slice of the 20,000 x 20,000 CDR3-like workload (lengths 8-25)."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    rng = random.Random(99)
    q = cdr3_like(rng, 200, 8, 25)
    r = cdr3_like(rng, 20000, 8, 25)
    for mode in ("global", "local"):
        _check(st, oracle_lib, q, r, m, mode, 11, 1)


def test_bad_arguments(gpu):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    with pytest.raises(ValueError, match="gap_open"):
        st.pairwise.score("AA", "AA", m, gap_open=-1)
    with pytest.raises(ValueError, match="mode"):
        st.pairwise.score("AA", "AA", m, mode="semiglobal")
    with pytest.raises(ValueError, match="alphabet|symbol"):
        st.pairwise.score("AA#", "AA", m)
    assert st.pairwise.score_matrix([], ["A"], m).shape == (0, 1)


# ------------------------------------------------ few long pairs: one wave per pair (8f-4)

def _sim_table(st, m):
    syms = st.amino_acids()
    return np.array([[m.similarity(a, b) for b in syms] for a in syms], np.int32)


@pytest.mark.parametrize("mode", ["global", "local"])
def test_long_pairs_run_one_wave_per_pair(gpu, oracle_lib, mode):
    """Queries of 500-3,000 residues against references of every length class, few enough
    pairs for the intra-pair wavefront kernel; scores and distances (self-scores of long
    strings included) against the oracle's int64 DP."""
    st = gpu
    rng = random.Random(seed_of("wave", mode))
    aa = "ACDEFGHIKLMNPQRSTVWY"
    qs = [rand_seq(rng, 500, 3000, aa) for _ in range(4)] + \
        [rand_seq(rng, L, L, aa) for L in (65, 512, 513, 1025, 2048, 2049, 4097)]
    rs = [rand_seq(rng, L, L, aa) for L in (1, 5, 63, 64, 65, 300)] + \
        [rand_seq(rng, 1000, 3000, aa) for _ in range(3)]
    m = st.SubstitutionMatrix.blosum62()
    sim = _sim_table(st, m)
    for go, ge in ((11, 1), (5, 5), (3, 2)):
        got, labels = launch_labels(lambda: np.asarray(
            st.pairwise.score_matrix(qs, rs, m, mode=mode, gap_open=go, gap_extend=ge)))
        assert f"gotoh_wave_{mode}" in labels, labels
        code, want = oracle_lib.align_matrix(qs, rs, sim, mode == "local", False, go, ge,
                                             threads=8)
        assert code == 0 and np.array_equal(got, np.asarray(want).reshape(got.shape)), (go, ge)
    got, labels = launch_labels(lambda: np.asarray(
        st.pairwise.dist_matrix(qs, rs, m, mode=mode)))
    assert f"gotoh_wave_self_{mode}" in labels and f"gotoh_wave_{mode}" in labels, labels
    code, want = oracle_lib.align_matrix(qs, rs, sim, mode == "local", True, 11, 1, threads=8)
    assert code == 0 and np.array_equal(got, np.asarray(want).reshape(got.shape))


def test_long_pairs_wave_equals_the_lane_per_pair_kernel(gpu):
    """The same long queries scored by both long-query kernels: few references send them to
    the wavefront kernel, many to the lane-per-pair multi-stripe kernel; the rows agree."""
    st = gpu
    rng = random.Random(12)
    aa = "ACDEFGHIKLMNPQRSTVWY"
    qs = [rand_seq(rng, 90, 700, aa) for _ in range(6)]
    rs = [rand_seq(rng, 8, 400, aa) for _ in range(50_000)]  # 300,000 pairs: multi-stripe
    m = st.SubstitutionMatrix.pam250()
    for mode in ("global", "local"):
        big, labels = launch_labels(lambda: np.asarray(
            st.pairwise.score_matrix(qs, rs, m, mode=mode)))
        assert f"gotoh_i16x2_stripes_{mode}" in labels or f"gotoh_i32_stripes_{mode}" in labels
        few, labels = launch_labels(lambda: np.asarray(
            st.pairwise.score_matrix(qs, rs[:3000], m, mode=mode)))
        assert f"gotoh_wave_{mode}" in labels, labels
        assert np.array_equal(big[:, :3000], few)
