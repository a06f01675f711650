"""Pins the CPU oracle (oracle/seqtree_oracle.c) before it is trusted as the GPU checker.

Three independent anchors, chosen so a plausible mistake (a dropped term, a swapped operand,
an off-by-one in a gap run, a wrong boundary) fails at least one:

* the reference's own C++ compiled from its sources (oracle/_ref), on random inputs across
  matrices, gap settings, priors and lengths (skipped where the reference is not built);
* the reference's known answers (tests/python/test_pairwise.py:112-163, 213-234, 321-326;
  test_distance.py:40-71; test_matrix.cpp:13-38; test_gapblock.py:534-536; docstrings);
* brute force: every gap-block layout enumerated, every global alignment enumerated (local =
  best global alignment of any pair of substrings), and Levenshtein by breadth-first search
  over single edits.
"""
from __future__ import annotations

import functools
import random

import numpy as np
import pytest

from conftest import AA, rand_seq, seed_of

AA24 = "ARNDCQEGHILKMFPSTWYVBZX*"


def _tables(reference, name):
    sim, pen, scale = reference.tables(name)
    return sim, pen, scale


# ----------------------------------------------------------------------------- brute force

def brute_block(q, r, pen, go, ge, prior=None):
    """Every single contiguous gap-block layout (the block sits in the longer sequence)."""
    m, n = len(q), len(r)
    d = abs(m - n)
    shorter, longer = min(m, n), max(m, n)
    best = None
    for i in range(shorter + 1):
        head = sum(pen[q[j]][r[j]] for j in range(i))
        if m >= n:
            tail = sum(pen[q[j + d]][r[j]] for j in range(i, n))
        else:
            tail = sum(pen[q[j]][r[j + d]] for j in range(i, m))
        pr = prior(i, d, longer) if (prior is not None and d > 0) else 0
        cand = head + tail + pr
        best = cand if best is None else min(best, cand)
    return best + (0 if d == 0 else go + (d - 1) * ge)


def alignments(m, n):
    """All global alignments of lengths m and n as column strings over {'M','X','Y'}
    (M: both residues, X: query residue vs gap, Y: gap vs ref residue)."""
    if m == 0 and n == 0:
        yield ""
        return
    if m and n:
        for rest in alignments(m - 1, n - 1):
            yield rest + "M"
    if m:
        for rest in alignments(m - 1, n):
            yield rest + "X"
    if n:
        for rest in alignments(m, n - 1):
            yield rest + "Y"


def score_alignment(cols, q, r, sim, go, ge):
    """A gap run continues only with the same gap type; X then Y is two opens."""
    i = j = 0
    total, prev = 0, None
    for c in cols:
        if c == "M":
            total += sim[q[i]][r[j]]
            i += 1
            j += 1
        elif c == "X":
            total -= ge if prev == "X" else go
            i += 1
        else:
            total -= ge if prev == "Y" else go
            j += 1
        prev = c
    return total


@functools.lru_cache(maxsize=None)
def _aligns(m, n):
    return tuple(alignments(m, n))


def brute_global(q, r, sim, go, ge):
    return max(score_alignment(a, q, r, sim, go, ge) for a in _aligns(len(q), len(r)))


def brute_local(q, r, sim, go, ge):
    best = 0
    for a0 in range(len(q) + 1):
        for a1 in range(a0, len(q) + 1):
            for b0 in range(len(r) + 1):
                for b1 in range(b0, len(r) + 1):
                    best = max(best, brute_global(q[a0:a1], r[b0:b1], sim, go, ge))
    return best


def bfs_edit_distance(a, b, alphabet="AC"):
    """Fewest single-character edits from a to b, by breadth-first search."""
    if a == b:
        return 0
    seen = {a}
    frontier = [a]
    dist = 0
    limit = len(a) + len(b) + 1
    while frontier and dist < limit:
        dist += 1
        nxt = []
        for s in frontier:
            cands = set()
            for p in range(len(s) + 1):
                for ch in alphabet:
                    cands.add(s[:p] + ch + s[p:])
            for p in range(len(s)):
                cands.add(s[:p] + s[p + 1:])
                for ch in alphabet:
                    cands.add(s[:p] + ch + s[p + 1:])
            for c in cands:
                if c == b:
                    return dist
                if c not in seen and len(c) <= max(len(a), len(b)) + 1:
                    seen.add(c)
                    nxt.append(c)
        frontier = nxt
    raise AssertionError("unreachable")


# ----------------------------------------------------------------------------- matrices

def test_gram_transform_and_scale_anchors(oracle_lib):
    b62 = np.zeros((24, 24), np.int32)
    # anchors of the reference's tests (test_matrix.cpp:13-38, test_matrix.py:89-101) are
    # checked on the product tables in test_host.py; here the transform itself is pinned on a
    # hand-made grid whose penalties follow from the definition by hand
    grid = np.array([[4, -1, 0], [-1, 5, 2], [0, 2, 3]], np.int32)
    pen = oracle_lib.pen_from_sim(grid)
    assert pen.tolist() == [[0, 11, 7], [11, 0, 4], [7, 4, 0]]
    # clamp at zero: an off-diagonal above both diagonals
    assert oracle_lib.pen_from_sim(np.array([[1, 5], [5, 1]], np.int32)).tolist() == [[0, 0], [0, 0]]
    # wrapping int32: 2^30 + 2^30 - 2 * (-2^30) wraps to 0 in int32
    big = np.array([[2**30, -(2**30)], [-(2**30), 2**30]], np.int64).astype(np.int32)
    assert oracle_lib.pen_from_sim(big)[0, 1] == 0
    # the median of the off-diagonal penalties is the element of rank n/2
    pen4 = np.array([[0, 1, 2, 3], [1, 0, 4, 5], [2, 4, 0, 6], [3, 5, 6, 0]], np.int32)
    assert oracle_lib.scale(pen4) == sorted([1, 2, 3, 1, 4, 5, 2, 4, 6, 3, 5, 6])[6]
    assert oracle_lib.scale(b62[:1, :1]) == 0


@pytest.mark.parametrize("name", ["blosum62", "blosum45", "blosum80", "pam250", "pam100",
                                  "structural"])
def test_pen_and_scale_match_reference(oracle_lib, reference, name):
    sim, pen, scale = _tables(reference, name)
    assert (oracle_lib.pen_from_sim(sim) == pen).all()
    assert oracle_lib.scale(pen) == scale


# ----------------------------------------------------------------------------- gap-block

def _pen_dict(pen, alphabet=AA24):
    return {a: {b: int(pen[i][j]) for j, b in enumerate(alphabet)} for i, a in enumerate(alphabet)}


def test_gapblock_cell_against_layout_enumeration(oracle_lib):
    rng = random.Random(7)
    pen = np.abs(np.random.default_rng(7).integers(0, 25, size=(24, 24))).astype(np.int32)
    np.fill_diagonal(pen, 0)
    pd = _pen_dict(pen)
    gap_costs = [(1, 1), (4, 1), (28, 1), (0, 3), (2, 5), (5, 5), (0, 0)]
    central = lambda i, d, m: 0 if d == 0 else int(21 * abs(2 * i + d - m) // 2)  # noqa
    from seqtree_amd.gapblock import _prior_cube

    for prior in (None, central):
        q = [rand_seq(rng, 1, 22) for _ in range(24)]
        r = [rand_seq(rng, 1, 22) for _ in range(24)]
        w = max(len(s) for s in q + r)
        for go, ge in gap_costs:
            cube = _prior_cube(prior, w) if prior else None
            got = oracle_lib.gapblock_matrix(q, r, pen, go, ge, cube, w)
            for i, a in enumerate(q):
                for k, b in enumerate(r):
                    assert got[i, k] == brute_block(a, b, pd, go, ge, prior), (a, b, go, ge)


def test_gapblock_known_answer_unit_cost(oracle_lib):
    pen = (np.arange(4)[:, None] != np.arange(4)[None, :]).astype(np.int32)
    got = oracle_lib.gapblock_matrix(["ACGT", "ACTT"], ["ACGT"], pen, 1, 1, alphabet="nt")
    assert got.tolist() == [[0], [1]]


@pytest.mark.parametrize("name", ["blosum62", "pam100", "structural"])
def test_gapblock_matches_reference(oracle_lib, reference, name):
    from seqtree_amd.gapblock import _prior_cube

    sim, pen, scale = _tables(reference, name)
    rng = random.Random(seed_of(name))
    q = [rand_seq(rng, 0, 30) for _ in range(50)]
    r = [rand_seq(rng, 0, 30) for _ in range(70)]
    w = max(len(s) for s in q + r)
    priors = [None,
              lambda i, d, m: 0 if d == 0 else int(21 * abs(2 * i + d - m) // 2),
              lambda i, d, m: 0 if d == 0 else 7 * abs(i - 4),
              lambda i, d, m: 0 if d == 0 else (0 if i in (3, 4, m - d - 4, m - d - 3) else 1 << 20)]
    for prior in priors:
        cube = _prior_cube(prior, w) if prior else []
        for go, ge in ((2 * scale, 1), (0, 3), (5, 5)):
            a = oracle_lib.gapblock_matrix(q, r, pen, go, ge, cube if prior else None, w,
                                           threads=4)
            b = reference.gapblock_matrix(q, r, name, go, ge, cube, w, threads=4)
            assert (a == b).all()


# ----------------------------------------------------------------------------- Gotoh

def _sim_dict(sim, alphabet=AA24):
    return {a: {b: int(sim[i][j]) for j, b in enumerate(alphabet)} for i, a in enumerate(alphabet)}


@pytest.mark.parametrize("go,ge", [(11, 1), (5, 5), (1, 3), (0, 0), (2, 7), (0, 4)])
def test_gotoh_against_alignment_enumeration(oracle_lib, go, ge):
    rng = np.random.default_rng(go * 31 + ge)
    grid = rng.integers(-6, 9, size=(24, 24)).astype(np.int32)  # asymmetric on purpose
    sd = _sim_dict(grid)
    r = random.Random(go * 131 + ge)
    for _ in range(60):
        a, b = rand_seq(r, 0, 5), rand_seq(r, 0, 5)
        assert oracle_lib.gotoh(a, b, grid, False, go, ge) == brute_global(a, b, sd, go, ge)
    for _ in range(25):
        a, b = rand_seq(r, 0, 4), rand_seq(r, 0, 4)
        assert oracle_lib.gotoh(a, b, grid, True, go, ge) == brute_local(a, b, sd, go, ge)


def test_gotoh_known_answers(oracle_lib, reference):
    sim, _, _ = _tables(reference, "blosum62")
    g = lambda q, r, **kw: oracle_lib.gotoh(q, r, sim, **kw)  # noqa: E731
    for L in (1, 2, 3, 5):
        assert g("AAA", "AAA" + "A" * L, go=11, ge=1) == 12 - (11 + (L - 1))
    for L in (1, 2, 3):
        assert g("AAA", "AAA" + "A" * L, go=5, ge=5) == 12 - 5 * L
    assert g("AAA", "KKAAA", go=11, ge=1) == 0
    assert g("WWW", "KKK", local=True) == 0
    assert g("WWWAAAWWW", "KKKAAAKKK", local=True) == 12
    assert g("CASSLGQAYEQYF", "CASSPGQAYEQF") == 45
    assert g("CASSLGQAYEQYF", "CASSPGQAYEQF", go=12) == 44
    q, r = "CASSLGQAYEQYF", "CASSPGQAYEQF"
    assert g(q, r, go=10**8) == -99_999_944
    assert g(q, r, go=2**30) == -1_073_741_768
    assert g(q, r, go=2**31 - 1) == -2_147_483_591
    assert g("", "") == 0 and g("", "AAA") == -13 and g("", "AAA", local=True) == 0
    code, _ = oracle_lib.align_matrix(["A" * 100], ["C"], sim, False, False, 10**8, 10**8)
    assert code == 2


@pytest.mark.parametrize("local", [False, True])
@pytest.mark.parametrize("dist", [False, True])
@pytest.mark.parametrize("name,go,ge", [("blosum62", 11, 1), ("blosum45", 5, 5),
                                        ("blosum80", 1, 3), ("pam250", 0, 0),
                                        ("structural", 100, 5), ("blosum62", 10**8, 1)])
def test_gotoh_matches_reference(oracle_lib, reference, name, go, ge, local, dist):
    sim, _, _ = _tables(reference, name)
    rng = random.Random(seed_of(name, go, ge, local, dist))
    q = [rand_seq(rng, 0, 28) for _ in range(30)]
    r = [rand_seq(rng, 0, 28) for _ in range(35)]
    code, a = oracle_lib.align_matrix(q, r, sim, local, dist, go, ge, threads=4)
    try:
        b = reference.align_matrix(q, r, name, local, dist, go, ge, threads=4)
        assert code == 0 and (a == b).all()
    except Exception as e:  # the reference raised: the oracle must report the same overflow
        assert getattr(e, "code", None) == 2 and code == 2


def test_gotoh_custom_asymmetric_matches_reference(oracle_lib, reference):
    grid = np.random.default_rng(3).integers(-9, 12, size=(24, 24)).astype(np.int32)
    rng = random.Random(3)
    q = [rand_seq(rng, 0, 40) for _ in range(20)]
    r = [rand_seq(rng, 0, 40) for _ in range(25)]
    for local in (False, True):
        for dist in (False, True):
            code, a = oracle_lib.align_matrix(q, r, grid, local, dist, 7, 2)
            b = reference.align_matrix(q, r, grid, local, dist, 7, 2)
            assert code == 0 and (a == b).all()


# ----------------------------------------------------------------------------- distances

def test_levenshtein_against_bfs(oracle_lib):
    rng = random.Random(2)
    for _ in range(150):
        a, b = rand_seq(rng, 0, 4, "AC"), rand_seq(rng, 0, 4, "AC")
        assert oracle_lib.levenshtein(a, b) == bfs_edit_distance(a, b)


def test_distance_known_answers(oracle_lib):
    for a, b, d in [("kitten", "sitting", 3), ("flaw", "lawn", 2),
                    ("CASSLGQAYEQYF", "CASSPGQAYEQF", 2), ("", "", 0), ("abc", "", 3),
                    ("", "abc", 3), ("abc", "abc", 0), ("ABC", "abc", 3)]:
        assert oracle_lib.levenshtein(a, b) == d
    assert oracle_lib.hamming("ABC", "abc") == 3
    assert oracle_lib.hamming("AAAA", "ATAT") == 2
    assert oracle_lib.hamming("AAAA", "AAA") == -1
    _, m = oracle_lib.distance_matrix("hamming", ["AAAA", "AAAT"], ["AAAA", "TTTT"])
    assert m.tolist() == [[0, 4], [1, 3]]


def test_distances_match_reference(oracle_lib, reference):
    rng = random.Random(9)
    a = [rand_seq(rng, 0, 70, "ACGT") for _ in range(40)]
    b = [rand_seq(rng, 0, 70, "ACGT") for _ in range(45)]
    assert (oracle_lib.distance_matrix("levenshtein", a, b)[1]
            == reference.distance_matrix("levenshtein", a, b)).all()
    h = [rand_seq(rng, 12, 12, "ACGT") for _ in range(60)]
    assert (oracle_lib.distance_matrix("hamming", h, h[:33])[1]
            == reference.distance_matrix("hamming", h, h[:33])).all()


def test_brute_helpers_are_not_vacuous():
    """The enumerators must see every layout: Delannoy numbers and a hand-checked case."""
    assert [len(_aligns(n, n)) for n in range(5)] == [1, 3, 13, 63, 321]
    sd = {a: {b: (4 if a == b else -1) for b in "AK"} for a in "AK"}
    assert brute_global("AAA", "KKAAA", sd, 11, 1) == 12 - 12
    assert brute_local("KAAK", "AA", sd, 11, 1) == 8


# -------------------------------------------------------------------------- traceback (8f-2)

def aligned_strings(q, r, ops, qb, rb):
    """The aligned strings an op string spells out from residues q[qb:], r[rb:]."""
    aq, ar, i, j = [], [], qb, rb
    for o in ops:
        if o in "MS":
            aq.append(q[i]); ar.append(r[j]); i += 1; j += 1
        elif o == "D":
            aq.append(q[i]); ar.append("-"); i += 1
        else:
            aq.append("-"); ar.append(r[j]); j += 1
    return "".join(aq), "".join(ar)


def rescore(aq, ar, sd, go, ge):
    """Gotoh score of an explicit alignment (the reference's own test re-scorer, restated)."""
    total, prev = 0, None
    for x, y in zip(aq, ar):
        if x == "-":
            total -= ge if prev == "I" else go
            prev = "I"
        elif y == "-":
            total -= ge if prev == "D" else go
            prev = "D"
        else:
            total += sd[x][y]
            prev = None
    return total


@pytest.mark.parametrize("local", [False, True])
@pytest.mark.parametrize("name,go,ge", [("blosum62", 11, 1), ("blosum62", 5, 5),
                                        ("pam250", 1, 3), ("blosum45", 0, 0),
                                        ("blosum62", 1, 10), ("structural", 2, 100)])
def test_align_pair_matches_reference(oracle_lib, reference, name, go, ge, local):
    """Score, ops and aligned strings equal the compiled reference's align_pair, including
    tie-heavy inputs (few residue kinds) where the back-pointer order decides the path."""
    sim, _, _ = _tables(reference, name)
    rng = random.Random(seed_of(name, go, ge, local))
    cases = [(rand_seq(rng, 0, 22), rand_seq(rng, 0, 22)) for _ in range(70)]
    cases += [(rand_seq(rng, 1, 14, "AAC"), rand_seq(rng, 1, 14, "AAC")) for _ in range(70)]
    cases += [("", ""), ("", "ACD"), ("WYK", ""), ("W", "W"), ("WWW", "KKK")]
    for q, r in cases:
        s, ops, qb, rb = oracle_lib.align_pair(q, r, sim, local, go, ge)
        want = reference.align_pair(q, r, name, local, go, ge)
        assert (s, ops) == (want[0], want[3]), (q, r)
        assert aligned_strings(q, r, ops, qb, rb) == (want[1], want[2]), (q, r)


def test_align_pair_known_answer_and_rescore(oracle_lib, reference):
    """The reference's docstring example (python/seqtree/pairwise.py:109-112) and its test that
    a traceback re-scores to the reported score (tests/python/test_pairwise.py:194-212)."""
    sim, _, _ = _tables(reference, "blosum62")
    q, r = "CASSLGQAYEQYF", "CASSPGQAYEQF"
    s, ops, qb, rb = oracle_lib.align_pair(q, r, sim)
    assert aligned_strings(q, r, ops, qb, rb) == ("CASSLGQAYEQYF", "CASSPGQAYEQ-F")
    assert s == 45
    sd = _sim_dict(sim)
    rng = random.Random(17)
    for local in (False, True):
        for go, ge in [(11, 1), (5, 5), (1, 3), (0, 0), (1, 10), (2, 100)]:
            for _ in range(30):
                a, b = rand_seq(rng, 4, 20), rand_seq(rng, 4, 20)
                s, ops, qb, rb = oracle_lib.align_pair(a, b, sim, local, go, ge)
                aq, ar = aligned_strings(a, b, ops, qb, rb)
                assert rescore(aq, ar, sd, go, ge) == s
                assert s == oracle_lib.gotoh(a, b, sim, local, go, ge)
                if not local:
                    assert aq.replace("-", "") == a and ar.replace("-", "") == b


# ------------------------------------------------------------ Index::align (penalty space)

def _penalty_of(cols, q, r, pen, go, ge):
    """Penalty of an explicit global alignment: substitution penalties plus go for the first
    column of each gap run and ge for every further one (X then Y opens a new run)."""
    return -score_alignment(cols, q, r, {a: {b: -v for b, v in row.items()}
                                         for a, row in pen.items()}, go, ge)


def _unit_dict(alphabet=AA24):
    return {a: {b: int(a != b) for b in alphabet} for a in alphabet}


def _ops_to_cols(ops):
    return ops.replace("S", "M").replace("D", "X").replace("I", "Y")


INDEX_REFS = ["CASSLGQAYEQYF", "CASSPGTGVYGYTF", "CASRQGAWDTQYF", "CASSFRSSYNEQFF",
              "CAWSVSGGGTDTQYF", "CASSLAPGATNEKLFF", "CSARDRTGNGYTF", "CASSIRSSYEQYF"]


def test_index_align_against_alignment_enumeration(oracle_lib, reference):
    """The score is the minimum penalty over every global alignment, and the returned ops
    spell an alignment that costs exactly that (unit cost and BLOSUM62 penalties)."""
    _, pen62, _ = _tables(reference, "blosum62")
    rng = random.Random(23)
    for pen, pd in ((None, _unit_dict()), (pen62, _pen_dict(pen62))):
        for go, ge in [(1, 1), (3, 1), (11, 1), (1, 4), (0, 0), (2, 0)]:
            for _ in range(25):
                q, r = rand_seq(rng, 0, 5, "ACDK"), rand_seq(rng, 0, 5, "ACDK")
                s, ops = oracle_lib.index_align(q, r, pen, go, ge)
                best = min(_penalty_of(c, q, r, pd, go, ge) for c in _aligns(len(q), len(r)))
                assert s == best, (q, r, go, ge)
                assert _penalty_of(_ops_to_cols(ops), q, r, pd, go, ge) == s, (q, r, ops)
                i = j = 0
                for o in ops:  # 'M' exactly where the aligned residues are equal
                    if o in "MS":
                        assert (o == "M") == (q[i] == r[j]), (q, r, ops)
                    i += o in "MSD"
                    j += o in "MSI"
                assert (i, j) == (len(q), len(r))


@pytest.mark.parametrize("matrix,go,ge", [(None, 1, 1), ("blosum62", 11, 1), ("blosum62", 14, 1),
                                          ("blosum62", 1, 1), ("structural", 7, 3),
                                          ("pam250", 0, 0), (None, 2, 5), ("blosum45", 40, 1)])
def test_index_align_matches_reference(oracle_lib, reference, matrix, go, ge):
    """Score and ops equal the compiled reference's Index::align, tie-heavy inputs included
    (few residue kinds, where the strict-< state order decides among equal paths)."""
    pen = None if matrix is None else _tables(reference, matrix)[1]
    rng = random.Random(seed_of("index", matrix, go, ge))
    qs = [rand_seq(rng, 0, 24) for _ in range(60)] + \
        [rand_seq(rng, 1, 14, "AAC") for _ in range(60)] + ["", "W", "WWW", "KAK"]
    rs = [rand_seq(rng, 0, 24) for _ in range(30)] + \
        [rand_seq(rng, 1, 14, "AAC") for _ in range(30)] + ["", "W", "KKK"]
    pq = list(range(len(qs))) * 2
    pr = [rng.randrange(len(rs)) for _ in pq]
    want = reference.index_align_pairs(qs, rs, pq, pr, matrix, go, ge)
    for a, b, w in zip(pq, pr, want):
        s, ops = oracle_lib.index_align(qs[a], rs[b], pen, go, ge)
        assert (s, ops) == (w[0], w[3]), (qs[a], rs[b])
        # the reference's aligned strings are the raw residues in op order
        assert aligned_strings(qs[a], rs[b], ops, 0, 0) == (w[1], w[2])


def test_index_align_known_answers(oracle_lib, reference):
    """The reference's own Index.align tests: tests/python/test_engine_guards.py:50-80,
    test_matrix.py:28-40 and test_sanity.py:86-98."""
    _, pen62, scale62 = _tables(reference, "blosum62")
    ref0 = INDEX_REFS[0]
    cheap = oracle_lib.index_align("CASSLGYEQYF", ref0, pen62, 11, 1)
    dear = oracle_lib.index_align("CASSLGYEQYF", ref0, pen62, 11, 9)
    assert cheap[0] < dear[0] and dear[0] - cheap[0] == 8
    lin = oracle_lib.index_align("CASSLGYEQYF", ref0, None, 1, 1)
    assert lin[0] == 2 and lin[1].count("I") + lin[1].count("D") == 2
    degenerate = oracle_lib.index_align("CASSIRSSYEQYF", ref0, pen62, 1, 1)
    sane = oracle_lib.index_align("CASSIRSSYEQYF", ref0, pen62, scale62, 1)
    assert degenerate[1].count("I") > 0 and sane[1] == "M" * 4 + "S" * 4 + "M" * 5
    want = {None: 2, "structural": 18, "pam250": 24, "blosum62": 28, "pam100": 30}
    for name, expected in want.items():
        pen = None if name is None else _tables(reference, name)[1]
        assert oracle_lib.index_align("CASSLAPGATNEKLFF", "CASSLELGATNEKLFF", pen, 100,
                                      1)[0] == expected
    assert oracle_lib.index_align("CASSPGQAYEQF", "CASSPGQAYEQF", pen62, 1, 1) == (0, "M" * 12)
    assert oracle_lib.index_align("CASSPGQAYEQY", "CASSPGQAYEQF", pen62, 2 * scale62,
                                  1)[1] == "M" * 11 + "S"
    assert oracle_lib.index_align("", "", None, 1, 1) == (0, "")
    assert oracle_lib.index_align("", "ACD", None, 3, 2) == (7, "III")
