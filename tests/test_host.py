"""Host-side logic of the product, no GPU needed: the C-ABI library and its exports, the
codec and substitution matrices, the gap priors and prior cube, the Python scalar
gap-block scorer, and argument validation (order, exception types and messages) -- which the
C++ API performs before any device work, exactly as the reference does."""
from __future__ import annotations

import ctypes
import random
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import AA, ROOT, rand_seq

HEADER = ROOT / "include" / "seqtree_hip.h"
HIPLIB = ROOT / "seqtree_amd" / "libseqtree_hip.so"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(st_\w+)\s*\(", text, flags=re.M)


def test_c_abi_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(HIPLIB))
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    lib.st_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.st_version()


def test_c_abi_rejects_bad_alphabet_without_a_device():
    lib = ctypes.CDLL(str(HIPLIB))

    class Status(ctypes.Structure):
        _fields_ = [("code", ctypes.c_int32), ("message", ctypes.c_char * 256)]

    st = Status()
    lib.st_align_matrix.restype = ctypes.c_int32
    off = (ctypes.c_uint64 * 2)(0, 1)
    code = (ctypes.c_uint8 * 1)(0)
    out = (ctypes.c_int32 * 1)()
    sim = (ctypes.c_int32 * 1)(0)
    rc = lib.st_align_matrix(code, off, ctypes.c_uint64(1), code, off, ctypes.c_uint64(1), sim,
                             ctypes.c_uint32(33), 0, 0, 11, 1, out, ctypes.byref(st))
    assert rc == 1 and st.code == 1 and b"alphabet size" in st.message
    # empty sides succeed without touching a device
    rc = lib.st_align_matrix(code, off, ctypes.c_uint64(0), code, off, ctypes.c_uint64(1), sim,
                             ctypes.c_uint32(24), 0, 0, 11, 1, out, ctypes.byref(st))
    assert rc == 0


# ----------------------------------------------------------------------------- matrices

def test_alphabets(st):
    assert st.amino_acids() == "ARNDCQEGHILKMFPSTWYVBZX*"
    assert st.alphabet_symbols("nt") == "ACGT"
    assert st.alphabet_symbols("iupac") == "ACGTRYSWKMBDHVN"
    with pytest.raises(ValueError, match="unknown alphabet"):
        st.alphabet_symbols("xyz")


def test_matrix_anchors(st):
    """Known answers of the reference's tests (test_matrix.cpp:13-38, test_matrix.py:89-101,
    test_pairwise.py:294-307, 364-379; scale(B62) == 14 at _bindings.cpp:293)."""
    SM = st.SubstitutionMatrix
    b62 = SM.blosum62()
    assert b62.size() == 24
    assert b62.penalty("A", "R") == 11 and b62.penalty("W", "Y") == 14
    assert b62.penalty("A", "E") + b62.penalty("P", "L") == 28
    assert SM.pam250().penalty("A", "E") == 6 and SM.pam100().penalty("A", "E") == 9
    assert SM.structural().penalty("A", "E") == 6
    assert (SM.blosum45().similarity("A", "A"), SM.blosum45().similarity("W", "W")) == (5, 15)
    assert (b62.similarity("A", "A"), b62.similarity("W", "W")) == (4, 11)
    assert (SM.blosum80().similarity("A", "A"), SM.blosum80().similarity("W", "W")) == (7, 16)
    assert (b62.similarity("C", "C"), b62.similarity("W", "W")) == (9, 11)
    assert [SM.blosum62().scale(), SM.blosum45().scale(), SM.blosum80().scale(),
            SM.pam250().scale(), SM.pam100().scale(), SM.structural().scale()] == \
        [14, 16, 24, 12, 17, 7]
    u = SM.unit(4)
    assert u.size() == 4
    for residue in ("C", "G", "T"):
        with pytest.raises(ValueError, match="out of range"):
            u.similarity(residue, residue)
        with pytest.raises(ValueError, match="out of range"):
            u.penalty(residue, residue)
    with pytest.raises(ValueError):
        b62.penalty("L", "AB")
    with pytest.raises(ValueError):
        b62.similarity("#", "A")
    with pytest.raises(ValueError, match="1..32"):
        SM.from_similarity([])
    with pytest.raises(ValueError, match="square"):
        SM.from_similarity([[1, 2], [3]])


@pytest.mark.parametrize("name", ["blosum62", "blosum45", "blosum80", "pam250", "pam100",
                                  "structural"])
def test_builtin_tables_equal_the_reference(st, reference, name):
    sim, pen, scale = reference.tables(name)
    m = getattr(st.SubstitutionMatrix, name)()
    syms = st.amino_acids()
    assert [[m.similarity(a, b) for b in syms] for a in syms] == sim.tolist()
    assert [[m.penalty(a, b) for b in syms] for a in syms] == pen.tolist()
    assert m.scale() == scale


def test_structural_regenerates_from_the_mj_energies(st):
    """The structural grid follows from the Miyazawa-Jernigan contact energies by the
    published construction (tests/golden/mj_structural.py)."""
    from golden.mj_structural import structural_grid

    m = st.SubstitutionMatrix.structural()
    syms = st.amino_acids()
    assert [[m.similarity(a, b) for b in syms] for a in syms] == structural_grid()


def test_from_similarity_matches_the_reference(st, reference):
    rng = np.random.default_rng(1)
    for size in (1, 4, 24, 32):
        grid = rng.integers(-2**31, 2**31, size=(size, size), dtype=np.int64).astype(np.int32)
        for g in (grid, (grid % 20).astype(np.int32)):
            m = st.SubstitutionMatrix.from_similarity(g.tolist())
            sim, pen, scale = reference.tables(g)
            assert m.scale() == scale
            if size <= 24:
                syms = st.amino_acids()[:size]
                assert [[m.penalty(a, b) for b in syms] for a in syms] == pen.tolist()


# ----------------------------------------------------------------------------- priors

def test_prior_known_answers(st):
    g = st.gapblock
    assert g.gap_cost(0, 1, 9) == 0 and g.gap_cost(1, 11, 1) == 11 and g.gap_cost(3, 11, 1) == 13
    with pytest.raises(ValueError):
        g.gap_cost(-1, 1, 1)
    c = g.central_prior(21)
    assert c(0, 1, 6) == 52 and c(0, 2, 6) == 42
    for m in range(6, 20):
        for d in range(0, 5):
            for i in range(0, m - d + 1):
                old = 0 if d == 0 else int(21 * abs(2 * i - (m - d)) // 2)
                assert c(i, d, m) == old
    w = [0.9, 0.8, 0.1, 0.0, 0.1, 0.8, 0.9]
    p = g.profile_prior(100, lambda j, m: w[j])
    assert p(2, 3, 7) == int(100 * (0.1 + 0.0 + 0.1))
    assert p(0, 3, 7) == int(100 * (0.9 + 0.8 + 0.1))
    assert p(0, 0, 7) == 0
    assert g.profile_prior(10, [0.0, 1.0, 0.0])(1, 1, 3) == 10
    assert g.profile_prior(10, [0.0, 1.0, 0.0])(0, 1, 3) == 0
    pos = g.positions_prior((3, 4, -4, -3))
    for i in range(13):
        assert pos(i, 2, 14) == (0 if i in {3, 4, 8, 9} else g.UNREACHABLE)
    clamp = g.positions_prior((9, -9))
    assert [i for i in range(5) if clamp(i, 2, 6) == 0] == [0, 4]
    with pytest.raises(ValueError, match="at least one start"):
        g.positions_prior(())
    for factory in (g.central_prior, lambda x: g.profile_prior(x, lambda j, m: 1.0),
                    lambda x: g.frame_prior(x, 4)):
        with pytest.raises(ValueError):
            factory(-1)
    with pytest.raises(ValueError, match="c must be"):
        g.frame_prior(1, -1)


def test_prior_cube_layout(st):
    from seqtree_amd.gapblock import _prior_cube

    prior = lambda i, d, m: 1000 * m + 10 * d + i  # noqa: E731
    w = 6
    cube = _prior_cube(prior, w)
    w1 = w + 1
    assert len(cube) == w1 ** 3
    for m in range(w1):
        for d in range(w1):
            for i in range(w1):
                v = cube[(m * w1 + d) * w1 + i]
                assert v == (prior(i, d, m) if (1 <= d <= m and i <= m - d) else 0)
    with pytest.raises(ValueError, match="negative"):
        _prior_cube(lambda i, d, m: -1, 3)


def _brute_block(q, r, pen, go, ge, prior=None):
    m, n = len(q), len(r)
    d = abs(m - n)
    best = None
    for i in range(min(m, n) + 1):
        head = sum(pen[q[j], r[j]] for j in range(i))
        tail = (sum(pen[q[j + d], r[j]] for j in range(i, n)) if m >= n
                else sum(pen[q[j], r[j + d]] for j in range(i, m)))
        cand = head + tail + (prior(i, d, max(m, n)) if prior and d else 0)
        best = cand if best is None else min(best, cand)
    return best + (0 if d == 0 else go + (d - 1) * ge)


def test_scalar_gapblock_score_against_enumeration(st):
    m = st.SubstitutionMatrix.blosum62()
    pen = {(a, b): m.penalty(a, b) for a in AA for b in AA}
    rng = random.Random(7)
    gap_costs = [(1, 1), (4, 1), (28, 1), (0, 3), (2, 5), (5, 5), (0, 0)]
    for prior in (None, st.central_prior(20), st.gapblock.profile_prior(40, lambda j, n: abs(j - (n - 1) / 2) / n)):
        for _ in range(600):
            q, r = rand_seq(rng, 1, 22), rand_seq(rng, 1, 22)
            go, ge = rng.choice(gap_costs)
            s, i = st.gapblock_score(q, r, m, go, ge, prior, _pen=pen)
            assert s == _brute_block(q, r, pen, go, ge, prior)
            d = abs(len(q) - len(r))
            if d and prior is None:  # the reported start reproduces the score
                head = sum(pen[q[j], r[j]] for j in range(i))
                tail = (sum(pen[q[j + d], r[j]] for j in range(i, len(r))) if len(q) >= len(r)
                        else sum(pen[q[j], r[j + d]] for j in range(i, len(q))))
                assert head + tail + st.gap_cost(d, go, ge) == s
    assert st.gapblock_score("CASSLGQAYEQYF", "CASSLGQAYEQYF", m) == (0, 0)
    assert st.gapblock_score("CAST", "CAST", None)[0] == 0
    with pytest.raises(ValueError):
        st.gapblock_score("CAST", "CAT", m, gap_open=-1)


# ----------------------------------------------------------------------------- validation

def test_gapblock_validation_order_and_messages(st):
    m = st.SubstitutionMatrix.blosum62()
    core = st.gapblock._gapblock_matrix
    with pytest.raises(ValueError, match="gap_open and gap_extend must be >= 0"):
        core(["CA#"], ["CA"], "aa", m, -1, 1, [], 0, 0)  # gap check wins over the bad symbol
    with pytest.raises(ValueError, match="matrix size does not match the alphabet"):
        core(["ACGT"], ["ACGT"], "nt", m, 1, 1, [], 0, 0)
    with pytest.raises(ValueError, match="symbol '#' is not in the alphabet"):
        core(["CAS#"], ["CASS"], "aa", m, 28, 1, [], 0, 0)
    with pytest.raises(ValueError, match=r"\(prior_width \+ 1\)\^3"):
        core(["CAS"], ["CA"], "aa", m, 28, 1, [0] * 7, 3, 0)
    with pytest.raises(ValueError, match="longer than the prior table's width"):
        core(["CASSS"], ["CA"], "aa", m, 28, 1, [0] * 64, 3, 0)
    with pytest.raises(ValueError, match="unknown alphabet"):
        core(["CAS"], ["CA"], "xyz", m, 28, 1, [], 0, 0)
    # an empty side returns before symbols or the prior are looked at (gapblock.cpp:79-81)
    assert core([], ["CA#"], "aa", m, 28, 1, [1], 5, 0).shape == (0, 1)
    assert core(["CA#"], [], "aa", m, 28, 1, [1], 5, 0).shape == (1, 0)
    with pytest.raises(ValueError, match="gap_open"):
        st.score_matrix(["CAS"], ["CAS"], m, gap_open=-1)
    with pytest.raises(ValueError, match="negative"):
        st.score_matrix(["CAS"], ["CA"], m, 28, gap_prior=lambda i, d, n: -1)


def test_pairwise_validation_order_and_messages(st):
    m = st.SubstitutionMatrix.blosum62()
    pw = st.pairwise
    with pytest.raises(ValueError, match="gap_open and gap_extend must be >= 0"):
        pw.score_matrix(["A#"], ["A"], m, gap_open=-1)
    with pytest.raises(ValueError, match="matrix size does not match"):
        pw.score_matrix(["ACGT"], ["ACGT"], m, alphabet="nt")
    with pytest.raises(ValueError, match="invalid symbol '#' in sequence"):
        pw.dist_matrix(["AC#"], ["AC"], m)
    with pytest.raises(ValueError, match="mode"):
        pw.score("AA", "AA", m, mode="semiglobal")
    with pytest.raises(ValueError, match="gap_open"):
        pw.score("AA", "AA", m, gap_open=-1)
    with pytest.raises(ValueError, match="symbol"):
        pw.score("AA#", "AA", m)
    assert pw.score_matrix([], ["A#"], m).shape == (0, 1)
    assert pw.dist_matrix(["A"], [], m).shape == (1, 0)


def test_distance_validation_and_empty(st):
    d = st.distance
    with pytest.raises(ValueError, match=r"equal-length sequences \(got 3 and 4\)"):
        d.hamming_matrix(["AAAA", "AAA"], ["AAAA"])
    with pytest.raises(ValueError, match=r"\(got 4 and 3\)"):
        d.hamming_matrix(["AAAA"], ["AAAA", "AAA"])
    with pytest.raises(ValueError, match="equal-length"):
        d.hamming("AAAA", "AAA")
    assert np.asarray(d.hamming_matrix([], ["A"])).shape == (0, 1)
    assert np.asarray(d.hamming_matrix(["A"], [])).shape == (1, 0)
    assert np.asarray(d.levenshtein_matrix([], [])).shape == (0, 0)
    assert set(d.__all__) == {"hamming", "levenshtein", "hamming_matrix", "levenshtein_matrix"}


def test_empty_score_matrix_buffer(st):
    sm = st.score_matrix([], ["ACGT"])
    mv = memoryview(sm)
    assert mv.format == "i" and mv.shape == (0, 1)
    assert len(sm) == 0 and repr(sm) == "ScoreMatrix(0, 1)"
    with pytest.raises(IndexError):
        sm.row(0)
    with pytest.raises(IndexError):
        sm[0, 0]


def test_compute_without_a_device_fails_loudly(st):
    if st.device_count() > 0:
        pytest.skip("a device is present; the loud-failure path is for GPU-less hosts")
    with pytest.raises(RuntimeError, match="no HIP device"):
        st.score_matrix(["CAS"], ["CA"], st.SubstitutionMatrix.blosum62(), 28)
    with pytest.raises(RuntimeError, match="no HIP device"):
        st.distance.levenshtein_matrix(["A"], ["B"])


def test_index_and_search_params_validation(st):
    """Index.build / Index.align / SearchParams raise as the reference's binding does
    (src/trie.cpp:60-66, src/index.cpp:109-113, src/_bindings.cpp:42-121, 352-397), before
    any device work."""
    with pytest.raises(ValueError, match=r"invalid symbol '1' in reference #1"):
        st.Index.build(["CASS", "CA1S"])
    idx = st.Index.build(["CASSLGQAYEQYF", "CASSPGTGVYGYTF"], "aa")
    assert len(idx) == 2 and idx.ref_seq(1) == "CASSPGTGVYGYTF"
    with pytest.raises(IndexError):
        idx.ref_seq(2)
    with pytest.raises(ValueError, match="unknown mode 'local'"):
        st.SearchParams(mode="local")
    with pytest.raises(ValueError, match="unknown engine"):
        st.SearchParams(engine="fast")
    with pytest.raises(ValueError, match="unknown matrix 'foo'"):
        st.SearchParams(matrix="foo")
    with pytest.raises(TypeError):
        st.SearchParams(matrix=3)
    p = st.SearchParams(matrix="BLOSUM62", gap_open=11)
    assert (p.matrix, p.gap_open, p.gap_extend, p.engine, p.mode) == ("BLOSUM62", 11, 1, "auto",
                                                                       "all")
    p.matrix = st.SubstitutionMatrix.blosum62()
    assert isinstance(p.matrix, st.SubstitutionMatrix)
    with pytest.raises(ValueError, match="invalid symbol '!' in query"):
        idx.align(0, "CASS!", st.SearchParams(gap_open=-1))
    with pytest.raises(ValueError, match=">= 0"):
        idx.align(0, "CASS", st.SearchParams(gap_open=-1))
    with pytest.raises(ValueError, match="matrix size does not match"):
        idx.align(0, "CASS", st.SearchParams(matrix=st.SubstitutionMatrix.unit(4)))
    with pytest.raises(IndexError):
        idx.align(2, "CASS", st.SearchParams())
    with pytest.raises(ValueError, match="requires the amino-acid alphabet"):
        st.Index.build(["ACGT"], "nt").align(0, "ACG", st.SearchParams(matrix="BLOSUM62"))
    with pytest.raises(ValueError, match="differ in length"):
        idx.align_batch([0, 1], ["CASS"], st.SearchParams())
    assert idx.align_batch([], [], st.SearchParams()) == []


def test_pair_list_validation(st):
    m = st.SubstitutionMatrix.blosum62()
    with pytest.raises(ValueError, match="indexes past"):
        st.gapblock.score_pairs(["CAS"], ["CA"], [0, 1], [0, 0], m, 28)
    with pytest.raises(ValueError, match="same length"):
        st.gapblock.score_pairs(["CAS"], ["CA"], [0, 0], [0], m, 28)
    with pytest.raises(ValueError, match="symbol '#'"):
        st.gapblock.score_threshold(["CA#"], ["CA"], 10, m, 28)
    assert len(st.gapblock.score_pairs(["CAS"], ["CA"], [], [], m, 28)) == 0
    qi, ri, sc = st.gapblock.score_threshold([], ["CA"], 10, m, 28)
    assert len(qi) == len(ri) == len(sc) == 0


def test_batch_encoded_set_validation():
    from seqtree_amd.batch import EncodedSet, _encode

    syms = "ACGT"
    codes, off = _encode(EncodedSet(np.array([0, 1, 2, 3], np.uint8),
                                    np.array([0, 2, 4], np.uint64)), syms)
    assert list(off) == [0, 2, 4] and list(codes) == [0, 1, 2, 3]
    with pytest.raises(ValueError, match="outside the alphabet"):
        _encode(EncodedSet(np.array([0, 4], np.uint8), np.array([0, 2], np.uint64)), syms)
    with pytest.raises(ValueError, match="malformed"):
        _encode(EncodedSet(np.array([0, 1], np.uint8), np.array([0, 3], np.uint64)), syms)
    codes, off = _encode(EncodedSet(np.zeros(0, np.uint8), np.zeros(1, np.uint64)), syms)
    assert list(off) == [0]


def test_synthetic_generators_are_seeded_and_shaped():
    from seqtree_amd import synthetic as syn

    a = syn.families(np.random.default_rng(1), 50, 4, 10, 20)
    b = syn.families(np.random.default_rng(1), 50, 4, 10, 20)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    codes, off, fam = a
    lens = np.diff(off.astype(np.int64))
    assert lens.min() >= 10 and lens.max() <= 20 and len(fam) == 200
    seqs = syn.decode(codes, off)
    assert all(s[0] == "C" and s[-1] == "F" for s in seqs)
    assert set("".join(seqs)) <= set(syn.BG_RESIDUES)
    c2, o2, l2 = syn.cdr3_csr(np.random.default_rng(2), 100, 6, 30)
    assert np.array_equal(np.diff(o2.astype(np.int64)), l2) and l2.min() >= 6 and l2.max() <= 30
    P = syn.padded(c2, o2, 30)
    assert np.array_equal(P[3, :l2[3]], c2[int(o2[3]):int(o2[4])]) and not P[3, l2[3]:].any()


def test_prefilter_pair_sampling_and_calibration():
    from seqtree_amd import prefilter as pf

    fam = np.repeat(np.arange(100), 4)
    a, b = pf.sample_pairs(np.random.default_rng(3), fam, 10_000, 0.5)
    near = slice(0, 5000)
    assert (fam[a[near]] == fam[b[near]]).all() and (a[near] != b[near]).all()
    assert ((a >= 0) & (a < 400) & (b >= 0) & (b < 400)).all()
    # the screening threshold keeps the requested share of close pairs, never below cutoff
    exact = np.array([10, 20, 30, 40, 200, 300])
    model = np.array([12.0, 35.0, 29.0, 90.0, 80.0, 400.0])
    assert pf.calibrate_tau(model, exact, 45, quantile=1.0) == 90.0
    assert pf.calibrate_tau(model, exact, 45, quantile=0.0) == 45.0
    assert pf.calibrate_tau(model, exact, 5) == 5.0
    w = pf.pair_weights(np.array([0.0, 50.0, 100.0]))
    assert w[0] > w[1] > w[2] > 0
    assert pf.recall(np.array([1, 2]), np.array([3, 4]), np.array([1, 2, 5]),
                     np.array([3, 9, 9])) == 1 / 3


def test_array_prior_cubes_equal_the_per_cell_evaluation(st):
    """The built-in priors' array form (used to build the cube fast) against calling the prior
    cell by cell, for several parameters and widths."""
    from seqtree_amd.gapblock import _prior_cube

    priors = [st.central_prior(0), st.central_prior(7), st.central_prior(21),
              st.frame_prior(3, 4), st.frame_prior(10, 0), st.frame_prior(1, 40),
              st.positions_prior((3, 4, -4, -3)), st.positions_prior((9, -9)),
              st.positions_prior((0,)), st.positions_prior((-1, 100))]
    for prior in priors:
        assert hasattr(prior, "cells")
        for width in (0, 1, 5, 20, 33):
            fast = _prior_cube(prior, width)
            slow = _prior_cube(lambda i, d, m, p=prior: p(i, d, m), width)
            assert fast == slow, (prior, width)
            assert all(type(v) is int for v in fast[:50])
    # parameters the array form does not take fall back to the per-cell path
    assert not hasattr(st.central_prior(1.5), "cells")
    assert _prior_cube(st.central_prior(1.5), 6) == _prior_cube(
        lambda i, d, m: st.central_prior(1.5)(i, d, m), 6)
