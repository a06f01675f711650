"""GPU parity for Gotoh alignment with traceback (reference: src/pairwise.cpp:141-267,
`align_pair`; python/seqtree/pairwise.py:93-116, `align`). Every score, op string and aligned
string must equal the compiled reference's align_pair (or the oracle's restatement of it,
which tests/test_oracle.py pins to the reference), for the int32 kernels with LDS rows, the
int32 kernels with rows in scratch (references past 64 residues) and the int64 kernels
(absurd gap costs)."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import cdr3_like, launch_labels, rand_seq, seed_of

pytestmark = pytest.mark.gpu

GAPS = [(11, 1), (5, 5), (1, 3), (0, 0), (1, 10), (2, 100)]


def _strings(q, r, ops, qb, rb):
    aq, ar, i, j = [], [], qb, rb
    for o in ops:
        if o in "MS":
            aq.append(q[i]); ar.append(r[j]); i += 1; j += 1
        elif o == "D":
            aq.append(q[i]); ar.append("-"); i += 1
        else:
            aq.append("-"); ar.append(r[j]); j += 1
    return "".join(aq), "".join(ar)


def _sim(st, m):
    syms = st.amino_acids()
    return np.array([[m.similarity(a, b) for b in syms] for a in syms], np.int32)


@pytest.mark.parametrize("mode", ["global", "local"])
@pytest.mark.parametrize("go,ge", GAPS)
def test_align_matches_the_reference(gpu, reference, mode, go, ge):
    """One pair at a time through the reference's API, including tie-heavy three-letter inputs
    where the back-pointer order decides which optimal path comes back."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    rng = random.Random(seed_of(mode, go, ge))
    cases = [(rand_seq(rng, 1, 24), rand_seq(rng, 1, 24)) for _ in range(25)]
    cases += [(rand_seq(rng, 1, 12, "AAC"), rand_seq(rng, 1, 12, "AAC")) for _ in range(25)]
    cases += [("WWW", "KKK"), ("CASSLGQAYEQYF", "CASSPGQAYEQF")]
    for q, r in cases:
        a = st.pairwise.align(q, r, m, mode=mode, gap_open=go, gap_extend=ge)
        want = reference.align_pair(q, r, "blosum62", mode == "local", go, ge)
        assert (a.score, a.aligned_query, a.aligned_ref, a.ops) == want, (q, r)


def test_align_pairs_batch_equals_the_oracle(gpu, oracle_lib):
    """Thousands of pairs in one batch (CDR3-like and random, ragged lengths, both modes),
    each equal to the oracle's traceback."""
    st = gpu
    m = st.SubstitutionMatrix.pam250()
    sim = _sim(st, m)
    rng = random.Random(41)
    q = cdr3_like(rng, 120, 6, 30) + [rand_seq(rng, 1, 40) for _ in range(40)]
    r = cdr3_like(rng, 300, 6, 30) + [rand_seq(rng, 1, 50) for _ in range(60)]
    npr = np.random.default_rng(41)
    pq = npr.integers(0, len(q), 2500).astype(np.uint32)
    pr = npr.integers(0, len(r), 2500).astype(np.uint32)
    for mode, go, ge in (("global", 11, 1), ("local", 11, 1), ("global", 1, 10)):
        got = st.pairwise.align_pairs(q, r, pq, pr, m, mode=mode, gap_open=go, gap_extend=ge)
        assert len(got) == len(pq)
        for t in range(0, len(pq), 7):
            a, b = q[pq[t]], r[pr[t]]
            s, ops, qb, rb = oracle_lib.align_pair(a, b, sim, mode == "local", go, ge)
            aq, ar = _strings(a, b, ops, qb, rb)
            g = got[t]
            assert (g.score, g.ops, g.aligned_query, g.aligned_ref) == (s, ops, aq, ar), t


@pytest.mark.parametrize("go,ge", [(11, 1), (1, 3)])
def test_reference_residues_at_every_alignment(gpu, oracle_lib, go, ge):
    """The register-row kernels read each reference's residues as aligned dwords: references of
    1 to 12 residues start at every byte offset modulo 4, the set ends with references of 1 to
    3 residues (the last dword holds the set's last byte), and every (query, reference) pair
    must equal the oracle (four-bit kernel for gap_open >= gap_extend, byte kernel otherwise)."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    sim = _sim(st, m)
    rng = random.Random(seed_of("align", go, ge))
    r = [rand_seq(rng, n, n) for n in list(range(1, 13)) * 3] + ["W", "CA", "KKQ"]
    q = [rand_seq(rng, n, n) for n in (1, 2, 5, 9, 16, 25)]
    pq = np.repeat(np.arange(len(q)), len(r)).astype(np.uint32)
    pr = np.tile(np.arange(len(r)), len(q)).astype(np.uint32)
    for mode in ("global", "local"):
        got, labels = launch_labels(lambda: st.pairwise.align_pairs(
            q, r, pq, pr, m, mode=mode, gap_open=go, gap_extend=ge))
        assert f"traceback_{'nib' if go >= ge else 'reg'}_{mode}" in labels, labels
        for t in range(len(pq)):
            a, b = q[pq[t]], r[pr[t]]
            s, ops, qb, rb = oracle_lib.align_pair(a, b, sim, mode == "local", go, ge)
            aq, ar = _strings(a, b, ops, qb, rb)
            g = got[t]
            assert (g.score, g.ops, g.aligned_query, g.aligned_ref) == (s, ops, aq, ar), (mode, a, b)


def test_large_lists_run_in_shape_order(gpu, oracle_lib):
    """From 4,096 pairs on, the pairs are processed in shape order (sorted by clamped
    lengths); every output must still land at its own pair's index. Lengths past the clamp
    (31) and repeated pairs included."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    sim = _sim(st, m)
    rng = random.Random(47)
    q = cdr3_like(rng, 300, 6, 30) + [rand_seq(rng, 60, 70) for _ in range(10)]
    r = cdr3_like(rng, 400, 6, 30) + [rand_seq(rng, 60, 66) for _ in range(10)] + [""]
    npr = np.random.default_rng(47)
    n = 6000
    pq = npr.integers(0, len(q), n).astype(np.uint32)
    pr = npr.integers(0, len(r), n).astype(np.uint32)
    pq[:50], pr[:50] = 3, 5
    for mode in ("global", "local"):
        got = st.pairwise.align_pairs(q, r, pq, pr, m, mode=mode)
        assert len(got) == n
        for t in range(n):
            a, b = q[pq[t]], r[pr[t]]
            s_, ops, qb, rb = oracle_lib.align_pair(a, b, sim, mode == "local", 11, 1)
            g = got[t]
            assert (g.score, g.ops) == (s_, ops), t
            assert (g.aligned_query, g.aligned_ref) == _strings(a, b, ops, qb, rb), t


@pytest.mark.parametrize("go,ge", [(11, 1), (5, 5), (2, 1), (1, 3), (0, 0)])
def test_register_rows_kernel(gpu, oracle_lib, go, ge):
    """References of at most 32 residues and queries of at most 64 take the register-row
    kernels: every reference length 1..32 (the 4- and 8-column groups and the last partial
    ones), queries up to 64, lists below and above the shape-order threshold. gap_open >=
    gap_extend takes the four-bit kernel (go = ge included), gap_open < gap_extend the
    direction-byte one."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    sim = _sim(st, m)
    rng = random.Random(seed_of("reg", go, ge))
    q = [rand_seq(rng, 1, 64) for _ in range(60)] + ["W" * 64, "A"]
    r = [rand_seq(rng, L, L) for L in range(1, 33)] + [rand_seq(rng, 1, 12, "AAC")
                                                     for _ in range(20)] + [""]
    npr = np.random.default_rng(seed_of("reg", go, ge))
    for n in (700, 5000):
        pq = npr.integers(0, len(q), n).astype(np.uint32)
        pr = npr.integers(0, len(r), n).astype(np.uint32)
        for mode in ("global", "local"):
            got, labels = launch_labels(lambda: st.pairwise.align_pairs(
                q, r, pq, pr, m, mode=mode, gap_open=go, gap_extend=ge))
            kernel = "nib" if go >= ge else "reg"
            assert f"traceback_{kernel}_{mode}" in labels, labels
            for t in range(0, n, 1 if n < 1000 else 3):
                a, b = q[pq[t]], r[pr[t]]
                s_, ops, qb, rb = oracle_lib.align_pair(a, b, sim, mode == "local", go, ge)
                g = got[t]
                assert (g.score, g.ops) == (s_, ops), (t, a, b)
                assert (g.aligned_query, g.aligned_ref) == _strings(a, b, ops, qb, rb), t


def test_long_sequences_take_the_scratch_rows(gpu, oracle_lib):
    """References past 64 residues: the int32 rows live in per-workgroup scratch."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    sim = _sim(st, m)
    rng = random.Random(43)
    q = [rand_seq(rng, 60, 140) for _ in range(12)]
    r = [rand_seq(rng, 65, 160) for _ in range(12)]
    pq = np.arange(12, dtype=np.uint32)
    for mode in ("global", "local"):
        got = st.pairwise.align_pairs(q, r, pq, pq, m, mode=mode)
        for t in range(12):
            s, ops, qb, rb = oracle_lib.align_pair(q[t], r[t], sim, mode == "local", 11, 1)
            assert (got[t].score, got[t].ops) == (s, ops)
            assert (got[t].aligned_query, got[t].aligned_ref) == _strings(q[t], r[t], ops, qb, rb)


def test_absurd_gap_costs_run_in_int64_and_overflow_is_refused(gpu, reference):
    """The reference's int64 known answers (tests/python/test_pairwise.py:215-228) through the
    traceback path, and its OverflowError where the score does not fit int32."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    q, r = "CASSLGQAYEQYF", "CASSPGQAYEQF"
    for go, want in ((10**8, -99_999_944), (2**30, -1_073_741_768), (2**31 - 1, -2_147_483_591)):
        a = st.pairwise.align(q, r, m, gap_open=go)
        assert a.score == want
        assert (a.score, a.aligned_query, a.aligned_ref, a.ops) == \
            reference.align_pair(q, r, "blosum62", False, go, 1)
    with pytest.raises(OverflowError, match="32 bits"):
        st.pairwise.align("A" * 400, "W" * 600, m, gap_open=2**31 - 1, gap_extend=2**31 - 1)


def test_empty_sides_keep_the_callers_characters(gpu, reference):
    """With an empty side the reference copies the caller's own characters (lower case
    included) instead of decoded symbols; a local alignment of nothing is empty."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    for q, r in (("", "acd"), ("wyk", ""), ("", "")):
        for mode in ("global", "local"):
            a = st.pairwise.align(q, r, m, mode=mode)
            assert (a.score, a.aligned_query, a.aligned_ref, a.ops) == \
                reference.align_pair(q, r, "blosum62", mode == "local", 11, 1)
    a = st.pairwise.align("cassl", "CASSL", m)
    assert (a.aligned_query, a.aligned_ref, a.ops) == ("CASSL", "CASSL", "MMMMM")
    a = st.pairwise.align("WWW", "KKK", m, mode="local")
    assert (a.score, a.aligned_query, a.ops) == (0, "", "")


def test_errors(gpu):
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    with pytest.raises(IndexError):
        st.pairwise.align_pairs(["CAS"], ["CAS"], [0], [1], m)
    with pytest.raises(ValueError, match="invalid symbol"):
        st.pairwise.align("CA1", "CAS", m)
    with pytest.raises(ValueError, match="gap_open"):
        st.pairwise.align("CAS", "CAS", m, gap_open=-1)
    assert repr(st.pairwise.align("CAS", "CAS", m)).startswith("Alignment(score=")
