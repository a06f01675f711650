"""GPU parity for the Hamming and Levenshtein matrices (reference: src/distance.cpp and
tests/python/test_distance.py). The checker is the CPU oracle."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import AA, rand_seq

pytestmark = pytest.mark.gpu


def _raw(rng, lo, hi):
    """Arbitrary bytes, NUL and high bytes included (passed to the API as bytes objects)."""
    return bytes(rng.randrange(256) for _ in range(rng.randint(lo, hi)))


def _check(st, oracle_lib, which, a, b):
    fn = st.distance.hamming_matrix if which == "hamming" else st.distance.levenshtein_matrix
    got = np.asarray(fn(a, b))
    code, want = oracle_lib.distance_matrix(which, a, b, threads=8)
    assert code == 0
    assert got.shape == want.shape and got.dtype == np.int32
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} cells differ, first {bad[:3].tolist()}"
    return got


@pytest.mark.parametrize("L", [0, 1, 3, 4, 12, 13, 20, 64, 65, 100])
def test_hamming_lengths(gpu, oracle_lib, L):
    st = gpu
    rng = random.Random(L)
    a = [rand_seq(rng, L, L, AA + "acgt") for _ in range(70)]
    b = [rand_seq(rng, L, L, AA + "acgt") for _ in range(1300)]
    _check(st, oracle_lib, "hamming", a, b)


@pytest.mark.parametrize("nb", [1, 3, 4, 1301, 1302, 1303, 1304])
def test_hamming_ref_counts(gpu, oracle_lib, nb):
    """Reference counts on and off the four-column store width, short and long rows."""
    st = gpu
    rng = random.Random(nb)
    for L in (12, 40):
        a = [rand_seq(rng, L, L, "ACGT") for _ in range(37)]
        b = [rand_seq(rng, L, L, "ACGT") for _ in range(nb)]
        _check(st, oracle_lib, "hamming", a, b)


def test_hamming_raw_bytes_and_case(gpu, oracle_lib):
    st = gpu
    rng = random.Random(4)
    a = [_raw(rng, 9, 9) for _ in range(40)]
    b = [_raw(rng, 9, 9) for _ in range(300)] + a[:5]
    _check(st, oracle_lib, "hamming", a, b)
    assert st.distance.hamming("ABC", "abc") == 3
    assert st.distance.hamming("AAAA", "ATAT") == 2
    assert st.distance.hamming("", "") == 0


def test_hamming_mismatch_message(gpu):
    st = gpu
    with pytest.raises(ValueError, match=r"equal-length sequences \(got 3 and 4\)"):
        st.distance.hamming_matrix(["AAAA", "AAA"], ["AAAA"])
    with pytest.raises(ValueError, match=r"\(got 4 and 5\)"):
        st.distance.hamming_matrix(["AAAA", "AAA"], ["AAAA", "AAAAA"])
    assert np.asarray(st.distance.hamming_matrix([], ["A"])).shape == (0, 1)


def test_umi_workload_slice(gpu, oracle_lib):
    """A slice of the 50,000 x 50,000 length-12 ACGT UMI workload."""
    st = gpu
    rng = random.Random(12)
    a = [rand_seq(rng, 12, 12, "ACGT") for _ in range(300)]
    b = [rand_seq(rng, 12, 12, "ACGT") for _ in range(50000)]
    d = _check(st, oracle_lib, "hamming", a, b)
    assert d.max() <= 12


@pytest.mark.parametrize("lo,hi", [(0, 8), (0, 64), (50, 70), (60, 140), (120, 200),
                                    (180, 260), (250, 320)])
def test_levenshtein_lengths(gpu, oracle_lib, lo, hi):
    """Myers with 1..4 pattern words and the two-row fallback beyond 256."""
    st = gpu
    rng = random.Random(lo * 1000 + hi)
    a = [rand_seq(rng, lo, hi, "ACGT") for _ in range(12)] + [""]
    b = [rand_seq(rng, 0, hi, "ACGT") for _ in range(1100)] + a[:3] + [""]
    _check(st, oracle_lib, "levenshtein", a, b)


def test_levenshtein_raw_bytes(gpu, oracle_lib):
    st = gpu
    rng = random.Random(8)
    a = [_raw(rng, 0, 70) for _ in range(20)]
    b = [_raw(rng, 0, 70) for _ in range(600)]
    _check(st, oracle_lib, "levenshtein", a, b)


def test_levenshtein_known(gpu):
    st = gpu
    lev = st.distance.levenshtein
    for a, b, d in [("kitten", "sitting", 3), ("flaw", "lawn", 2),
                    ("CASSLGQAYEQYF", "CASSPGQAYEQF", 2), ("", "", 0), ("abc", "", 3),
                    ("", "abc", 3), ("abc", "abc", 0), ("ABC", "abc", 3)]:
        assert lev(a, b) == d


def test_reads_workload_slice(gpu, oracle_lib):
    """A slice of the 20,000 x 20,000 ACGT reads workload: 40-60 bases, half derived."""
    st = gpu
    rng = random.Random(60)
    base = [rand_seq(rng, 40, 60, "ACGT") for _ in range(1000)]
    derived = []
    for s in base:
        t = list(s)
        for _ in range(rng.randint(1, 3)):
            op, p = rng.randrange(3), rng.randrange(len(t))
            if op == 0:
                t[p] = rng.choice("ACGT")
            elif op == 1:
                t.insert(p, rng.choice("ACGT"))
            elif len(t) > 1:
                del t[p]
        derived.append("".join(t))
    reads = base + derived
    rng.shuffle(reads)
    d = _check(st, oracle_lib, "levenshtein", reads[:150], reads)
    assert (d >= 0).all()


def test_symmetry_and_zero_diagonal(gpu):
    st = gpu
    rng = random.Random(7)
    seqs = [rand_seq(rng, 10, 10) for _ in range(20)]
    h = np.asarray(st.distance.hamming_matrix(seqs, seqs))
    assert (h.diagonal() == 0).all() and (h == h.T).all()
    lev = [rand_seq(rng, 0, 14) for _ in range(20)]
    lv = np.asarray(st.distance.levenshtein_matrix(lev, lev))
    assert (lv.diagonal() == 0).all() and (lv == lv.T).all()
