"""GPU parity for the penalty-space global alignment of a query to a stored reference
(reference: src/index.cpp:105-179, `Index::align`; bound at src/_bindings.cpp:197-201, 454).
Every score, op string and aligned string must equal the compiled reference's Index::align or
the oracle's restatement of it (which tests/test_oracle.py pins to the reference and to
enumeration), across unit cost and matrix penalties, gap settings, tie-heavy inputs, the
register / LDS-row / scratch-row int32 kernels and the int64 kernels (gap costs whose sums
pass the int32 range bound), the shape-ordered path for long pair lists, and empty sides."""
from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

from conftest import cdr3_like, launch_labels, rand_seq, seed_of

pytestmark = pytest.mark.gpu

REFS = ["CASSLGQAYEQYF", "CASSPGTGVYGYTF", "CASRQGAWDTQYF", "CASSFRSSYNEQFF",
        "CAWSVSGGGTDTQYF", "CASSLAPGATNEKLFF", "CSARDRTGNGYTF", "CASSIRSSYEQYF"]


def _pen(st, m):
    syms = st.amino_acids()
    return np.array([[m.penalty(a, b) for b in syms] for a in syms], np.int32)


def _strings(q, r, ops):
    aq, ar, i, j = [], [], 0, 0
    for o in ops:
        aq.append(q[i] if o != "I" else "-")
        ar.append(r[j] if o != "D" else "-")
        i += o != "I"
        j += o != "D"
    return "".join(aq), "".join(ar)


def test_reference_known_answers(gpu):
    """The reference's own Index.align tests (tests/python/test_engine_guards.py:50-80,
    test_matrix.py:28-40, test_sanity.py:86-98), through the same Python API."""
    st = gpu
    idx = st.Index.build(list(REFS), "aa")
    cheap = idx.align(0, "CASSLGYEQYF", st.SearchParams(matrix="BLOSUM62", gap_open=11,
                                                        gap_extend=1))
    dear = idx.align(0, "CASSLGYEQYF", st.SearchParams(matrix="BLOSUM62", gap_open=11,
                                                       gap_extend=9))
    assert cheap.score < dear.score and dear.score - cheap.score == 8
    a = idx.align(0, "CASSLGYEQYF", st.SearchParams(gap_open=1, gap_extend=1))
    assert a.score == 2 and a.ops.count("I") + a.ops.count("D") == 2
    m = st.SubstitutionMatrix.blosum62()
    degenerate = idx.align(0, "CASSIRSSYEQYF", st.SearchParams(matrix="BLOSUM62", gap_open=1))
    sane = idx.align(0, "CASSIRSSYEQYF", st.SearchParams(matrix="BLOSUM62", gap_open=m.scale()))
    assert degenerate.ops.count("I") > 0
    assert sane.ops == "M" * 4 + "S" * 4 + "M" * 5
    two = st.Index.build(["CASSLAPGATNEKLFF", "CASSLELGATNEKLFF"], alphabet="aa")
    for name, expected in {"identity": 2, "structural": 18, "PAM250": 24, "BLOSUM62": 28,
                           "PAM100": 30}.items():
        p = st.SearchParams(matrix=name, gap_open=100, engine="seqtm")
        assert two.align(1, "CASSLAPGATNEKLFF", p).score == expected, name
    one = st.Index.build(["CASSPGQAYEQF"], "aa")
    for p in (st.SearchParams(matrix=m, max_penalty=100, engine="seqtm"),
              st.SearchParams(max_penalty=100, engine="seqtm")):
        aln = one.align(0, "casspgqayeqf", p)
        assert aln.score == 0 and aln.ops == "M" * 12
        assert aln.aligned_query == "casspgqayeqf" and aln.aligned_ref == "CASSPGQAYEQF"
    p = st.SearchParams(matrix=m, max_penalty=200, gap_open=2 * m.scale(), gap_extend=1,
                        engine="seqtm")
    assert one.align(0, "CASSPGQAYEQY", p).ops == "M" * 11 + "S"


@pytest.mark.parametrize("matrix,go,ge", [("", 1, 1), ("BLOSUM62", 11, 1), ("BLOSUM62", 28, 1),
                                          ("structural", 7, 3), ("PAM250", 0, 0), ("", 2, 5),
                                          ("BLOSUM45", 40, 1), ("identity", 1, 4)])
def test_align_matches_the_reference(gpu, reference, matrix, go, ge):
    """Random and tie-heavy pairs (three-letter strings where the strict-< state order picks
    among equal paths), one call each, against the compiled reference."""
    st = gpu
    rng = random.Random(seed_of("gpu-index", matrix, go, ge))
    refs = [rand_seq(rng, 1, 24) for _ in range(20)] + [rand_seq(rng, 1, 12, "AAC")
                                                        for _ in range(20)] + ["W", "KKK"]
    idx = st.Index.build(refs, "aa")
    p = st.SearchParams(matrix=matrix, gap_open=go, gap_extend=ge)
    ref_kind = None if matrix in ("", "identity") else matrix.lower()
    qs = [rand_seq(rng, 1, 24) for _ in range(20)] + [rand_seq(rng, 1, 12, "AAC")
                                                      for _ in range(20)] + ["WWW", "K"]
    pr = [rng.randrange(len(refs)) for _ in qs]
    want = reference.index_align_pairs(qs, refs, list(range(len(qs))), pr, ref_kind, go, ge)
    for q, b, w in zip(qs, pr, want):
        a = idx.align(b, q, p)
        assert (a.score, a.aligned_query, a.aligned_ref, a.ops) == w, (q, refs[b])


def test_batch_kernels_equal_the_oracle(gpu, oracle_lib):
    """One batch per kernel: CDR3-like pairs (register kernels: four bits per cell when
    gap_open >= gap_extend, direction bytes otherwise), references of 33-64 residues
    (LDS rows), longer ones (rows in scratch), absurd gap costs (int64 cells) and a list of
    more than 4,096 pairs (shape order); every output at its own index."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    pen = _pen(st, m)
    rng = random.Random(77)
    cases = [
        ("traceback_nib_penalty", cdr3_like(rng, 200, 6, 30), cdr3_like(rng, 300, 6, 30),
         11, 1, 3000),
        ("traceback_reg_penalty", cdr3_like(rng, 200, 6, 30), cdr3_like(rng, 300, 6, 30),
         2, 5, 3000),
        ("traceback_i32_penalty", cdr3_like(rng, 60, 20, 70), [rand_seq(rng, 33, 64)
                                                               for _ in range(50)], 11, 1, 700),
        ("traceback_i32_penalty", [rand_seq(rng, 1, 90) for _ in range(40)],
         [rand_seq(rng, 65, 120) for _ in range(30)], 5, 2, 300),
        ("traceback_i64_penalty", cdr3_like(rng, 40, 6, 30), cdr3_like(rng, 40, 6, 30),
         20_000_000, 3_000_000, 400),
        ("traceback_nib_penalty", cdr3_like(rng, 500, 6, 30), cdr3_like(rng, 800, 6, 30),
         14, 1, 6000),
    ]
    for label, qs, refs, go, ge, n in cases:
        idx = st.Index.build(refs, "aa")
        pr = [rng.randrange(len(refs)) for _ in range(n)]
        pq = [rng.randrange(len(qs)) for _ in range(n)]
        batch = [qs[a] for a in pq]
        p = st.SearchParams(matrix=m, gap_open=go, gap_extend=ge)
        got, labels = launch_labels(lambda: idx.align_batch(pr, batch, p))
        assert label in labels, (label, labels)
        assert len(got) == n
        for t in range(0, n, 3):
            s, ops = oracle_lib.index_align(batch[t], refs[pr[t]], pen, go, ge)
            g = got[t]
            assert (g.score, g.ops) == (s, ops), (label, t)
            assert (g.aligned_query, g.aligned_ref) == _strings(batch[t], refs[pr[t]], ops)


def test_unit_cost_nucleotides_and_empty_sides(gpu, oracle_lib):
    """Unit cost (no matrix) over the nucleotide codec, empty queries and references (all
    gaps, or the empty alignment), and case folding through the codec."""
    st = gpu
    rng = random.Random(5)
    refs = ["", "ACGT", "acgtac"] + [rand_seq(rng, 0, 20, "ACGT") for _ in range(30)]
    qs = ["", "A", "ACGTACGT", "tttt"] + [rand_seq(rng, 0, 20, "ACGT") for _ in range(30)]
    idx = st.Index.build(refs, "nt")
    pairs = [(i, j) for i in range(len(qs)) for j in range(len(refs))]
    p = st.SearchParams(gap_open=3, gap_extend=1)
    got = idx.align_batch([j for _, j in pairs], [qs[i] for i, _ in pairs], p)
    for (i, j), g in zip(pairs, got):
        s, ops = oracle_lib.index_align(qs[i], refs[j], None, 3, 1, alphabet="nt")
        assert (g.score, g.ops) == (s, ops), (qs[i], refs[j])
        assert (g.aligned_query, g.aligned_ref) == _strings(qs[i], refs[j], ops)


def test_errors_in_the_reference_order(gpu):
    st = gpu
    idx = st.Index.build(list(REFS), "aa")
    with pytest.raises(ValueError, match="invalid symbol"):
        idx.align(0, "CASS!GQAYEQYF", st.SearchParams(gap_open=-1))  # symbols first
    with pytest.raises(ValueError, match=">= 0"):
        idx.align(0, "CASSLGYEQYF", st.SearchParams(gap_open=-1))
    with pytest.raises(IndexError):
        idx.align(len(REFS), "CASS", st.SearchParams())
    with pytest.raises(ValueError, match="matrix size does not match"):
        idx.align(0, "CASS", st.SearchParams(matrix=st.SubstitutionMatrix.unit(4)))


def _device_csr(seqs, codec):
    from seqtree_amd._device import DeviceArray

    codes = np.array([codec[c] for c in "".join(seqs)] or [0], np.uint8)
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return DeviceArray.from_numpy(codes), DeviceArray.from_numpy(off)


def test_device_entry_points_and_the_length_bound(gpu, oracle_lib):
    """st_index_align_pairs_device equals the oracle; st_align_pairs_device and
    st_index_align_pairs_device refuse a pair longer than the stated bounds with
    ST_INVALID_ARGUMENT instead of writing past their scratch."""
    st = gpu
    from seqtree_amd import _device
    from seqtree_amd._device import DeviceArray, Status

    L = _device.lib()
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
    L.st_index_align_pairs_device.restype = i32
    L.st_index_align_pairs_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, vp, u64,
                                              vp, u32, i32, i32, vp, u64, vp, vp, vp,
                                              ctypes.POINTER(Status)]
    L.st_align_pairs_device.restype = i32
    L.st_align_pairs_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, vp, u64, vp,
                                        u32, i32, i32, i32, vp, u64, vp, vp, vp, vp, vp,
                                        ctypes.POINTER(Status)]
    syms = st.amino_acids()
    codec = {c: i for i, c in enumerate(syms)}
    m = st.SubstitutionMatrix.blosum62()
    pen = _pen(st, m)
    rng = random.Random(9)
    qs, rs = cdr3_like(rng, 50, 6, 30), cdr3_like(rng, 70, 6, 30)
    n = 500
    pq = np.array([rng.randrange(len(qs)) for _ in range(n)], np.uint32)
    pr = np.array([rng.randrange(len(rs)) for _ in range(n)], np.uint32)
    qc, qo = _device_csr(qs, codec)
    rc, ro = _device_csr(rs, codec)
    dpq, dpr = DeviceArray.from_numpy(pq), DeviceArray.from_numpy(pr)
    qmax, rmax = max(map(len, qs)), max(map(len, rs))
    stride = qmax + rmax
    ops = DeviceArray((n * stride,), np.uint8)
    score = DeviceArray((n,), np.int32)
    n_ops = DeviceArray((n,), np.uint32)
    stt = Status()
    rc_ = L.st_index_align_pairs_device(qc.ptr, qo.ptr, len(qs), qmax, rc.ptr, ro.ptr, len(rs),
                                        rmax, dpq.ptr, dpr.ptr, n, pen.ctypes.data, len(syms),
                                        11, 1, ops.ptr, stride, score.ptr, n_ops.ptr, None,
                                        ctypes.byref(stt))
    assert rc_ == 0, stt.message
    _device.synchronize()
    o, s, c = ops.numpy().reshape(n, stride), score.numpy(), n_ops.numpy()
    for t in range(n):
        want = oracle_lib.index_align(qs[pq[t]], rs[pr[t]], pen, 11, 1)
        assert (int(s[t]), bytes(o[t, :c[t]]).decode()) == want, t
    # a bound one below the true maximum: refused, nothing written past the scratch
    stt = Status()
    rc_ = L.st_index_align_pairs_device(qc.ptr, qo.ptr, len(qs), qmax, rc.ptr, ro.ptr, len(rs),
                                        rmax - 1, dpq.ptr, dpr.ptr, n, pen.ctypes.data,
                                        len(syms), 11, 1, ops.ptr, stride, score.ptr,
                                        n_ops.ptr, None, ctypes.byref(stt))
    assert rc_ == 1 and b"longer than" in stt.message
    sim = np.array([[m.similarity(a, b) for b in syms] for a in syms], np.int32)
    s64 = DeviceArray((n,), np.int64)
    meta = DeviceArray((3 * n,), np.uint32)
    stt = Status()
    rc_ = L.st_align_pairs_device(qc.ptr, qo.ptr, len(qs), qmax - 1, rc.ptr, ro.ptr, len(rs),
                                  rmax, dpq.ptr, dpr.ptr, n, sim.ctypes.data, len(syms), 0, 11,
                                  1, ops.ptr, stride, s64.ptr, meta.ptr, meta.ptr + 4 * n,
                                  meta.ptr + 8 * n, None, ctypes.byref(stt))
    assert rc_ == 1 and b"longer than" in stt.message
    # and the correct bounds still work after a refused call
    stt = Status()
    rc_ = L.st_align_pairs_device(qc.ptr, qo.ptr, len(qs), qmax, rc.ptr, ro.ptr, len(rs),
                                  rmax, dpq.ptr, dpr.ptr, n, sim.ctypes.data, len(syms), 0, 11,
                                  1, ops.ptr, stride, s64.ptr, meta.ptr, meta.ptr + 4 * n,
                                  meta.ptr + 8 * n, None, ctypes.byref(stt))
    assert rc_ == 0, stt.message
    got = s64.numpy()
    for t in range(0, n, 5):
        assert int(got[t]) == oracle_lib.align_pair(qs[pq[t]], rs[pr[t]], sim, False, 11, 1)[0]
