"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle and the compiled reference.

* :class:`Oracle` wraps ``oracle/liboracle.so``, the plain-C restatement of the reference's
  all-vs-all scoring algorithms (``oracle/seqtree_oracle.c``).
* :class:`Reference` wraps ``oracle/_ref/libseqtree_ref.so``, the reference's own C++ sources
  (``src/{gapblock,pairwise,distance,substitution_matrix,codec}.cpp``, plus ``index.cpp`` and
  the trie and engine sources it links against, for ``Index::align``) compiled by
  ``oracle/Makefile`` together with ``oracle/ref_shim.cpp``. It exists only where the
  reference tree was available at build time.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use this
package, and only as the checker / CPU baseline; the shipped library never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libseqtree_ref.so"
REFERENCE_TREE = Path("/root/reference")

_c_u8p = ctypes.POINTER(ctypes.c_uint8)
_c_u64p = ctypes.POINTER(ctypes.c_uint64)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_u32p = ctypes.POINTER(ctypes.c_uint32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)

ALPHABETS = {"aa": 0, "nt": 1, "iupac": 2}


def build(reference: bool = True) -> None:
    """Compile liboracle.so and, when the reference tree is present, the reference library."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if reference and REFERENCE_TREE.is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _bytes(s) -> bytes:
    """str -> UTF-8 bytes, the conversion pybind11 applies to a std::string argument."""
    return s.encode("utf-8") if isinstance(s, str) else bytes(s)


def pack(seqs, encode=None):
    """CSR form of a list of str/bytes: (uint8 bytes, uint64 offsets[n+1])."""
    raw = [_bytes(s) for s in seqs]
    off = np.zeros(len(raw) + 1, dtype=np.uint64)
    if raw:
        off[1:] = np.cumsum([len(b) for b in raw], dtype=np.uint64)
    data = np.frombuffer(b"".join(raw), dtype=np.uint8).copy() if raw else np.zeros(0, np.uint8)
    if encode is not None:
        data = encode(data)
    if data.size == 0:
        data = np.zeros(1, np.uint8)  # keep a valid pointer for empty inputs
    return data, off


def _p(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """The plain-C restatement (see ``seqtree_oracle.c``)."""

    def __init__(self, path: Path = LIB):
        if not Path(path).exists():
            build(reference=False)
        lib = ctypes.CDLL(str(path))
        lib.or_encode.restype = ctypes.c_int64
        lib.or_encode.argtypes = [ctypes.c_int32, _c_u8p, ctypes.c_uint64, _c_u8p]
        lib.or_pen_from_sim.argtypes = [ctypes.c_int32, _c_i32p, _c_i32p]
        lib.or_scale.restype = ctypes.c_int32
        lib.or_scale.argtypes = [ctypes.c_int32, _c_i32p]
        lib.or_gotoh.restype = ctypes.c_int64
        lib.or_gotoh.argtypes = [_c_u8p, ctypes.c_uint64, _c_u8p, ctypes.c_uint64, _c_i32p,
                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        lib.or_align_pair.restype = None
        lib.or_align_pair.argtypes = [
            _c_u8p, ctypes.c_uint64, _c_u8p, ctypes.c_uint64, _c_i32p, ctypes.c_int32,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
            _c_u8p, _c_u64p, _c_u64p, _c_u64p]
        lib.or_index_align.restype = None
        lib.or_index_align.argtypes = [
            _c_u8p, ctypes.c_uint64, _c_u8p, ctypes.c_uint64, _c_i32p, ctypes.c_int32,
            ctypes.c_int32, ctypes.c_int32, _c_i32p, _c_u8p, _c_u64p]
        lib.or_hamming.restype = ctypes.c_int32
        lib.or_hamming.argtypes = [_c_u8p, ctypes.c_uint64, _c_u8p, ctypes.c_uint64]
        lib.or_levenshtein.restype = ctypes.c_int32
        lib.or_levenshtein.argtypes = [_c_u8p, ctypes.c_uint64, _c_u8p, ctypes.c_uint64]
        lib.or_gapblock_matrix.restype = ctypes.c_int
        lib.or_gapblock_matrix.argtypes = [
            _c_u8p, _c_u64p, ctypes.c_uint64, _c_u8p, _c_u64p, ctypes.c_uint64, _c_i32p,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _c_i32p, ctypes.c_uint32,
            ctypes.c_int, _c_i32p]
        lib.or_align_matrix.restype = ctypes.c_int
        lib.or_align_matrix.argtypes = [
            _c_u8p, _c_u64p, ctypes.c_uint64, _c_u8p, _c_u64p, ctypes.c_uint64, _c_i32p,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
            ctypes.c_int, _c_i32p]
        lib.or_distance_matrix.restype = ctypes.c_int
        lib.or_distance_matrix.argtypes = [
            ctypes.c_int, _c_u8p, _c_u64p, ctypes.c_uint64, _c_u8p, _c_u64p, ctypes.c_uint64,
            ctypes.c_int, _c_i32p]
        self.lib = lib

    # -- helpers ------------------------------------------------------------------------
    def encoder(self, alphabet: str):
        code = ALPHABETS[alphabet]

        def enc(data):
            out = np.empty_like(data)
            bad = self.lib.or_encode(code, _p(data, _c_u8p), data.size, _p(out, _c_u8p))
            if bad >= 0:
                raise ValueError(f"symbol {chr(data[bad])!r} is not in the alphabet")
            return out

        return enc

    def pen_from_sim(self, sim):
        sim = np.ascontiguousarray(sim, dtype=np.int32)
        A = sim.shape[0]
        pen = np.empty_like(sim)
        self.lib.or_pen_from_sim(A, _p(sim, _c_i32p), _p(pen, _c_i32p))
        return pen

    def scale(self, pen):
        pen = np.ascontiguousarray(pen, dtype=np.int32)
        return int(self.lib.or_scale(pen.shape[0], _p(pen, _c_i32p)))

    # -- scalar -------------------------------------------------------------------------
    def gotoh(self, q, r, sim, local=False, go=11, ge=1, alphabet="aa"):
        sim = np.ascontiguousarray(sim, dtype=np.int32)
        enc = self.encoder(alphabet)
        qc, qo = pack([q], enc)
        rc, ro = pack([r], enc)
        return int(self.lib.or_gotoh(_p(qc, _c_u8p), int(qo[1]), _p(rc, _c_u8p), int(ro[1]),
                                     _p(sim, _c_i32p), sim.shape[0], int(local), go, ge))

    def align_pair(self, q, r, sim, local=False, go=11, ge=1, alphabet="aa"):
        """(score as int64, ops, first aligned query residue, first aligned reference
        residue) of or_align_pair; q and r are strings (or bytes) in `alphabet`."""
        sim = np.ascontiguousarray(sim, dtype=np.int32)
        enc = self.encoder(alphabet)
        qc, qo = pack([q], enc)
        rc, ro = pack([r], enc)
        m, n = int(qo[1]), int(ro[1])
        ops = np.zeros(m + n + 1, np.uint8)
        score = ctypes.c_int64()
        meta = np.zeros(3, np.uint64)
        self.lib.or_align_pair(_p(qc, _c_u8p), m, _p(rc, _c_u8p), n, _p(sim, _c_i32p),
                               sim.shape[0], int(local), go, ge, ctypes.byref(score),
                               _p(ops, _c_u8p), _p(meta[0:], _c_u64p), _p(meta[1:], _c_u64p),
                               _p(meta[2:], _c_u64p))
        return score.value, bytes(ops[: int(meta[0])]).decode(), int(meta[1]), int(meta[2])

    def index_align(self, q, r, pen=None, go=1, ge=1, alphabet="aa"):
        """(penalty score, ops) of or_index_align, the reference's Index::align; pen is the
        penalty table (None for unit cost); q and r are strings (or bytes) in `alphabet`."""
        enc = self.encoder(alphabet)
        qc, qo = pack([q], enc)
        rc, ro = pack([r], enc)
        m, n = int(qo[1]), int(ro[1])
        tab = None if pen is None else np.ascontiguousarray(pen, dtype=np.int32)
        ops = np.zeros(m + n + 1, np.uint8)
        score = np.zeros(1, np.int32)
        cnt = np.zeros(1, np.uint64)
        self.lib.or_index_align(_p(qc, _c_u8p), m, _p(rc, _c_u8p), n,
                                None if tab is None else _p(tab, _c_i32p),
                                0 if tab is None else tab.shape[0], go, ge, _p(score, _c_i32p),
                                _p(ops, _c_u8p), _p(cnt, _c_u64p))
        return int(score[0]), bytes(ops[: int(cnt[0])]).decode()

    def hamming(self, a, b):
        x, xo = pack([a])
        y, yo = pack([b])
        return int(self.lib.or_hamming(_p(x, _c_u8p), int(xo[1]), _p(y, _c_u8p), int(yo[1])))

    def levenshtein(self, a, b):
        x, xo = pack([a])
        y, yo = pack([b])
        return int(self.lib.or_levenshtein(_p(x, _c_u8p), int(xo[1]), _p(y, _c_u8p),
                                           int(yo[1])))

    # -- matrices -----------------------------------------------------------------------
    def gapblock_matrix(self, q, r, pen, go, ge, prior=None, width=0, alphabet="aa",
                        threads=1):
        pen = np.ascontiguousarray(pen, dtype=np.int32)
        enc = self.encoder(alphabet)
        qc, qo = pack(q, enc)
        rc, ro = pack(r, enc)
        out = np.zeros((len(q), len(r)), dtype=np.int32)
        pr = None if prior is None or len(prior) == 0 else np.ascontiguousarray(prior, np.int32)
        rc_ = self.lib.or_gapblock_matrix(
            _p(qc, _c_u8p), _p(qo, _c_u64p), len(q), _p(rc, _c_u8p), _p(ro, _c_u64p), len(r),
            _p(pen, _c_i32p), pen.shape[0], go, ge, None if pr is None else _p(pr, _c_i32p),
            width, threads, _p(out, _c_i32p))
        assert rc_ == 0
        return out

    def align_matrix(self, q, r, sim, local=False, dist=False, go=11, ge=1, alphabet="aa",
                     threads=1):
        """Returns (code, matrix): code 0 ok, 2 overflow."""
        sim = np.ascontiguousarray(sim, dtype=np.int32)
        enc = self.encoder(alphabet)
        qc, qo = pack(q, enc)
        rc, ro = pack(r, enc)
        out = np.zeros((len(q), len(r)), dtype=np.int32)
        code = self.lib.or_align_matrix(
            _p(qc, _c_u8p), _p(qo, _c_u64p), len(q), _p(rc, _c_u8p), _p(ro, _c_u64p), len(r),
            _p(sim, _c_i32p), sim.shape[0], int(local), int(dist), go, ge, threads,
            _p(out, _c_i32p))
        return code, out

    def distance_matrix(self, which, a, b, threads=1):
        """which: 'hamming' or 'levenshtein'. Returns (code, matrix): code 1 = length mismatch."""
        x, xo = pack(a)
        y, yo = pack(b)
        out = np.zeros((len(a), len(b)), dtype=np.int32)
        code = self.lib.or_distance_matrix(
            0 if which == "hamming" else 1, _p(x, _c_u8p), _p(xo, _c_u64p), len(a),
            _p(y, _c_u8p), _p(yo, _c_u64p), len(b), threads, _p(out, _c_i32p))
        return code, out


MATRIX_KINDS = {"blosum62": 0, "blosum45": 1, "blosum80": 2, "pam250": 3, "pam100": 4,
                "structural": 5}


class ReferenceError_(Exception):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class Reference:
    """The reference's own C++ (compiled from /root/reference sources by oracle/Makefile)."""

    def __init__(self, path: Path = REF_LIB):
        if not Path(path).exists():
            raise FileNotFoundError(str(path))
        lib = ctypes.CDLL(str(path))
        lib.ref_matrix_tables.argtypes = [ctypes.c_int, ctypes.c_int32, _c_i32p, _c_i32p,
                                          _c_i32p, _c_i32p, _c_i32p]
        lib.ref_gapblock_matrix.argtypes = [
            ctypes.c_char_p, _c_u64p, ctypes.c_uint64, ctypes.c_char_p, _c_u64p, ctypes.c_uint64,
            ctypes.c_int, ctypes.c_int, _c_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
            _c_i32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, _c_i32p, ctypes.c_char_p,
            ctypes.c_int]
        lib.ref_align_matrix.argtypes = [
            ctypes.c_char_p, _c_u64p, ctypes.c_uint64, ctypes.c_char_p, _c_u64p, ctypes.c_uint64,
            ctypes.c_int, ctypes.c_int, _c_i32p, ctypes.c_int32, ctypes.c_int, ctypes.c_int,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int, _c_i32p, ctypes.c_char_p, ctypes.c_int]
        lib.ref_align_score.argtypes = [
            ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
            ctypes.c_int, _c_i32p, ctypes.c_int32, ctypes.c_int, ctypes.c_int32, ctypes.c_int32,
            _c_i32p, ctypes.c_char_p, ctypes.c_int]
        lib.ref_align_pair.argtypes = [
            ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
            ctypes.c_int, _c_i32p, ctypes.c_int32, ctypes.c_int, ctypes.c_int32, ctypes.c_int32,
            _c_i32p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _c_u64p,
            ctypes.c_char_p, ctypes.c_int]
        lib.ref_align_pairs.argtypes = [
            ctypes.c_char_p, _c_u64p, ctypes.c_uint64, ctypes.c_char_p, _c_u64p, ctypes.c_uint64,
            _c_u32p, _c_u32p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int, _c_u64p, ctypes.c_char_p, ctypes.c_int]
        lib.ref_index_align_pairs.argtypes = [
            ctypes.c_char_p, _c_u64p, ctypes.c_uint64, ctypes.c_char_p, _c_u64p, ctypes.c_uint64,
            _c_u32p, _c_u32p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _c_i32p,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int, _c_i32p,
            ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, _c_u64p,
            ctypes.c_char_p, ctypes.c_int]
        lib.ref_distance_matrix.argtypes = [
            ctypes.c_int, ctypes.c_char_p, _c_u64p, ctypes.c_uint64, ctypes.c_char_p, _c_u64p,
            ctypes.c_uint64, ctypes.c_int, _c_i32p, ctypes.c_char_p, ctypes.c_int]
        lib.ref_distance.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64,
                                     ctypes.c_char_p, ctypes.c_uint64, _c_i32p, ctypes.c_char_p,
                                     ctypes.c_int]
        self.lib = lib

    @staticmethod
    def _matrix(kind):
        """kind: None, a builtin name, ('unit', size) or a square similarity grid."""
        if kind is None:
            return -1, None, 0
        if isinstance(kind, str):
            return MATRIX_KINDS[kind], None, 0
        if isinstance(kind, tuple) and kind[0] == "unit":
            return 6, None, kind[1]
        grid = np.ascontiguousarray(kind, dtype=np.int32)
        return 7, grid, grid.shape[0]

    @staticmethod
    def _csr(seqs):
        raw = [_bytes(s) for s in seqs]
        off = np.zeros(len(raw) + 1, dtype=np.uint64)
        if raw:
            off[1:] = np.cumsum([len(b) for b in raw], dtype=np.uint64)
        return b"".join(raw) + b"\0", off

    def tables(self, kind):
        k, grid, size = self._matrix(kind)
        sim = np.zeros(32 * 32, np.int32)
        pen = np.zeros(32 * 32, np.int32)
        A = ctypes.c_int32()
        sc = ctypes.c_int32()
        rc = self.lib.ref_matrix_tables(k, size, None if grid is None else _p(grid, _c_i32p),
                                        _p(sim, _c_i32p), _p(pen, _c_i32p), ctypes.byref(A),
                                        ctypes.byref(sc))
        assert rc == 0
        n = A.value
        return sim[: n * n].reshape(n, n), pen[: n * n].reshape(n, n), sc.value

    def _err(self):
        return ctypes.create_string_buffer(512)

    def gapblock_matrix(self, q, r, matrix, go, ge, prior=(), width=0, alphabet="aa",
                        threads=1):
        k, grid, size = self._matrix(matrix)
        qb, qo = self._csr(q)
        rb, ro = self._csr(r)
        out = np.zeros((len(q), len(r)), np.int32)
        pr = np.ascontiguousarray(prior if len(prior) else [0], np.int32)
        err = self._err()
        rc = self.lib.ref_gapblock_matrix(
            qb, _p(qo, _c_u64p), len(q), rb, _p(ro, _c_u64p), len(r), ALPHABETS[alphabet], k,
            None if grid is None else _p(grid, _c_i32p), size, go, ge, _p(pr, _c_i32p),
            len(prior), width, threads, _p(out, _c_i32p), err, 512)
        if rc:
            raise ReferenceError_(rc, err.value.decode())
        return out

    def align_matrix(self, q, r, matrix, local=False, dist=False, go=11, ge=1, alphabet="aa",
                     threads=1):
        k, grid, size = self._matrix(matrix)
        qb, qo = self._csr(q)
        rb, ro = self._csr(r)
        out = np.zeros((len(q), len(r)), np.int32)
        err = self._err()
        rc = self.lib.ref_align_matrix(
            qb, _p(qo, _c_u64p), len(q), rb, _p(ro, _c_u64p), len(r), ALPHABETS[alphabet], k,
            None if grid is None else _p(grid, _c_i32p), size, int(local), int(dist), go, ge,
            threads, _p(out, _c_i32p), err, 512)
        if rc:
            raise ReferenceError_(rc, err.value.decode())
        return out

    def align_score(self, q, r, matrix, local=False, go=11, ge=1, alphabet="aa"):
        k, grid, size = self._matrix(matrix)
        out = ctypes.c_int32()
        err = self._err()
        qb, rb = _bytes(q), _bytes(r)
        rc = self.lib.ref_align_score(qb, len(qb), rb, len(rb), ALPHABETS[alphabet], k,
                                      None if grid is None else _p(grid, _c_i32p), size,
                                      int(local), go, ge, ctypes.byref(out), err, 512)
        if rc:
            raise ReferenceError_(rc, err.value.decode())
        return out.value

    def align_pair(self, q, r, matrix, local=False, go=11, ge=1, alphabet="aa"):
        """The reference's align_pair: (score, aligned_query, aligned_ref, ops)."""
        k, grid, size = self._matrix(matrix)
        qb, rb = _bytes(q), _bytes(r)
        cap = len(qb) + len(rb) + 1
        aq, ar, ops = (ctypes.create_string_buffer(cap) for _ in range(3))
        score = ctypes.c_int32()
        n = np.zeros(1, np.uint64)
        err = self._err()
        rc = self.lib.ref_align_pair(qb, len(qb), rb, len(rb), ALPHABETS[alphabet], k,
                                     None if grid is None else _p(grid, _c_i32p), size,
                                     int(local), go, ge, ctypes.byref(score), aq, ar, ops,
                                     _p(n, _c_u64p), err, 512)
        if rc:
            raise ReferenceError_(rc, err.value.decode())
        L = int(n[0])
        return score.value, aq.raw[:L].decode("latin-1"), ar.raw[:L].decode("latin-1"), \
            ops.raw[:L].decode()

    def align_pairs(self, q, r, pq, pr, matrix, local=False, go=11, ge=1, alphabet="aa",
                    threads=1):
        """align_pair over index pairs on `threads` workers; returns the total op count."""
        k, _, _ = self._matrix(matrix)
        qb, qo = self._csr(q)
        rb, ro = self._csr(r)
        pq = np.ascontiguousarray(pq, np.uint32)
        pr = np.ascontiguousarray(pr, np.uint32)
        total = np.zeros(1, np.uint64)
        err = self._err()
        rc = self.lib.ref_align_pairs(qb, _p(qo, _c_u64p), len(q), rb, _p(ro, _c_u64p), len(r),
                                      _p(pq, _c_u32p), _p(pr, _c_u32p), len(pq),
                                      ALPHABETS[alphabet], k, int(local), go, ge, threads,
                                      _p(total, _c_u64p), err, 512)
        if rc:
            raise ReferenceError_(rc, err.value.decode())
        return int(total[0])

    def index_align_pairs(self, q, r, pq, pr, matrix=None, go=1, ge=1, alphabet="aa",
                          threads=1, strings=True):
        """Index::align of (q[pq[t]], ref_id pr[t]) against Index.build(r): per pair
        (score, aligned_query, aligned_ref, ops), or with strings=False only the scores and
        the op counts (the CPU baseline)."""
        k, grid, size = self._matrix(matrix)
        qb, qo = self._csr(q)
        rb, ro = self._csr(r)
        pq = np.ascontiguousarray(pq, np.uint32)
        pr = np.ascontiguousarray(pr, np.uint32)
        P = len(pq)
        stride = (max(map(len, q), default=0) + max(map(len, r), default=0) + 1) if strings else 0
        bufs = [ctypes.create_string_buffer(max(1, P * stride)) for _ in range(3)] \
            if strings else [None, None, None]
        score = np.zeros(max(P, 1), np.int32)
        lens = np.zeros(max(P, 1), np.uint64)
        err = self._err()
        rc = self.lib.ref_index_align_pairs(
            qb, _p(qo, _c_u64p), len(q), rb, _p(ro, _c_u64p), len(r), _p(pq, _c_u32p),
            _p(pr, _c_u32p), P, ALPHABETS[alphabet], k,
            None if grid is None else _p(grid, _c_i32p), size, go, ge, threads,
            _p(score, _c_i32p), bufs[0], bufs[1], bufs[2], stride, _p(lens, _c_u64p), err, 512)
        if rc:
            raise ReferenceError_(rc, err.value.decode())
        if not strings:
            return score[:P], lens[:P]
        out = []
        for t in range(P):
            a, L = t * stride, int(lens[t])
            out.append((int(score[t]), bufs[1].raw[a:a + L].decode("latin-1"),
                        bufs[2].raw[a:a + L].decode("latin-1"), bufs[0].raw[a:a + L].decode()))
        return out

    def distance_matrix(self, which, a, b, threads=1):
        ab, ao = self._csr(a)
        bb, bo = self._csr(b)
        out = np.zeros((len(a), len(b)), np.int32)
        err = self._err()
        rc = self.lib.ref_distance_matrix(0 if which == "hamming" else 1, ab, _p(ao, _c_u64p),
                                          len(a), bb, _p(bo, _c_u64p), len(b), threads,
                                          _p(out, _c_i32p), err, 512)
        if rc:
            raise ReferenceError_(rc, err.value.decode())
        return out


def reference_available() -> bool:
    return REF_LIB.exists()


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1
