"""The allocation-free gap-block matrix entry point (st_gapblock_matrix_device_noalloc) against
the allocating one (st_gapblock_matrix_device), through the C-ABI: bit-identical matrices on
consecutive calls with a smaller second input and different tables (so stale scratch contents
cannot leak into a result), on the twin-row short path and on the long-query path, and a set
too large for the static scratch refused with ST_INVALID_ARGUMENT before any work."""
from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

from conftest import cdr3_like, rand_seq

pytestmark = pytest.mark.gpu


def _lib():
    from seqtree_amd import _device
    from seqtree_amd._device import Status

    L = _device.lib()
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
    args = [vp, vp, u64, u32, vp, vp, u64, u32, vp, u32, i32, i32, vp, u32, vp, vp,
            ctypes.POINTER(Status)]
    for name in ("st_gapblock_matrix_device", "st_gapblock_matrix_device_noalloc"):
        fn = getattr(L, name)
        fn.restype = i32
        fn.argtypes = args
    return L


class _Set:
    """A CSR sequence set in device memory (codec codes)."""

    def __init__(self, seqs, codec):
        from seqtree_amd._device import DeviceArray

        self.n = len(seqs)
        self.max_len = max(map(len, seqs), default=0)
        codes = np.array([codec[c] for c in "".join(seqs)] or [0], np.uint8)
        off = np.zeros(self.n + 1, np.uint64)
        off[1:] = np.cumsum([len(s) for s in seqs])
        self.codes = DeviceArray.from_numpy(codes)
        self.off = DeviceArray.from_numpy(off)


def _call(L, name, q, r, pen, go, ge, cube, width, out):
    from seqtree_amd import _device
    from seqtree_amd._device import Status

    st = Status()
    rc = getattr(L, name)(q.codes.ptr, q.off.ptr, q.n, q.max_len, r.codes.ptr, r.off.ptr, r.n,
                          r.max_len, pen.ctypes.data, pen.shape[0], go, ge,
                          None if cube is None else cube.ctypes.data, width, out.ptr, None,
                          ctypes.byref(st))
    _device.synchronize()
    return rc, st.message.decode()


def _pen(st, m):
    syms = st.amino_acids()
    if m is None:
        return np.array([[int(a != b) for b in syms] for a in syms], np.int32)
    return np.array([[m.penalty(a, b) for b in syms] for a in syms], np.int32)


def test_noalloc_equals_the_allocating_entry_point(gpu, oracle_lib):
    st = gpu
    from seqtree_amd._device import DeviceArray
    from seqtree_amd.gapblock import _prior_cube

    L = _lib()
    codec = {c: i for i, c in enumerate(st.amino_acids())}
    rng = random.Random(2024)
    cases = [
        # (queries, refs, matrix, gap_open, prior): twin short path, then a smaller input with
        # other tables, then queries past the short launch (profile path), unit cost
        (cdr3_like(rng, 300, 6, 30), cdr3_like(rng, 5000, 6, 30),
         st.SubstitutionMatrix.blosum62(), 28, st.central_prior(21)),
        (cdr3_like(rng, 40, 8, 18), cdr3_like(rng, 700, 8, 18),
         st.SubstitutionMatrix.structural(), 14, st.frame_prior(3, 4)),
        ([rand_seq(rng, 33, 60) for _ in range(50)], [rand_seq(rng, 10, 60) for _ in range(800)],
         None, 1, None),
        (cdr3_like(rng, 17, 6, 12), cdr3_like(rng, 90, 6, 12),
         st.SubstitutionMatrix.pam250(), 24, st.positions_prior((3, 4, -4, -3))),
    ]
    for n, (qs, rs, m, go, prior) in enumerate(cases):
        print(f"no-alloc case {n}: {len(qs)} x {len(rs)}", flush=True)
        q, r = _Set(qs, codec), _Set(rs, codec)
        pen = _pen(st, m)
        width = max(q.max_len, r.max_len)
        cube = None if prior is None else np.array(_prior_cube(prior, width), np.int32)
        want_d = DeviceArray((len(qs), len(rs)), np.int32)
        got_d = DeviceArray((len(qs), len(rs)), np.int32)
        rc, msg = _call(L, "st_gapblock_matrix_device", q, r, pen, go, 1, cube, width, want_d)
        assert rc == 0, msg
        rc, msg = _call(L, "st_gapblock_matrix_device_noalloc", q, r, pen, go, 1, cube, width,
                        got_d)
        assert rc == 0, msg
        want, got = want_d.numpy(), got_d.numpy()
        assert np.array_equal(got, want), n
        rows = sorted(rng.sample(range(len(qs)), 3))
        ref = oracle_lib.gapblock_matrix([qs[i] for i in rows], rs, pen, go, 1,
                                         cube, width if cube is not None else 0)
        assert np.array_equal(got[rows], np.asarray(ref).reshape(len(rows), len(rs))), n


def test_noalloc_refuses_sets_past_its_scratch(gpu):
    st = gpu
    from seqtree_amd._device import DeviceArray

    L = _lib()
    codec = {c: i for i, c in enumerate(st.amino_acids())}
    rng = np.random.default_rng(3)
    K = 800_000  # 30 residues each: about 35 MB of packed references, past 32 MiB
    codes = DeviceArray.from_numpy(rng.integers(0, 20, K * 30).astype(np.uint8))
    off = DeviceArray.from_numpy(np.arange(K + 1, dtype=np.uint64) * 30)
    big = _Set.__new__(_Set)
    big.n, big.max_len, big.codes, big.off = K, 30, codes, off
    q = _Set(["CASSLGQAYEQYF"], codec)
    out = DeviceArray((1, K), np.int32)
    pen = _pen(st, st.SubstitutionMatrix.blosum62())
    rc, msg = _call(L, "st_gapblock_matrix_device_noalloc", q, big, pen, 28, 1, None, 0, out)
    assert rc == 1 and "static scratch" in msg, (rc, msg)
    # the entry point still works afterwards, and agrees with the allocating one
    small = _Set(["CASSLGQAYEQF", "CASSPGQAYEQYF", "CSARDRTGNGYTF"], codec)
    a = DeviceArray((1, 3), np.int32)
    b = DeviceArray((1, 3), np.int32)
    assert _call(L, "st_gapblock_matrix_device_noalloc", q, small, pen, 28, 1, None, 0, a)[0] == 0
    assert _call(L, "st_gapblock_matrix_device", q, small, pen, 28, 1, None, 0, b)[0] == 0
    assert np.array_equal(a.numpy(), b.numpy())
