"""The host-memory entry points' transfer path: pinned staging, row blocks whose copy-out
overlaps the scoring of the next block, and the parallel copy into the caller's memory.
Matrices are checked against the device-memory entry points (which skip that path) and
against the oracle on sampled rows; the caller's buffers include page-aligned numpy arrays,
which the platform's in-place pinning of pageable memory is refused for."""
from __future__ import annotations

import ctypes
import random
import threading

import numpy as np
import pytest

from conftest import cdr3_like, rand_seq

pytestmark = pytest.mark.gpu


def _csr(seqs, enc):
    codes = np.array([enc[c] for c in "".join(seqs)] or [0], np.uint8)
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return codes, off


def test_c_abi_host_path_into_page_aligned_numpy_buffers(gpu, oracle_lib):
    st = gpu
    from seqtree_amd import _device
    from seqtree_amd._device import Status
    from seqtree_amd.gapblock import _prior_cube

    L = _device.lib()
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
    L.st_gapblock_matrix.restype = i32
    L.st_gapblock_matrix.argtypes = [vp, vp, u64, vp, vp, u64, vp, u32, i32, i32, vp, u32, vp,
                                     ctypes.POINTER(Status)]
    syms = st.amino_acids()
    enc = {c: i for i, c in enumerate(syms)}
    rng = random.Random(31)
    qs, rs = cdr3_like(rng, 1000, 10, 20), cdr3_like(rng, 10000, 10, 20)
    (qc, qo), (rc, ro) = _csr(qs, enc), _csr(rs, enc)
    m = st.SubstitutionMatrix.blosum62()
    pen = np.array([[m.penalty(a, b) for b in syms] for a in syms], np.int32)
    w = max(map(len, qs + rs))
    cube = np.array(_prior_cube(st.central_prior(21), w), np.int32)
    raw = np.empty(1000 * 10000 + 2048, np.int32)
    skip = (-raw.ctypes.data % 4096) // 4
    out = raw[skip:skip + 1000 * 10000].reshape(1000, 10000)  # a page-aligned 40 MB buffer
    assert out.ctypes.data % 4096 == 0
    stt = Status()
    rc_ = L.st_gapblock_matrix(qc.ctypes.data, qo.ctypes.data, 1000, rc.ctypes.data,
                               ro.ctypes.data, 10000, pen.ctypes.data, len(syms), 28, 1,
                               cube.ctypes.data, w, out.ctypes.data, ctypes.byref(stt))
    assert rc_ == 0, stt.message
    assert np.array_equal(out, np.asarray(st.score_matrix(qs, rs, m, 28, 1,
                                                          st.central_prior(21))))
    rows = [0, 517, 999]
    want = np.asarray(oracle_lib.gapblock_matrix([qs[i] for i in rows], rs, pen, 28, 1, cube, w))
    assert np.array_equal(out[rows], want.reshape(len(rows), -1))


def test_multi_block_matrices_equal_the_device_path(gpu, oracle_lib):
    """Outputs past one 128 MB row block and many 16 MB staging chunks."""
    st = gpu
    from seqtree_amd import _device
    from seqtree_amd._device import DeviceArray, Status

    L = _device.lib()
    vp, u64, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32
    L.st_hamming_matrix_device.restype = ctypes.c_int32
    L.st_hamming_matrix_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, vp,
                                           ctypes.POINTER(Status)]
    rng = random.Random(8)
    a = [rand_seq(rng, 12, 12, "ACGT") for _ in range(7000)]
    b = [rand_seq(rng, 12, 12, "ACGT") for _ in range(9000)]  # 252 MB of output
    got = np.asarray(st.distance.hamming_matrix(a, b))
    raw = {c: ord(c) for c in "ACGT"}
    (ac, ao), (bc, bo) = _csr(a, raw), _csr(b, raw)
    d = DeviceArray((len(a), len(b)), np.int32)
    dac, dao, dbc, dbo = (DeviceArray.from_numpy(x) for x in (ac, ao, bc, bo))
    stt = Status()
    assert L.st_hamming_matrix_device(dac.ptr, dao.ptr, len(a), 12, dbc.ptr, dbo.ptr, len(b), 12,
                                      d.ptr, None, ctypes.byref(stt)) == 0, stt.message
    _device.synchronize()
    assert np.array_equal(got, d.numpy())
    # Gotoh distances over three blocks, against the oracle on rows of each block
    m = st.SubstitutionMatrix.blosum62()
    syms = st.amino_acids()
    sim = np.array([[m.similarity(x, y) for y in syms] for x in syms], np.int32)
    qs, rs = cdr3_like(rng, 3000, 8, 25), cdr3_like(rng, 30000, 8, 25)  # 360 MB
    dist = np.asarray(st.pairwise.dist_matrix(qs, rs, m, mode="local"))
    rows = [0, 1001, 1999, 2999]
    code, want = oracle_lib.align_matrix([qs[i] for i in rows], rs, sim, True, True, 11, 1)
    assert code == 0 and np.array_equal(dist[rows], np.asarray(want).reshape(len(rows), -1))


def test_concurrent_host_calls(gpu, oracle_lib):
    """Two threads calling the host path at once (the GIL is released): each gets its own
    transfer context; the copy pool serves one copy at a time."""
    st = gpu
    m = st.SubstitutionMatrix.blosum62()
    rng = random.Random(4)
    sets = [(cdr3_like(rng, 700, 10, 20), cdr3_like(rng, 20000, 10, 20)) for _ in range(2)]
    want = [np.asarray(st.score_matrix(q, r, m, 28, 1)) for q, r in sets]
    got = [None, None]
    errors = []

    def work(k):
        try:
            for _ in range(3):
                got[k] = np.asarray(st.score_matrix(*sets[k], m, 28, 1))
                if not np.array_equal(got[k], want[k]):
                    errors.append(k)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors


def test_output_in_library_pinned_memory_equals_the_staged_path(gpu):
    """st_host_alloc memory takes the matrix by direct DMA (no staging): an output placed at an
    interior offset of such a buffer equals the same call into ordinary memory, the bytes
    around it stay untouched, and st_host_free releases it (NULL and foreign pointers are
    ignored)."""
    st = gpu
    from seqtree_amd import _device
    from seqtree_amd._device import Status

    L = _device.lib()
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
    L.st_host_alloc.restype = i32
    L.st_host_alloc.argtypes = [u64, ctypes.POINTER(vp), ctypes.POINTER(Status)]
    L.st_host_free.restype = None
    L.st_host_free.argtypes = [vp]
    L.st_hamming_matrix.restype = i32
    L.st_hamming_matrix.argtypes = [vp, vp, u64, vp, vp, u64, vp, ctypes.POINTER(Status)]
    rng = random.Random(12)
    a = [rand_seq(rng, 12, 12, "ACGT") for _ in range(3000)]
    b = [rand_seq(rng, 12, 12, "ACGT") for _ in range(12000)]  # 144 MB: two row blocks
    raw = {c: ord(c) for c in "ACGT"}
    (ac, ao), (bc, bo) = _csr(a, raw), _csr(b, raw)
    n = len(a) * len(b)
    stt = Status()
    ptr = vp()
    assert L.st_host_alloc(4 * (n + 64), ctypes.byref(ptr), ctypes.byref(stt)) == 0, stt.message
    try:
        buf = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int32)),
                                    shape=(n + 64,))
        buf[:] = -7
        out = buf[32:32 + n]
        assert L.st_hamming_matrix(ac.ctypes.data, ao.ctypes.data, len(a), bc.ctypes.data,
                                   bo.ctypes.data, len(b), out.ctypes.data,
                                   ctypes.byref(stt)) == 0, stt.message
        plain = np.empty(n, np.int32)
        assert L.st_hamming_matrix(ac.ctypes.data, ao.ctypes.data, len(a), bc.ctypes.data,
                                   bo.ctypes.data, len(b), plain.ctypes.data,
                                   ctypes.byref(stt)) == 0, stt.message
        assert np.array_equal(out, plain)
        assert np.all(buf[:32] == -7) and np.all(buf[32 + n:] == -7)
        want = np.asarray(st.distance.hamming_matrix(a[:5], b)).ravel()
        assert np.array_equal(out[:want.size], want)
    finally:
        L.st_host_free(ptr)
    L.st_host_free(None)
    L.st_host_free(ctypes.c_void_p(plain.ctypes.data))  # not the library's: ignored
