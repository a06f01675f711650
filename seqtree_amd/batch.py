"""Ragged batches of gap-block jobs in one persistent, graph-captured launch.

Each job is an ordinary gap-block score matrix (:func:`seqtree_amd.gapblock.score_matrix`
semantics: its own queries, references, matrix, gap costs and prior); the batch scores all of
them with one persistent kernel whose tile shape is chosen by a small autotuner, and replays
the captured hipGraph on every :meth:`GapblockBatch.run`. The reference has no batch API; each
job's result equals what its ``score_matrix`` would return.

Device memory for the encoded sequences and the output matrices comes from the library's
own allocator (``seqtree_amd._device``); the kernels and the plan live in
``libseqtree_hip.so`` behind the C-ABI of ``include/seqtree_hip.h``. Runs use the device's
null stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

from . import _core
from ._device import DeviceArray, synchronize
from .gapblock import GapPrior, _default_gap_open, _prior_cube

_LIB = Path(__file__).resolve().parent / "libseqtree_hip.so"


class _Status(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("message", ctypes.c_char * 256)]


class _Job(ctypes.Structure):
    _fields_ = [
        ("q_codes", ctypes.c_void_p), ("q_offsets", ctypes.c_void_p),
        ("n_queries", ctypes.c_uint64), ("q_max_len", ctypes.c_uint32),
        ("r_codes", ctypes.c_void_p), ("r_offsets", ctypes.c_void_p),
        ("n_refs", ctypes.c_uint64), ("r_max_len", ctypes.c_uint32),
        ("pen", ctypes.c_void_p), ("alphabet_size", ctypes.c_uint32),
        ("gap_open", ctypes.c_int32), ("gap_extend", ctypes.c_int32),
        ("prior", ctypes.c_void_p), ("prior_width", ctypes.c_uint32),
        ("out", ctypes.c_void_p),
    ]


def _lib():
    lib = ctypes.CDLL(str(_LIB))
    vp = ctypes.c_void_p
    lib.st_gapblock_batch_plan_create.restype = ctypes.c_int32
    lib.st_gapblock_batch_plan_create.argtypes = [ctypes.POINTER(_Job), ctypes.c_uint32,
                                                  ctypes.c_int32, vp, ctypes.POINTER(vp),
                                                  ctypes.POINTER(_Status)]
    lib.st_gapblock_batch_plan_run.restype = ctypes.c_int32
    lib.st_gapblock_batch_plan_run.argtypes = [vp, vp, ctypes.POINTER(_Status)]
    lib.st_gapblock_batch_plan_tiles.restype = ctypes.c_int32
    lib.st_gapblock_batch_plan_tiles.argtypes = [vp] + [ctypes.POINTER(ctypes.c_uint32)] * 3 + [
        ctypes.POINTER(ctypes.c_uint64)]
    lib.st_gapblock_batch_plan_twin_jobs.restype = ctypes.c_int32
    lib.st_gapblock_batch_plan_twin_jobs.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
    lib.st_gapblock_batch_plan_destroy.argtypes = [vp]
    return lib


@dataclass
class EncodedSet:
    """Strings already in codec codes, as CSR: ``codes`` (uint8) and ``offsets`` (uint64,
    n + 1 entries). Lets large batches skip the per-string Python encoding."""

    codes: np.ndarray
    offsets: np.ndarray

    def __len__(self) -> int:
        return len(self.offsets) - 1

    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64))


@dataclass
class GapblockJob:
    """One gap-block score matrix: the arguments of :func:`gapblock.score_matrix`.
    ``queries`` and ``refs`` are strings or an :class:`EncodedSet`."""

    queries: Sequence[str] | EncodedSet
    refs: Sequence[str] | EncodedSet
    matrix: _core.SubstitutionMatrix | None = None
    gap_open: int | None = None
    gap_extend: int = 1
    gap_prior: GapPrior | None = None
    alphabet: str = "aa"


def _encode(seqs, symbols: str) -> tuple[np.ndarray, np.ndarray]:
    if isinstance(seqs, EncodedSet):
        codes = np.ascontiguousarray(seqs.codes, np.uint8)
        off = np.ascontiguousarray(seqs.offsets, np.uint64)
        if off[0] != 0 or (np.diff(off.astype(np.int64)) < 0).any() or int(off[-1]) > len(codes):
            raise ValueError("malformed EncodedSet offsets")
        if int(off[-1]) and int(codes[: int(off[-1])].max()) >= len(symbols):
            raise ValueError("EncodedSet code outside the alphabet")
        return (codes if len(codes) else np.zeros(1, np.uint8)), off
    lut = np.full(256, 255, np.uint8)
    for code, ch in enumerate(symbols):
        lut[ord(ch)] = code
        lut[ord(ch.lower())] = code
    raw = [s.encode("latin-1") for s in seqs]
    off = np.zeros(len(raw) + 1, np.uint64)
    off[1:] = np.cumsum([len(b) for b in raw])
    data = lut[np.frombuffer(b"".join(raw) or b"\0", np.uint8)]
    if (data[: int(off[-1])] == 255).any():
        bad = next(ch for s in seqs for ch in s if lut[ord(ch)] == 255)
        raise ValueError(f"symbol '{bad}' is not in the alphabet")
    return data, off


class GapblockBatch:
    """A captured batch of gap-block jobs on the GPU.

    ``autotune=True`` times the candidate tile shapes on this batch (once per process and
    workload shape) before capturing the graph. ``run()`` replays it on ``stream`` (a raw
    ``hipStream_t`` handle, default the null stream); ``results()`` copies the matrices back.
    """

    def __init__(self, jobs: Sequence[GapblockJob], autotune: bool = False, stream: int = 0):
        self._lib = _lib()
        self._stream = stream
        self._keep = []  # host tables and device buffers referenced by the plan
        self.shapes = []
        self.outputs = []
        descs = (_Job * len(jobs))()
        for t, job in enumerate(jobs):
            symbols = _core.alphabet_symbols(job.alphabet)
            A = len(symbols)
            m = job.matrix
            if m is not None and m.size() != A:
                raise ValueError("matrix size does not match the alphabet")
            go = _default_gap_open(m) if job.gap_open is None else job.gap_open
            if go < 0 or job.gap_extend < 0:
                raise ValueError("gap_open and gap_extend must be >= 0")
            pen = np.array([[m.penalty(a, b) if m is not None else int(a != b)
                             for b in symbols] for a in symbols], np.int32) \
                if (m is None or job.alphabet == "aa") else \
                (1 - np.eye(A, dtype=np.int32))
            qb, qo = _encode(job.queries, symbols)
            rb, ro = _encode(job.refs, symbols)
            q_lens = np.diff(qo.astype(np.int64))
            r_lens = np.diff(ro.astype(np.int64))
            width = int(max(q_lens.max(initial=0), r_lens.max(initial=0)))
            cube = (np.array(_prior_cube(job.gap_prior, width), np.int32)
                    if job.gap_prior is not None else None)
            tq_b, tq_o = DeviceArray.from_numpy(qb), DeviceArray.from_numpy(qo)
            tr_b, tr_o = DeviceArray.from_numpy(rb), DeviceArray.from_numpy(ro)
            out = DeviceArray((len(job.queries), len(job.refs)), np.int32)
            self._keep += [pen, cube, tq_b, tq_o, tr_b, tr_o]
            self.outputs.append(out)
            self.shapes.append((len(job.queries), len(job.refs)))
            d = descs[t]
            d.q_codes, d.q_offsets = tq_b.ptr, tq_o.ptr
            d.n_queries = len(job.queries)
            d.q_max_len = int(q_lens.max(initial=0))
            d.r_codes, d.r_offsets = tr_b.ptr, tr_o.ptr
            d.n_refs = len(job.refs)
            d.r_max_len = int(r_lens.max(initial=0))
            d.pen = pen.ctypes.data
            d.alphabet_size = A
            d.gap_open, d.gap_extend = go, job.gap_extend
            d.prior = cube.ctypes.data if cube is not None else None
            d.prior_width = width
            d.out = out.ptr
        st = _Status()
        plan = ctypes.c_void_p()
        rc = self._lib.st_gapblock_batch_plan_create(descs, len(jobs), int(autotune),
                                                     ctypes.c_void_p(self._stream),
                                                     ctypes.byref(plan), ctypes.byref(st))
        if rc != 0:
            raise RuntimeError(f"st_gapblock_batch_plan_create: {st.message.decode()}")
        self._plan = plan

    def tiles(self) -> dict:
        q, th, wg, units = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        self._lib.st_gapblock_batch_plan_tiles(self._plan, ctypes.byref(q), ctypes.byref(th),
                                               ctypes.byref(wg), ctypes.byref(units))
        return {"queries_per_unit": q.value, "threads": th.value, "workgroups": wg.value,
                "work_units": units.value}

    def twin_jobs(self) -> int:
        """How many jobs run twin rows (two queries per 16-bit packed table)."""
        n = ctypes.c_uint32()
        self._lib.st_gapblock_batch_plan_twin_jobs(self._plan, ctypes.byref(n))
        return n.value

    def run(self, stream: int | None = None) -> None:
        st = _Status()
        s = self._stream if stream is None else stream
        rc = self._lib.st_gapblock_batch_plan_run(self._plan, ctypes.c_void_p(s),
                                                  ctypes.byref(st))
        if rc != 0:
            raise RuntimeError(f"st_gapblock_batch_plan_run: {st.message.decode()}")

    def results(self) -> list[np.ndarray]:
        synchronize()
        return [o.numpy() for o in self.outputs]

    def close(self) -> None:
        if getattr(self, "_plan", None):
            synchronize()
            self._lib.st_gapblock_batch_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass
