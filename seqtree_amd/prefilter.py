"""Learned prefilter for thresholded gap-block search on the GPU.

A small 1D-conv encoder (embedding 32 -> three width-5 convolutions of 128 channels with
ReLU -> mean pool -> 128-d) is trained so that the squared embedding distance
``|e(a) - e(b)|^2 = s(a,a) + s(b,b) - 2 s(a,b)`` tracks the exact gap-block score of (a, b),
on pairs whose labels come from the exact gap-block kernel itself. A search then

1. embeds the queries and the references (BF16 MFMA kernels, FP32 accumulation),
2. screens every (query, reference) pair with one BF16 MFMA Gram GEMM whose epilogue keeps
   the pairs with estimated distance <= tau (the cutoff plus a calibrated margin),
3. rescores the survivors with the exact gap-block pair kernel and keeps those <= cutoff.

Every reported score is exact; what the screen can cost is recall, which the exhaustive exact
pass (``st_gapblock_threshold_device``) measures. Training pairs are mined by that same
exhaustive pass over the training set (every close pair, a band of near misses, random
pairs), and the margin is calibrated twice: on held-out training pairs, and at search time on
the exact hits of a small exhaustively searched sample of the queries (``audit``). The reference has no prefilter: its
``GapBlockIndex.search`` (gapblock.py:567-673) answers the same question with a
deletion-variant trie.

All kernels and the model live in ``libseqtree_hip.so`` (``st_prefilter_*`` in
``include/seqtree_hip.h``); this module holds the data pipeline, the training loop and the
margin calibration.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _device
from ._device import DeviceArray, Status, check
from .gapblock import _prior_cube, central_prior
from .synthetic import AA_ORDER

EMB = 128
# Distances are learned in units of this many gap-block score points.
SCORE_UNIT = 100.0
# Share of the calibration pairs' exact hits (score <= cutoff) whose model distance the
# screening threshold keeps: 0.997 leaves margin over a 99.5 % recall target, since the
# threshold is also raised past the cutoff and past the held-out calibration's quantile.
DEFAULT_QUANTILE = 0.997
# The audit of a search (Searcher.query) wants at least this many exact hits behind its
# quantile, and searches at most this share of the queries exhaustively to get them.
AUDIT_MIN_HITS = 1000
AUDIT_MAX_SHARE = 0.1


def _bind():
    L = _device.lib()
    if getattr(L, "_pf_bound", False):
        return L
    vp, u32, u64, i32, f32 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                              ctypes.c_float)
    S = ctypes.POINTER(Status)
    L.st_prefilter_param_count.restype = u64
    L.st_prefilter_create.restype = i32
    L.st_prefilter_create.argtypes = [u64, ctypes.POINTER(vp), S]
    L.st_prefilter_destroy.argtypes = [vp]
    for name in ("st_prefilter_get_params", "st_prefilter_set_params", "st_prefilter_get_grads"):
        fn = getattr(L, name)
        fn.restype = i32
        fn.argtypes = [vp, vp, u64, S]
    L.st_prefilter_train_step.restype = i32
    L.st_prefilter_train_step.argtypes = [vp, vp, vp, u32, u32, vp, vp, vp, vp, u32, f32,
                                          ctypes.POINTER(f32), vp, S]
    L.st_prefilter_encode.restype = i32
    L.st_prefilter_encode.argtypes = [vp, vp, vp, u32, u32, vp, vp, vp, S]
    L.st_prefilter_screen.restype = i32
    L.st_prefilter_screen.argtypes = [vp, vp, u32, vp, vp, u32, f32, u64, vp, vp,
                                      ctypes.POINTER(u64), vp, S]
    L.st_prefilter_pair_distance.restype = i32
    L.st_prefilter_pair_distance.argtypes = [vp, vp, vp, vp, vp, vp, u64, vp, vp, S]
    L.st_select_le.restype = i32
    L.st_select_le.argtypes = [vp, u64, i32, u64, vp, ctypes.POINTER(u64), vp, S]
    L.st_select_pairs_le.restype = i32
    L.st_select_pairs_le.argtypes = [vp, vp, vp, u64, i32, u64, vp, vp, vp, ctypes.POINTER(u64),
                                     vp, S]
    L.st_gapblock_pairs_device.restype = i32
    L.st_gapblock_pairs_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, u32, i32, i32,
                                           vp, u32, vp, vp, u64, vp, vp, S]
    L.st_gapblock_threshold_device.restype = i32
    L.st_gapblock_threshold_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, u32, i32,
                                               i32, vp, u32, i32, u64, vp, vp, vp, vp, vp, S]
    L._pf_bound = True
    return L


@dataclass
class DeviceSet:
    """A CSR set of codec codes resident on the GPU."""

    codes: DeviceArray
    offsets: DeviceArray
    n: int
    max_len: int
    host: tuple | None = None  # (codes, offsets) on the host, when uploaded from there

    @classmethod
    def upload(cls, codes: np.ndarray, offsets: np.ndarray) -> "DeviceSet":
        lens = np.diff(offsets.astype(np.int64))
        c = codes if len(codes) else np.zeros(1, np.uint8)
        return cls(DeviceArray.from_numpy(np.ascontiguousarray(c, np.uint8)),
                   DeviceArray.from_numpy(np.ascontiguousarray(offsets, np.uint64)),
                   len(offsets) - 1, int(lens.max(initial=0)), (codes, offsets))

    def subset(self, idx: np.ndarray) -> "DeviceSet":
        """The sequences idx (host indices) as a new device set; needs the host copies."""
        codes, off = self.host
        off = off.astype(np.int64)
        lens = off[idx + 1] - off[idx]
        sub_off = np.zeros(len(idx) + 1, np.uint64)
        sub_off[1:] = np.cumsum(lens)
        total = int(sub_off[-1])
        take = np.repeat(off[idx] - sub_off[:-1].astype(np.int64), lens) + np.arange(total)
        return DeviceSet.upload(codes[take], sub_off)


@dataclass
class GapblockParams:
    """Exact gap-block scoring parameters (pen grid in codec order, gaps, prior cube)."""

    pen: np.ndarray
    gap_open: int
    gap_extend: int
    cube: np.ndarray | None
    width: int

    @classmethod
    def blosum62_central(cls, width: int) -> "GapblockParams":
        """BLOSUM62, gap_open = 2 * scale = 28, gap_extend 1, central prior 1.5 * scale = 21."""
        from . import _core

        m = _core.SubstitutionMatrix.blosum62()
        pen = np.array([[m.penalty(a, b) for b in AA_ORDER] for a in AA_ORDER], np.int32)
        lam = int(1.5 * m.scale())
        cube = np.array(_prior_cube(central_prior(lam), width), np.int32)
        return cls(pen, 2 * m.scale(), 1, cube, width)


def exact_pairs(q: DeviceSet, r: DeviceSet, pa: DeviceArray, pb: DeviceArray, n_pairs: int,
                gp: GapblockParams, out: DeviceArray | None = None) -> DeviceArray:
    """Exact gap-block scores of pairs (q[pa[t]], r[pb[t]]) on the GPU (device int32)."""
    L = _bind()
    out = out if out is not None else DeviceArray(max(n_pairs, 1), np.int32)
    st = Status()
    rc = L.st_gapblock_pairs_device(
        q.codes.ptr, q.offsets.ptr, q.n, q.max_len, r.codes.ptr, r.offsets.ptr, r.n, r.max_len,
        gp.pen.ctypes.data, len(AA_ORDER), gp.gap_open, gp.gap_extend,
        gp.cube.ctypes.data if gp.cube is not None else None, gp.width, pa.ptr, pb.ptr, n_pairs,
        out.ptr, None, ctypes.byref(st))
    check(rc, st, "st_gapblock_pairs_device")
    return out


def exact_threshold(q: DeviceSet, r: DeviceSet, cutoff: int, gp: GapblockParams,
                    capacity: int = 1 << 22, buffers: dict | None = None):
    """Exhaustive exact search: every pair with score <= cutoff, as host arrays
    (query, ref, score). Retries with more room when the first capacity is short. `buffers`
    (a dict, initially empty) keeps the device buffers for the next call."""
    L = _bind()
    bufs = buffers if buffers is not None else {}
    if "count" not in bufs:
        bufs["count"] = DeviceArray(1, np.uint64)  # the entry point counts in device memory
    count = bufs["count"]
    while True:
        if bufs.get("cap", 0) < capacity:
            bufs["hits"] = (DeviceArray(capacity, np.uint32), DeviceArray(capacity, np.uint32),
                            DeviceArray(capacity, np.int32))
            bufs["cap"] = capacity
        capacity = bufs["cap"]
        hq, hr, hs = bufs["hits"]
        st = Status()
        rc = L.st_gapblock_threshold_device(
            q.codes.ptr, q.offsets.ptr, q.n, q.max_len, r.codes.ptr, r.offsets.ptr, r.n,
            r.max_len, gp.pen.ctypes.data, len(AA_ORDER), gp.gap_open, gp.gap_extend,
            gp.cube.ctypes.data if gp.cube is not None else None, gp.width, cutoff, capacity,
            hq.ptr, hr.ptr, hs.ptr, count.ptr, None, ctypes.byref(st))
        check(rc, st, "st_gapblock_threshold_device")
        k = int(count.numpy()[0])
        if k <= capacity:
            return hq.numpy(k), hr.numpy(k), hs.numpy(k)
        capacity = k


class Prefilter:
    """The encoder model on the GPU (FP32 master parameters, BF16 operand copies)."""

    def __init__(self, seed: int = 0):
        self._L = _bind()
        self._h = ctypes.c_void_p()
        st = Status()
        check(self._L.st_prefilter_create(seed, ctypes.byref(self._h), ctypes.byref(st)), st,
              "st_prefilter_create")
        self.n_params = int(self._L.st_prefilter_param_count())

    def close(self):
        if self._h:
            self._L.st_prefilter_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        st = Status()
        check(self._L.st_prefilter_get_params(self._h, out.ctypes.data, out.size,
                                              ctypes.byref(st)), st, "get_params")
        return out

    def grads(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        st = Status()
        check(self._L.st_prefilter_get_grads(self._h, out.ctypes.data, out.size,
                                             ctypes.byref(st)), st, "get_grads")
        return out

    def load(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        st = Status()
        check(self._L.st_prefilter_set_params(self._h, p.ctypes.data, p.size, ctypes.byref(st)),
              st, "set_params")

    def train_step(self, seqs: DeviceSet, pa: DeviceArray, pb: DeviceArray, y: DeviceArray,
                   w: DeviceArray, n_pairs: int, lr: float, max_len: int | None = None) -> float:
        loss = ctypes.c_float()
        st = Status()
        rc = self._L.st_prefilter_train_step(
            self._h, seqs.codes.ptr, seqs.offsets.ptr, seqs.n, max_len or seqs.max_len, pa.ptr,
            pb.ptr, y.ptr, w.ptr, n_pairs, lr, ctypes.byref(loss), None, ctypes.byref(st))
        check(rc, st, "st_prefilter_train_step")
        return float(loss.value)

    def encode(self, seqs: DeviceSet, max_len: int | None = None, out: tuple | None = None):
        """(embeddings n x 128 BF16 as a uint16 device array, squared norms float32). `out`
        may pass (embeddings, norms) buffers of at least n rows to write into."""
        if out is not None and out[0].shape[0] >= seqs.n and out[1].shape[0] >= seqs.n:
            emb, norm = out
        else:
            emb = DeviceArray((seqs.n, EMB), np.uint16)
            norm = DeviceArray(seqs.n, np.float32)
        st = Status()
        rc = self._L.st_prefilter_encode(self._h, seqs.codes.ptr, seqs.offsets.ptr, seqs.n,
                                         max_len or seqs.max_len, emb.ptr, norm.ptr, None,
                                         ctypes.byref(st))
        check(rc, st, "st_prefilter_encode")
        return emb, norm


def screen(q_emb, q_norm, nq: int, r_emb, r_norm, nr: int, tau: float, capacity: int,
           out: tuple | None = None):
    """Device (query, ref) index arrays of the pairs with estimated distance <= tau (distance
    units), and their count. `out` may pass preallocated (hit_q, hit_r) buffers of `capacity`
    entries; the buffers grow (and the screen repeats) when the capacity is short."""
    L = _bind()
    while True:
        if out is None or out[0].shape[0] < capacity:
            out = (DeviceArray(capacity, np.uint32), DeviceArray(capacity, np.uint32))
        hq, hr = out
        count = ctypes.c_uint64()
        st = Status()
        rc = L.st_prefilter_screen(q_emb.ptr, q_norm.ptr, nq, r_emb.ptr, r_norm.ptr, nr, tau,
                                   hq.shape[0], hq.ptr, hr.ptr, ctypes.byref(count), None,
                                   ctypes.byref(st))
        check(rc, st, "st_prefilter_screen")
        if count.value <= hq.shape[0]:
            return hq, hr, int(count.value)
        capacity = int(count.value)
        out = None


def pair_distances(q_emb, q_norm, r_emb, r_norm, pq: np.ndarray, pr: np.ndarray) -> np.ndarray:
    """Model distances (score units) of the pairs (queries pq[t], references pr[t]) of two
    encoded sets, as the screen estimates them."""
    L = _bind()
    n = len(pq)
    if n == 0:
        return np.zeros(0, np.float32)
    dq = DeviceArray.from_numpy(np.ascontiguousarray(pq, np.uint32))
    dr = DeviceArray.from_numpy(np.ascontiguousarray(pr, np.uint32))
    out = DeviceArray(n, np.float32)
    st = Status()
    check(L.st_prefilter_pair_distance(q_emb.ptr, q_norm.ptr, r_emb.ptr, r_norm.ptr, dq.ptr,
                                       dr.ptr, n, out.ptr, None, ctypes.byref(st)), st,
          "st_prefilter_pair_distance")
    return out.numpy() * np.float32(SCORE_UNIT)


def select_le(values: DeviceArray, n: int, cutoff: int) -> np.ndarray:
    """Host indices t < n with values[t] <= cutoff (sorted)."""
    L = _bind()
    cap = 1 << 16
    while True:
        idx = DeviceArray(cap, np.uint32)
        count = ctypes.c_uint64()
        st = Status()
        check(L.st_select_le(values.ptr, n, cutoff, cap, idx.ptr, ctypes.byref(count), None,
                             ctypes.byref(st)), st, "st_select_le")
        if count.value <= cap:
            return np.sort(idx.numpy(count.value))
        cap = int(count.value)


# ------------------------------------------------------------------------------ training
@dataclass
class TrainingData:
    seqs_codes: np.ndarray
    seqs_off: np.ndarray
    pair_a: np.ndarray
    pair_b: np.ndarray
    score: np.ndarray  # exact gap-block scores


def label_pairs(codes: np.ndarray, off: np.ndarray, pa: np.ndarray, pb: np.ndarray,
                gp: GapblockParams) -> np.ndarray:
    """Exact gap-block scores of (s[pa[t]], s[pb[t]]) within one set, on the GPU."""
    ds = DeviceSet.upload(codes, off)
    da = DeviceArray.from_numpy(pa.astype(np.uint32))
    db = DeviceArray.from_numpy(pb.astype(np.uint32))
    return exact_pairs(ds, ds, da, db, len(pa), gp).numpy()[:len(pa)]


def sample_pairs(rng: np.random.Generator, fam: np.ndarray, n_pairs: int, near_frac: float):
    """near_frac of the pairs inside a family (distinct members), the rest uniform."""
    n = len(fam)
    n_near = int(n_pairs * near_frac)
    starts = np.searchsorted(fam, fam)  # first member of each sequence's family
    sizes = np.bincount(fam)[fam]
    a = rng.integers(0, n, size=n_near)
    off = rng.integers(1, np.maximum(sizes[a], 2), size=n_near)
    b = starts[a] + (a - starts[a] + off) % np.maximum(sizes[a], 1)
    ra = rng.integers(0, n, size=n_pairs - n_near)
    rb = rng.integers(0, n, size=n_pairs - n_near)
    return np.concatenate([a, ra]).astype(np.int64), np.concatenate([b, rb]).astype(np.int64)


def mine_pairs(rng: np.random.Generator, codes: np.ndarray, off: np.ndarray,
               gp: GapblockParams, n_pairs: int, near_frac: float, cutoff: int,
               mine_cutoff: int):
    """Training pairs of one set: every pair with exact score <= mine_cutoff comes from one
    exhaustive thresholded pass over the set (the exact kernel), and near_frac of the pairs
    are drawn from those -- every pair <= cutoff up to half of them, the rest from the band
    (cutoff, mine_cutoff] -- the remainder uniformly from all pairs. Returns (a, b, the
    number of mined pairs <= cutoff)."""
    s_all = DeviceSet.upload(codes, off)
    a, b, sc = exact_threshold(s_all, s_all, mine_cutoff, gp)
    keep = a < b  # each unordered pair once, no self pairs
    a, b, sc = a[keep], b[keep], sc[keep]
    close = np.flatnonzero(sc <= cutoff)
    band = np.flatnonzero(sc > cutoff)
    n_near = int(n_pairs * near_frac)
    k_close = min(len(close), n_near // 2)
    pick = np.concatenate([rng.choice(close, k_close, replace=False),
                           rng.choice(band, min(len(band), n_near - k_close), replace=False)])
    n = len(off) - 1
    n_rand = n_pairs - len(pick)
    pa = np.concatenate([a[pick], rng.integers(0, n, size=n_rand)]).astype(np.int64)
    pb = np.concatenate([b[pick], rng.integers(0, n, size=n_rand)]).astype(np.int64)
    return pa, pb, len(close)


def pair_weights(score: np.ndarray) -> np.ndarray:
    """Relative squared error: (d - y)^2 / (y + 0.1)^2 in distance units, so close pairs --
    the ones a cutoff search keeps -- dominate the fit."""
    y = score / SCORE_UNIT
    return (1.0 / (y + 0.1) ** 2).astype(np.float32)


def train(model: Prefilter, data: TrainingData, epochs: int, batch_pairs: int = 4096,
          lr: float = 2e-3, seed: int = 0, max_len: int | None = None, log=None) -> list[float]:
    """Adam over the labelled pairs; each step embeds the distinct sequences of its batch."""
    rng = np.random.default_rng(seed)
    lens = np.diff(data.seqs_off.astype(np.int64))
    max_len = max_len or int(lens.max())
    y_all = (data.score / SCORE_UNIT).astype(np.float32)
    w_all = pair_weights(data.score)
    from .synthetic import padded

    table = padded(data.seqs_codes, data.seqs_off, max_len)
    n = len(data.pair_a)
    steps = (n + batch_pairs - 1) // batch_pairs
    total = epochs * steps
    # device staging buffers sized for the largest batch
    cap_seqs = 2 * batch_pairs
    d_codes = DeviceArray(cap_seqs * max_len, np.uint8)
    d_off = DeviceArray(cap_seqs + 1, np.uint64)
    d_pa, d_pb = DeviceArray(batch_pairs, np.uint32), DeviceArray(batch_pairs, np.uint32)
    d_y, d_w = DeviceArray(batch_pairs, np.float32), DeviceArray(batch_pairs, np.float32)
    losses = []
    k = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        ep_loss = 0.0
        for s in range(steps):
            sel = order[s * batch_pairs:(s + 1) * batch_pairs]
            a, b = data.pair_a[sel], data.pair_b[sel]
            uniq, inv = np.unique(np.concatenate([a, b]), return_inverse=True)
            ia, ib = inv[:len(sel)], inv[len(sel):]
            ul = lens[uniq]
            off = np.zeros(len(uniq) + 1, np.uint64)
            off[1:] = np.cumsum(ul)
            codes = table[uniq][np.arange(max_len)[None, :] < ul[:, None]]
            _upload(d_codes, codes)
            _upload(d_off, off)
            for dst, src in ((d_pa, ia.astype(np.uint32)), (d_pb, ib.astype(np.uint32)),
                             (d_y, y_all[sel]), (d_w, w_all[sel])):
                _upload(dst, src)
            batch = DeviceSet(d_codes, d_off, len(uniq), max_len)
            frac = k / max(total - 1, 1)
            step_lr = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * frac)))  # cosine decay
            ep_loss += model.train_step(batch, d_pa, d_pb, d_y, d_w, len(sel), step_lr, max_len)
            k += 1
        losses.append(ep_loss / steps)
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean weighted loss {losses[-1]:.4f}")
    return losses


def _upload(dst: DeviceArray, src: np.ndarray) -> None:
    src = np.ascontiguousarray(src)
    assert src.nbytes <= dst.nbytes
    st = Status()
    check(_device.lib().st_memcpy(dst.ptr, src.ctypes.data, src.nbytes, 1, ctypes.byref(st)), st,
          "st_memcpy")


def embed_distances(model: Prefilter, codes, off, pa, pb, max_len: int) -> np.ndarray:
    """Model distances (BF16 embeddings, FP32 arithmetic) of pairs within one set, in score
    units."""
    ds = DeviceSet.upload(codes, off)
    emb, norm = model.encode(ds, max_len)
    e = _bf16_to_f32(emb.numpy())
    d = ((e[pa] - e[pb]) ** 2).sum(axis=1)
    return d * SCORE_UNIT


def _bf16_to_f32(u16: np.ndarray) -> np.ndarray:
    return (u16.astype(np.uint32) << 16).view(np.float32)


def calibrate_tau(model_dist: np.ndarray, exact: np.ndarray, cutoff: int,
                  quantile: float = DEFAULT_QUANTILE) -> float:
    """Screening threshold (score units) that keeps `quantile` of the held-out pairs whose
    exact score is <= cutoff: the cutoff itself or the quantile of their model distance,
    whichever is larger."""
    close = model_dist[exact <= cutoff]
    if len(close) == 0:
        return float(cutoff)
    return float(max(cutoff, np.quantile(close, quantile)))


def calibrate_on_set(model: Prefilter, codes: np.ndarray, off: np.ndarray, gp: GapblockParams,
                     cutoff: int, quantile: float, max_len: int):
    """Screening threshold from a calibration set of sequences the model was not trained on:
    every pair of the set with exact score <= cutoff (one exhaustive thresholded pass), and the
    `quantile` of their model distances (calibrate_tau). Returns (tau, the number of close
    pairs). Sequences unseen in training make the margin hold for new queries and references,
    which held-out pairs of training sequences understate."""
    s_all = DeviceSet.upload(codes, off)
    a, b, sc = exact_threshold(s_all, s_all, cutoff, gp)
    keep = a < b
    a, b, sc = a[keep].astype(np.int64), b[keep].astype(np.int64), sc[keep]
    dist = embed_distances(model, codes, off, a, b, max_len)
    return calibrate_tau(dist, sc, cutoff, quantile), len(sc)


@dataclass
class SearchResult:
    query: np.ndarray
    ref: np.ndarray
    score: np.ndarray
    candidates: int
    seconds: dict
    tau: float = 0.0           # the screening threshold used (score units)
    audit_queries: int = 0     # queries searched exhaustively to calibrate it
    audit_hits: int = 0        # their exact hits
    audit_within: float = 1.0  # fraction of those hits the threshold keeps


class Searcher:
    """Thresholded gap-block search through the prefilter, with device buffers kept across
    calls. ``index`` embeds the references once; ``query`` embeds a query set, screens it
    against the index, rescores the candidates exactly and returns the hits <= cutoff."""

    def __init__(self, model: Prefilter, gp: GapblockParams, max_len: int,
                 capacity: int = 1 << 24, hit_capacity: int = 1 << 21):
        self.model, self.gp, self.max_len = model, gp, max_len
        self._cand = None
        self._cap = capacity
        self._scores = None
        self._hits = None
        self._hit_cap = hit_capacity
        self.refs = None
        self.refs_built = False
        self._q_buffers = None
        self._count = ctypes.c_uint64()
        self._audit_buffers: dict = {}

    def index(self, r: DeviceSet) -> float:
        t0 = time.perf_counter()
        self.refs = r
        # the buffers of the previous index are written over when they hold r
        prev = (self.r_emb, self.r_norm) if self.refs_built else None
        self.r_emb, self.r_norm = self.model.encode(r, self.max_len, out=prev)
        self.refs_built = True
        _device.synchronize()
        return time.perf_counter() - t0

    def query(self, q: DeviceSet, cutoff: int, tau_score: float, audit: float = 0.0,
              quantile: float = DEFAULT_QUANTILE, seed: int = 0) -> SearchResult:
        """With audit > 0, a random sample of that fraction of the queries (q needs its host
        copies) is first searched exhaustively with the exact kernel; the screening threshold
        becomes the larger of tau_score and the `quantile` of the model distances of the
        sample's exact hits, so it is calibrated on this query and reference distribution
        rather than only on the training pairs'. A quantile near 1 is only as good as the
        number of hits behind it, so a sample with fewer than AUDIT_MIN_HITS hits is grown
        once, by the hit rate it saw, up to AUDIT_MAX_SHARE of the queries."""
        L = _bind()
        t = {}
        t0 = time.perf_counter()
        qe, qn = self.model.encode(q, self.max_len, out=self._q_buffers)
        self._q_buffers = (qe, qn)
        _device.synchronize()
        t["encode_queries"] = time.perf_counter() - t0
        tau = float(tau_score)
        n_audit, audit_hits, within = 0, 0, 1.0
        if audit > 0:
            t0 = time.perf_counter()
            n_audit = max(1, min(q.n, int(round(audit * q.n))))
            rng = np.random.default_rng(seed)
            # a sample without replacement drawn directly: microseconds, where a permutation
            # of all the queries took about 0.7 ms of host time at 50,000 queries
            idx = np.sort(rng.choice(q.n, n_audit, replace=False))
            aq, ar, _ = exact_threshold(q.subset(idx), self.refs, cutoff, self.gp,
                                        capacity=1 << 18, buffers=self._audit_buffers)
            pq, pr = idx[aq], ar
            cap = max(n_audit, int(AUDIT_MAX_SHARE * q.n))
            if len(aq) < AUDIT_MIN_HITS and n_audit < cap:
                # more queries not yet sampled, sized by the hit rate seen so far
                want = n_audit * AUDIT_MIN_HITS / max(len(aq), 1)
                rest = np.setdiff1d(np.arange(q.n), idx, assume_unique=True)
                k = min(cap, int(np.ceil(want))) - n_audit
                more = np.sort(rng.choice(rest, min(k, len(rest)), replace=False))
                if len(more):
                    bq, br, _ = exact_threshold(q.subset(more), self.refs, cutoff, self.gp,
                                                capacity=1 << 18, buffers=self._audit_buffers)
                    pq, pr = np.concatenate([pq, more[bq]]), np.concatenate([pr, br])
                    n_audit += len(more)
            d = pair_distances(qe, qn, self.r_emb, self.r_norm, pq, pr)
            audit_hits = len(d)
            if audit_hits:
                # a little headroom for the screen's different FP32 summation order
                tau = max(tau, float(np.quantile(d, quantile)) * 1.001)
                within = float((d <= tau).mean())
            t["audit"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        cq, cr, nc = screen(qe, qn, q.n, self.r_emb, self.r_norm, self.refs.n,
                            tau / SCORE_UNIT, self._cap, self._cand)
        self._cand, self._cap = (cq, cr), cq.shape[0]
        t["screen"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        if self._scores is None or self._scores.shape[0] < max(nc, 1):
            self._scores = DeviceArray(max(nc, self._cap), np.int32)
        scores = exact_pairs(q, self.refs, cq, cr, nc, self.gp, self._scores)
        while True:
            if self._hits is None or self._hits[0].shape[0] < self._hit_cap:
                self._hits = (DeviceArray(self._hit_cap, np.uint32),
                              DeviceArray(self._hit_cap, np.uint32),
                              DeviceArray(self._hit_cap, np.int32))
            hq, hr, hs = self._hits
            st = Status()
            check(L.st_select_pairs_le(scores.ptr, cq.ptr, cr.ptr, nc, cutoff, self._hit_cap,
                                       hq.ptr, hr.ptr, hs.ptr, ctypes.byref(self._count), None,
                                       ctypes.byref(st)), st, "st_select_pairs_le")
            k = int(self._count.value)
            if k <= self._hit_cap:
                break
            self._hit_cap = k
        t["rescore"] = time.perf_counter() - t0
        return SearchResult(hq.numpy(k), hr.numpy(k), hs.numpy(k), nc, t, tau, n_audit,
                            audit_hits, within)


def search(model: Prefilter, q: DeviceSet, r: DeviceSet, cutoff: int, tau_score: float,
           gp: GapblockParams, max_len: int, audit: float = 0.0,
           quantile: float = DEFAULT_QUANTILE) -> SearchResult:
    """One-shot search: index the references, then query (see Searcher)."""
    s = Searcher(model, gp, max_len)
    t_index = s.index(r)
    res = s.query(q, cutoff, tau_score, audit=audit, quantile=quantile)
    res.seconds["encode_refs"] = t_index
    return res


def recall(found_q, found_r, exact_q, exact_r) -> float:
    """Fraction of the exact hits (exact_q, exact_r) that the prefilter found."""
    if len(exact_q) == 0:
        return 1.0
    key = lambda a, b: (a.astype(np.int64) << 32) | b.astype(np.int64)  # noqa: E731
    return float(np.isin(key(exact_q, exact_r), key(found_q, found_r)).mean())
