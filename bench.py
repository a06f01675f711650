#!/usr/bin/env python
"""All-vs-all scoring throughput on one MI355X.

Default workload (one "step", the headline line): the 20,000 x 20,000 all-vs-all Gotoh affine
distance matrix (dist_matrix: d = s(a,a) + s(b,b) - 2 s(a,b)) of synthetic CDR3-like amino-acid
strings ('C' + body + 'F', total length uniform in 8..25, body residues i.i.d. at Robinson &
Robinson background frequencies, numpy PCG64 seed 2), BLOSUM62, gap_open 11, gap_extend 1,
computed twice -- global (Needleman-Wunsch) and local (Smith-Waterman) -- into two
device-resident int32 matrices. Inputs (codec-encoded CSR
strings) are already in HBM when the timed region starts; each step runs the full device path
(reference packing, query bucketing, scoring kernels) through the C-ABI entry point
st_align_matrix_device on one HIP stream.

--workload selects the other kernels of the path on their own configurations:
  gapblock     1,000 x 10,000 CDR3-like (lengths 10..20), BLOSUM62, gap_open 2*scale = 28,
               gap_extend 1, central prior lambda 21
  hamming      50,000 x 50,000 random ACGT UMIs of length 12
  levenshtein  20,000 x 20,000 ACGT reads of length 40..60, half derived from the other half
               by 1-3 random substitutions / insertions / deletions
  batch        64 ragged gap-block jobs (512..8,192 x 10,000..200,000, lengths 6..30,
               BLOSUM62 / structural) in one captured persistent launch
  traceback    Gotoh alignment with traceback (align_pair) of 1,000,000 candidate pairs
               drawn from two 20,000-string CDR3-like sets, BLOSUM62 11/1, global and local
  counts       per-query histograms of exact gap-block scores, 1,000 queries x 250,000
               control (the control pass of threshold_for_evalue), ceiling 64
  germline     4,000 x 4,000 Gotoh dist_matrix of germline-V-length strings (82..98),
               BLOSUM62 11/1, global and local (the multi-stripe register kernel)
  longpairs    64 x 64 Gotoh dist_matrix of 2,000-residue strings, global and local (few long
               pairs: the one-wave-per-pair wavefront kernel)
  prefilter    learned prefilter, 50,000 queries x 1,000,000 references, gap-block cutoff 48,
               against the exhaustive exact thresholded pass; a step is one search against
               reference embeddings built once (their build time and the one-shot pipeline
               are reported beside it)

Prints one JSON line: pairs/s (value), cell updates per second, the roofline of the dominant
kernel measured with HIP events on the stream it runs on, and the reference's multithreaded CPU
path timed on a bounded sample on this host.

By default every workload runs and one line carries them all: the Gotoh dist_matrix (BASELINE
C2) as the headline value, every other configuration under "configs" (value, step time,
roofline, CPU baseline), the end-to-end drop-in timings and the scalar-call latencies.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--workload all|gotoh|...]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

AA_ORDER = "ARNDCQEGHILKMFPSTWYVBZX*"
# Background amino-acid frequencies of Robinson & Robinson (PNAS 88:8880, 1991), as tabulated
# in NCBI BLAST, for the 20 standard residues in the order below.
BG_RESIDUES = "ARNDCQEGHILKMFPSTWYV"
BG_FREQ = np.array([0.07805, 0.05129, 0.04487, 0.05364, 0.01925, 0.04264, 0.06295, 0.07377,
                    0.02199, 0.05142, 0.09019, 0.05744, 0.02243, 0.03856, 0.05203, 0.07120,
                    0.05841, 0.01330, 0.03216, 0.06441])
BG_FREQ = BG_FREQ / BG_FREQ.sum()

N_QUERIES = 20_000
N_REFS = 20_000
LEN_LO, LEN_HI = 8, 25
GAP_OPEN, GAP_EXTEND = 11, 1
SEED = 2

# Integer lane-operations per DP cell of the register kernel's cell update (see
# seqtree_amd/csrc/gotoh.hip, "H form"): M = Hg_diag + s', X = max(Hg_up, X_up - ge),
# Y = max(Hg_left, Y_left - ge), Hg = max(M, X, Y) - go -> 8 operations; local adds the 0
# floor and the running best (+2). A packed int16 instruction performs two lane-operations
# (one per query half), so the count is the same for both cell widths.
K_MIN = {"global": 8.0, "local": 10.0}
# VALU lane-operation peak of MI355X: the published FP32 vector rate, 157.3 TFLOP/s, is
# 78.6e12 FMA lane-operations/s = 256 CUs x 128 lanes x 2.4 GHz. The on-device probe
# (st_probe_valu, reported in every VALU line) reaches ~87% of it with two-operand VGPR
# instructions (v_add_u32, v_max_i16); three-operand, packed and SGPR forms issue at half that
# rate, so a kernel built from them tops out near 50% of this peak.
VALU_PEAK_TOPS = 157.3 / 2


GERMLINE_N, GERMLINE_LEN = 4000, (82, 98)


def run_germline(torch, dev, lib, stream, args) -> dict:
    """Germline-V-length all-vs-all Gotoh distances (the reference benchmark's 'germline' shape,
    bench/bench_pairwise.py: lengths 90 +- 8, residues uniform over the 20 amino acids,
    BLOSUM62 11/1): 4,000 x 4,000, global and local dist_matrix. Every query is longer than
    one register stripe, so the multi-stripe register kernel carries the work."""
    rng = np.random.default_rng(SEED + 11)
    aa = np.array([AA_ORDER.index(c) for c in BG_RESIDUES], np.uint8)
    lens = rng.integers(GERMLINE_LEN[0], GERMLINE_LEN[1] + 1, size=2 * GERMLINE_N)
    seqs = [aa[rng.integers(0, 20, size=int(L))] for L in lens]
    qs, rs = seqs[:GERMLINE_N], seqs[GERMLINE_N:]
    Q, R = DeviceSet(torch, dev, qs), DeviceSet(torch, dev, rs)
    sim, _, _ = matrix_tables("blosum62")
    outs = {mode: torch.empty((Q.n, R.n), dtype=torch.int32, device=dev)
            for mode in ("global", "local")}
    st = Status()

    def step():
        for mode in ("global", "local"):
            rc = lib.lib.st_align_matrix_device(
                Q.bytes.data_ptr(), Q.off.data_ptr(), Q.n, Q.max_len, R.bytes.data_ptr(),
                R.off.data_ptr(), R.n, R.max_len, sim.ctypes.data, len(AA_ORDER),
                int(mode == "local"), 1, GAP_OPEN, GAP_EXTEND, outs[mode].data_ptr(),
                ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
            check(rc, st, "st_align_matrix_device")

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    pairs_per_step = 2 * Q.n * R.n
    cells_per_mode = float(Q.lens.sum()) * float(R.lens.sum())
    label, launches, avg_s = dominant(kernels)
    roofline = None
    if "_stripes_" in label:
        mode = label.rsplit("_", 1)[1]
        roofline = valu_roofline(label, launches, avg_s, cells_per_mode * K_MIN[mode],
                                 cells_per_launch=cells_per_mode, ops_per_cell=K_MIN[mode])
    cpu = None
    if args.cpu_rows > 0:
        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        cpu = cpu_reference("gotoh", [to_str(s) for s in qs], [to_str(s) for s in rs],
                            max(1, args.cpu_rows // 10), seconds=args.cpu_seconds, sim=sim)
    line = base_line(
        "all-vs-all pairs/s (Gotoh NW+SW dist_matrix, germline-V lengths)",
        pairs_per_step / (ms_step / 1e3), args.steps, args.warmup, ms_step,
        {"workload": "gotoh_dist_matrix_germline_4000x4000_len82-98_blosum62_go11_ge1_"
                     "global+local",
         "queries": Q.n, "refs": R.n, "lengths": list(GERMLINE_LEN), "seed": SEED + 11,
         "pairs_per_step": pairs_per_step, "cells_per_step": 2 * cells_per_mode,
         "parallelism": "single GPU"},
        kernels, roofline, cpu, gcups=2 * cells_per_mode / (ms_step / 1e3) / 1e9)
    line["dtype"] = "int16"
    line["arithmetic"] = GOTOH_ARITHMETIC
    return line


LONG_N, LONG_LEN = 64, 2000


def run_longpairs(torch, dev, lib, stream, args) -> dict:
    """Few long pairs: 64 x 64 all-vs-all Gotoh distances of 2,000-residue strings (residues
    uniform over the 20 amino acids, BLOSUM62 11/1), global and local. 4,096 pairs would give
    a lane-per-pair kernel 64 waves; the intra-pair wavefront kernel runs a wave per pair."""
    rng = np.random.default_rng(SEED + 13)
    aa = np.array([AA_ORDER.index(c) for c in BG_RESIDUES], np.uint8)
    seqs = [aa[rng.integers(0, 20, size=LONG_LEN)] for _ in range(2 * LONG_N)]
    qs, rs = seqs[:LONG_N], seqs[LONG_N:]
    Q, R = DeviceSet(torch, dev, qs), DeviceSet(torch, dev, rs)
    sim, _, _ = matrix_tables("blosum62")
    outs = {mode: torch.empty((Q.n, R.n), dtype=torch.int32, device=dev)
            for mode in ("global", "local")}
    st = Status()

    def step():
        for mode in ("global", "local"):
            rc = lib.lib.st_align_matrix_device(
                Q.bytes.data_ptr(), Q.off.data_ptr(), Q.n, Q.max_len, R.bytes.data_ptr(),
                R.off.data_ptr(), R.n, R.max_len, sim.ctypes.data, len(AA_ORDER),
                int(mode == "local"), 1, GAP_OPEN, GAP_EXTEND, outs[mode].data_ptr(),
                ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
            check(rc, st, "st_align_matrix_device")

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    pairs_per_step = 2 * Q.n * R.n
    cells_per_mode = float(Q.lens.sum()) * float(R.lens.sum())
    label, launches, avg_s = dominant(kernels)
    roofline = None
    if label.startswith("gotoh_wave_") and "self" not in label:
        mode = label.rsplit("_", 1)[1]
        roofline = valu_roofline(label, launches, avg_s, cells_per_mode * K_MIN[mode],
                                 cells_per_launch=cells_per_mode, ops_per_cell=K_MIN[mode])
    cpu = None
    if args.cpu_rows > 0:
        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        cpu = cpu_reference("gotoh", [to_str(x) for x in qs], [to_str(x) for x in rs], 2,
                            seconds=args.cpu_seconds, sim=sim)
    line = base_line(
        "all-vs-all pairs/s (Gotoh NW+SW dist_matrix, few long pairs)",
        pairs_per_step / (ms_step / 1e3), args.steps, args.warmup, ms_step,
        {"workload": "gotoh_dist_matrix_long_64x64_len2000_blosum62_go11_ge1_global+local",
         "queries": Q.n, "refs": R.n, "length": LONG_LEN, "seed": SEED + 13,
         "pairs_per_step": pairs_per_step, "cells_per_step": 2 * cells_per_mode,
         "parallelism": "single GPU"},
        kernels, roofline, cpu, gcups=2 * cells_per_mode / (ms_step / 1e3) / 1e9)
    line["dtype"] = "int32"
    line["arithmetic"] = ("int32 Gotoh cells, one wave per pair with the query's rows spread "
                          "over the lanes (range-proven)")
    return line


GOTOH_ARITHMETIC = ("packed int16x2 Gotoh cells (two queries per register; exact under the "
                    "host's range proof, int32 / int64 kernels past it)")


def end_to_end(call, pairs: float, what: str, cpu) -> dict:
    """The drop-in call as a user makes it (Python strings in, host ScoreMatrix out): one
    warm-up, then the best of three, timed with perf_counter as the reference's own
    benchmarks do (bench/bench_score_matrix.py:35-37, 70-77)."""
    call()
    best = min(_timed(call) for _ in range(3))
    out = {"pairs_per_s": pairs / best, "ms_per_call": round(1e3 * best, 3), "what": what}
    if cpu:
        out["vs_cpu_baseline"] = round(pairs / best / cpu["value"], 2)
    return out


def _timed(call) -> float:
    t0 = time.perf_counter()
    call()
    return time.perf_counter() - t0


def scalar_latency(args) -> dict:
    """Latency of the scalar entry points (one pair per call), median of 200 calls, against the
    reference's own C++ scalar functions (oracle/_ref, through ctypes) on the host."""
    import oracle
    import seqtree_amd as st

    m = st.SubstitutionMatrix.blosum62()
    q, r = "CASSLGQAYEQYF", "CASSPGQAYEQF"
    idx = st.Index.build([r], "aa")
    params = st.SearchParams(matrix="BLOSUM62", gap_open=11, gap_extend=1)
    umi_a, umi_b = "ACGTACGTACGT", "ACGTTCGTACGA"
    read_a, read_b = "ACGT" * 12, "ACGA" * 12

    def med(fn, n=200):
        fn()
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return round(1e6 * float(np.median(ts)), 1)

    out = {"unit": "us per call (median of 200)",
           "gpu": {"pairwise.score": med(lambda: st.pairwise.score(q, r, m)),
                   "pairwise.align": med(lambda: st.pairwise.align(q, r, m)),
                   "Index.align": med(lambda: idx.align(0, q, params)),
                   "distance.hamming": med(lambda: st.distance.hamming(umi_a, umi_b)),
                   "distance.levenshtein": med(lambda: st.distance.levenshtein(read_a, read_b))}}
    if oracle.reference_available():
        R = oracle.Reference()
        out["reference_cpu"] = {
            "align_score": med(lambda: R.align_score(q, r, "blosum62")),
            "align_pair": med(lambda: R.align_pair(q, r, "blosum62")),
            "Index.align": med(lambda: R.index_align_pairs([q], [r], [0], [0], "blosum62", 11,
                                                           1)),
            "hamming": med(lambda: R.distance_matrix("hamming", [umi_a], [umi_b])),
            "levenshtein": med(lambda: R.distance_matrix("levenshtein", [read_a], [read_b]))}
        out["reference_note"] = ("through ctypes (a few us of call overhead); Index.align "
                                 "includes building a one-reference Index")
    return out


def cdr3_set(rng: np.random.Generator, n: int, lo: int, hi: int) -> list[np.ndarray]:
    """Codec codes of n strings 'C' + body + 'F' with total length uniform in [lo, hi]."""
    lens = rng.integers(lo, hi + 1, size=n)
    codes_bg = np.array([AA_ORDER.index(c) for c in BG_RESIDUES], np.uint8)
    c_code, f_code = AA_ORDER.index("C"), AA_ORDER.index("F")
    out = []
    for L in lens:
        body = codes_bg[rng.choice(20, size=L - 2, p=BG_FREQ)]
        out.append(np.concatenate([[c_code], body, [f_code]]).astype(np.uint8))
    return out


def cdr3_csr(rng: np.random.Generator, n: int, lo: int, hi: int):
    """cdr3_set's strings generated in bulk, as CSR (codes, offsets, lengths)."""
    lens = rng.integers(lo, hi + 1, size=n).astype(np.int64)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    codes_bg = np.array([AA_ORDER.index(c) for c in BG_RESIDUES], np.uint8)
    codes = np.empty(int(off[-1]), np.uint8)
    body = np.ones(len(codes), bool)
    starts = off[:-1].astype(np.int64)
    body[starts] = False
    body[starts + lens - 1] = False
    codes[body] = codes_bg[rng.choice(20, size=int(body.sum()), p=BG_FREQ)]
    codes[starts] = AA_ORDER.index("C")
    codes[starts + lens - 1] = AA_ORDER.index("F")
    return codes, off, lens


def dna_set(rng: np.random.Generator, n: int, lo: int, hi: int) -> list[np.ndarray]:
    """Raw ACGT byte strings with length uniform in [lo, hi]."""
    acgt = np.frombuffer(b"ACGT", np.uint8)
    return [acgt[rng.integers(0, 4, size=int(L))] for L in rng.integers(lo, hi + 1, size=n)]


def mutate(rng: np.random.Generator, s: np.ndarray) -> np.ndarray:
    """1-3 random substitutions, insertions or deletions."""
    acgt = np.frombuffer(b"ACGT", np.uint8)
    t = list(s)
    for _ in range(int(rng.integers(1, 4))):
        op, p = int(rng.integers(0, 3)), int(rng.integers(0, len(t)))
        if op == 0:
            t[p] = acgt[rng.integers(0, 4)]
        elif op == 1:
            t.insert(p, acgt[rng.integers(0, 4)])
        elif len(t) > 1:
            del t[p]
    return np.array(t, np.uint8)


def csr(seqs: list[np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return np.concatenate(seqs).astype(np.uint8), off


def matrix_tables(name: str) -> tuple[np.ndarray, np.ndarray, int]:
    import seqtree_amd as st

    m = getattr(st.SubstitutionMatrix, name)()
    sim = np.array([[m.similarity(a, b) for b in AA_ORDER] for a in AA_ORDER], np.int32)
    pen = np.array([[m.penalty(a, b) for b in AA_ORDER] for a in AA_ORDER], np.int32)
    return sim, pen, m.scale()


class Lib:
    def __init__(self):
        self.lib = ctypes.CDLL(str(ROOT / "seqtree_amd" / "libseqtree_hip.so"))
        L = self.lib
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
        L.st_align_matrix_device.restype = i32
        L.st_align_matrix_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, u32, i32,
                                             i32, i32, i32, vp, vp, vp]
        L.st_gapblock_matrix_device.restype = i32
        L.st_gapblock_matrix_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, u32, i32,
                                                i32, vp, u32, vp, vp, vp]
        for name in ("st_hamming_matrix_device", "st_levenshtein_matrix_device"):
            fn = getattr(L, name)
            fn.restype = i32
            fn.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, vp, vp]
        L.st_align_pairs_device.restype = i32
        L.st_align_pairs_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, vp, u64, vp,
                                            u32, i32, i32, i32, vp, u64, vp, vp, vp, vp, vp, vp]
        L.st_gapblock_histogram_device.restype = i32
        L.st_gapblock_histogram_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, u32,
                                                   i32, i32, vp, u32, u32, vp, vp, vp]
        L.st_probe_mfma.restype = i32
        L.st_probe_mfma.argtypes = [ctypes.POINTER(ctypes.c_double), vp]
        L.st_probe_hbm_write.restype = i32
        L.st_probe_hbm_write.argtypes = [u64, ctypes.POINTER(ctypes.c_double), vp]
        L.st_probe_lds.restype = i32
        L.st_probe_lds.argtypes = [i32, ctypes.POINTER(ctypes.c_double), vp]
        L.st_probe_valu.restype = i32
        L.st_probe_valu.argtypes = [i32, ctypes.c_char_p, u64, ctypes.POINTER(ctypes.c_double),
                                    vp]
        L.st_timing_enable.argtypes = [i32]
        L.st_timing_collect.restype = i32
        L.st_timing_collect.argtypes = [ctypes.c_char_p, u64]

    def timing(self) -> dict[str, tuple[int, float]]:
        buf = ctypes.create_string_buffer(1 << 16)
        n = self.lib.st_timing_collect(buf, len(buf))
        if n < 0:
            raise RuntimeError("unresolved timing events")
        out = {}
        for line in buf.value.decode().splitlines():
            label, count, ms = line.split()
            out[label] = (int(count), float(ms))
        return out


class Status(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("message", ctypes.c_char * 256)]


def check(rc: int, st: Status, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {rc} {st.message.decode()}")


class DeviceSet:
    """A CSR string set resident in HBM (torch tensors only hold the memory)."""

    def __init__(self, torch, dev, seqs: list[np.ndarray]):
        b, o = csr(seqs)
        self.n = len(seqs)
        self.lens = np.diff(o.astype(np.int64))
        self.max_len = int(self.lens.max())
        self.bytes = torch.from_numpy(b).to(dev)
        self.off = torch.from_numpy(o.view(np.int64)).to(dev)


def time_steps(torch, lib: Lib, stream, step, steps: int, warmup: int):
    """Untimed warmup, then `steps` steps bracketed by HIP events on `stream` with the
    library's launch timing off (the step time: no per-launch events on the host path), then
    `steps` more with it on, whose per-kernel launch times come from the library's own events
    on the same stream."""
    for _ in range(warmup):
        step()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    lib.lib.st_timing_enable(1)
    lib.timing()  # drop anything recorded before
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    lib.lib.st_timing_enable(0)
    return ms, lib.timing()


def traffic_for(label: str):
    """HBM bytes per dispatch measured with rocprofv3 --pmc (profiles/traffic.json), or None."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    return json.loads(path.read_text()).get(label)


def dominant(kernels):
    label, (launches, ms) = max(kernels.items(), key=lambda kv: kv[1][1])
    return label, launches, ms / launches / 1e3


def host_cpu() -> dict:
    """The host the CPU baseline runs on: CPU model, logical CPUs and those this process may
    use."""
    import os

    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


# The north-star comparison is the reference's CPU path on 16 cores; it is also timed on every
# core this process may use (the reference's threads=0).
CPU_THREADS = 16


def timed_rate(once, seconds: float):
    """One untimed warm-up call, then calls of `once` (returning its work units) repeated until
    `seconds` of wall time: (units per second, calls, units, seconds)."""
    once()
    units, calls = 0, 0
    t0 = time.perf_counter()
    while True:
        units += once()
        calls += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            return units / dt, calls, units, dt


def cpu_reference(kind: str, q, r, rows: int, seconds: float = 3.0, **kw) -> dict:
    """The reference's own C++ (oracle/_ref, compiled -O3 from its sources) on the first `rows`
    queries against every reference, with 16 worker threads and with every usable core; the
    plain-C port where the compiled reference is absent."""
    import oracle

    host = host_cpu()
    qs = q[:rows]
    have_ref = oracle.reference_available()
    impl = oracle.Reference() if have_ref else oracle.Oracle()

    def once(threads) -> int:
        if kind == "gotoh":
            for local in (False, True):
                impl.align_matrix(qs, r, "blosum62" if have_ref else kw["sim"], local, True,
                                  GAP_OPEN, GAP_EXTEND, threads=threads)
            return 2 * len(qs) * len(r)
        if kind == "gapblock":
            impl.gapblock_matrix(qs, r, "blosum62" if have_ref else kw["pen"], kw["go"], 1,
                                 kw["cube"], kw["width"], threads=threads)
            return len(qs) * len(r)
        impl.distance_matrix(kind, qs, r, threads=threads)
        return len(qs) * len(r)

    rate, calls, pairs, dt = timed_rate(lambda: once(CPU_THREADS), seconds)
    out = {"value": rate, "unit": "pairs/s", "cores": CPU_THREADS,
           "kind": "reference" if have_ref else "port",
           "sample": f"first {len(qs)} queries x all {len(r)} references of the same workload"
                     f"{' (global + local)' if kind == 'gotoh' else ''}, {calls} calls after "
                     f"one warm-up: {pairs} pairs in {dt:.2f} s wall",
           "host": host}
    if host["usable_cpus"] != CPU_THREADS:
        all_rate, _, _, _ = timed_rate(lambda: once(host["usable_cpus"]), seconds)
        out["all_cores"] = {"value": all_rate, "cores": host["usable_cpus"]}
    return out


def base_line(metric, value, steps, warmup, ms_step, config, kernels, roofline, cpu, **extra):
    line = {"metric": metric, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms_step, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic"}
    line.update(extra)
    line.update({"config": config,
                 "kernels_ms": {k: round(v[1] / steps, 4) for k, v in sorted(kernels.items())},
                 "roofline": roofline, "cpu_baseline": cpu})
    return line


def valu_probe(lib) -> dict:
    """Measured issue rate (T lane-instructions/s) of each probed VALU instruction kind."""
    out = {}
    for kind in range(16):
        name = ctypes.create_string_buffer(32)
        rate = ctypes.c_double()
        st = Status()
        if lib.lib.st_probe_valu(kind, name, 32, ctypes.byref(rate), ctypes.byref(st)) != 0:
            break
        out[name.value.decode()] = round(rate.value / 1e12, 2)
    return out


def valu_roofline(label, launches, avg_s, ops_per_launch, **info):
    achieved = ops_per_launch / avg_s / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "Tops/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "traffic": traffic_for(label), "kernel": label, "launches": launches, "avg_launch_ms": round(avg_s * 1e3, 4),
            **info}


# ------------------------------------------------------------------------------ workloads

def run_gotoh(torch, dev, lib, stream, args) -> dict:
    rng = np.random.default_rng(SEED)
    qs = cdr3_set(rng, N_QUERIES, LEN_LO, LEN_HI)
    rs = cdr3_set(rng, N_REFS, LEN_LO, LEN_HI)
    Q, R = DeviceSet(torch, dev, qs), DeviceSet(torch, dev, rs)
    sim, _, _ = matrix_tables("blosum62")
    outs = {mode: torch.empty((Q.n, R.n), dtype=torch.int32, device=dev)
            for mode in ("global", "local")}
    st = Status()

    def step():
        for mode in ("global", "local"):
            rc = lib.lib.st_align_matrix_device(
                Q.bytes.data_ptr(), Q.off.data_ptr(), Q.n, Q.max_len, R.bytes.data_ptr(),
                R.off.data_ptr(), R.n, R.max_len, sim.ctypes.data, len(AA_ORDER),
                int(mode == "local"), 1, GAP_OPEN, GAP_EXTEND, outs[mode].data_ptr(),
                ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
            check(rc, st, "st_align_matrix_device")

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    pairs_per_step = 2 * Q.n * R.n
    cells_per_mode = float(Q.lens.sum()) * float(R.lens.sum())
    # dominant kernel: its algorithmic work per launch is sum(m) * sum(n) over the queries of
    # its length class, times the per-cell instruction count of its cell update
    label, launches, avg_s = dominant(kernels)
    roofline = None
    if label.startswith("gotoh_i"):
        _, cell, rtag, mode = label.split("_")
        rows = int(rtag[1:])
        in_class = (Q.lens <= rows) & (Q.lens > rows - 4)
        cells = float(Q.lens[in_class].sum()) * float(R.lens.sum())
        roofline = valu_roofline(label, launches, avg_s, cells * K_MIN[mode],
                                 cells_per_launch=cells, ops_per_cell=K_MIN[mode])
    cpu = None
    if args.cpu_rows > 0:
        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        cpu = cpu_reference("gotoh", [to_str(s) for s in qs], [to_str(s) for s in rs],
                            args.cpu_rows, seconds=args.cpu_seconds, sim=sim)
    e2e = None
    if args.end_to_end:
        import seqtree_amd as st

        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        q_strs, r_strs = [to_str(x) for x in qs], [to_str(x) for x in rs]
        m = st.SubstitutionMatrix.blosum62()
        e2e = end_to_end(lambda: [st.pairwise.dist_matrix(q_strs, r_strs, m, mode=md)
                                  for md in ("global", "local")], pairs_per_step,
                         "seqtree_amd.pairwise.dist_matrix, global then local: str lists in, "
                         "host ScoreMatrix out", cpu)
    line = base_line(
        "all-vs-all pairs/s (Gotoh NW+SW dist_matrix)", pairs_per_step / (ms_step / 1e3),
        args.steps, args.warmup, ms_step,
        {"workload": "gotoh_dist_matrix_cdr3_20000x20000_blosum62_go11_ge1_global+local",
         "queries": Q.n, "refs": R.n, "lengths": [LEN_LO, LEN_HI], "seed": SEED,
         "pairs_per_step": pairs_per_step, "cells_per_step": 2 * cells_per_mode,
         "parallelism": "single GPU"},
        kernels, roofline, cpu, gcups=2 * cells_per_mode / (ms_step / 1e3) / 1e9,
        end_to_end=e2e)
    line["dtype"] = "int16"
    line["arithmetic"] = GOTOH_ARITHMETIC
    return line


def run_gapblock(torch, dev, lib, stream, args) -> dict:
    from seqtree_amd.gapblock import _prior_cube, central_prior

    rng = np.random.default_rng(SEED + 1)
    qs = cdr3_set(rng, 1000, 10, 20)
    rs = cdr3_set(rng, 10000, 10, 20)
    Q, R = DeviceSet(torch, dev, qs), DeviceSet(torch, dev, rs)
    _, pen, scale = matrix_tables("blosum62")
    go, lam = 2 * scale, int(1.5 * scale)
    width = max(Q.max_len, R.max_len)
    cube = np.array(_prior_cube(central_prior(lam), width), np.int32)
    out = torch.empty((Q.n, R.n), dtype=torch.int32, device=dev)
    st = Status()

    def step():
        rc = lib.lib.st_gapblock_matrix_device(
            Q.bytes.data_ptr(), Q.off.data_ptr(), Q.n, Q.max_len, R.bytes.data_ptr(),
            R.off.data_ptr(), R.n, R.max_len, pen.ctypes.data, len(AA_ORDER), go, 1,
            cube.ctypes.data, width, out.data_ptr(), ctypes.c_void_p(stream.cuda_stream),
            ctypes.byref(st))
        check(rc, st, "st_gapblock_matrix_device")

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    pairs = Q.n * R.n
    # algorithmic lookups: L per pair at d == 0, 2L at d > 0; the one-pass
    # minimum spends 5 integer ops per block position when d > 0 (two prefix sums, their
    # difference, the prior, the min) and 1 at d == 0
    L = np.minimum(Q.lens[:, None], R.lens[None, :])
    d = np.abs(Q.lens[:, None] - R.lens[None, :])
    lookups = float(np.where(d > 0, 2 * L, L).sum())
    ops = float(np.where(d > 0, 5 * L, L).sum())
    label, launches, avg_s = dominant(kernels)
    roofline = valu_roofline(label, launches, avg_s, ops, lookups_per_launch=lookups,
                             ops_per_position=5,
                             lds_bytes_per_lookup=2 if "twin" in label else 4
                             ) if label.startswith("gapblock") else None
    cpu = None
    if args.cpu_rows > 0:
        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        cpu = cpu_reference("gapblock", [to_str(s) for s in qs], [to_str(s) for s in rs],
                            min(args.cpu_rows, Q.n), seconds=args.cpu_seconds, pen=pen, go=go,
                            cube=cube.tolist(), width=width)
    e2e = None
    if args.end_to_end:
        import seqtree_amd as st

        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        q_strs, r_strs = [to_str(x) for x in qs], [to_str(x) for x in rs]
        m = st.SubstitutionMatrix.blosum62()
        prior = central_prior(lam)
        e2e = end_to_end(lambda: st.score_matrix(q_strs, r_strs, m, go, 1, prior), pairs,
                         "seqtree_amd.score_matrix: str lists in, prior cube built, host "
                         "ScoreMatrix out", cpu)
    return base_line(
        "all-vs-all pairs/s (gap-block score_matrix)", pairs / (ms_step / 1e3), args.steps,
        args.warmup, ms_step,
        {"workload": "gapblock_score_matrix_cdr3_1000x10000_blosum62_go28_ge1_central21",
         "queries": Q.n, "refs": R.n, "lengths": [10, 20], "seed": SEED + 1,
         "pairs_per_step": pairs, "lookups_per_step": lookups, "parallelism": "single GPU"},
        kernels, roofline, cpu, glookups=lookups / (ms_step / 1e3) / 1e9, end_to_end=e2e)


# Ragged batch: 64 jobs, log-normal set sizes (median 2,000 queries and 50,000 references,
# sigma 0.8) clipped to [512, 8192] and [10^4, 2*10^5], lengths 6-30, matrices alternating
# BLOSUM62 (gap_open 2*14, central prior 21) and structural (2*7, 10).
BATCH_JOBS = 64
BATCH_N = (np.log(2000.0), 0.8, 512, 8192)
BATCH_K = (np.log(50000.0), 0.8, 10_000, 200_000)
BATCH_LEN = (6, 30)


def batch_jobs(seed: int):
    import seqtree_amd as st
    from seqtree_amd.batch import EncodedSet, GapblockJob

    rng = np.random.default_rng(seed)
    mats = [st.SubstitutionMatrix.blosum62(), st.SubstitutionMatrix.structural()]
    jobs, hists = [], []
    for t in range(BATCH_JOBS):
        n = int(np.clip(round(rng.lognormal(BATCH_N[0], BATCH_N[1])), BATCH_N[2], BATCH_N[3]))
        k = int(np.clip(round(rng.lognormal(BATCH_K[0], BATCH_K[1])), BATCH_K[2], BATCH_K[3]))
        qc, qo, ql = cdr3_csr(rng, n, *BATCH_LEN)
        rc, ro, rl = cdr3_csr(rng, k, *BATCH_LEN)
        m = mats[t % 2]
        jobs.append(GapblockJob(EncodedSet(qc, qo), EncodedSet(rc, ro), m, 2 * m.scale(), 1,
                                st.central_prior(int(1.5 * m.scale()))))
        hists.append((np.bincount(ql, minlength=BATCH_LEN[1] + 1).astype(np.float64),
                      np.bincount(rl, minlength=BATCH_LEN[1] + 1).astype(np.float64)))
    return jobs, hists


def run_batch(torch, dev, lib, stream, args) -> dict:
    from seqtree_amd.batch import GapblockBatch

    seed = SEED + 3
    jobs, hists = batch_jobs(seed)
    pairs = float(sum(len(j.queries) * len(j.refs) for j in jobs))
    # algorithmic work from the length histograms: per pair L lookups at d == 0, 2L at d > 0,
    # and 5 integer operations per block position at d > 0 (1 at d == 0), as for one job
    ls = np.arange(BATCH_LEN[1] + 1, dtype=np.float64)
    Lmin = np.minimum(ls[:, None], ls[None, :])
    dpos = ls[:, None] != ls[None, :]
    look_w = np.where(dpos, 2 * Lmin, Lmin)
    ops_w = np.where(dpos, 5 * Lmin, Lmin)
    lookups = float(sum(hq @ look_w @ hr for hq, hr in hists))
    ops = float(sum(hq @ ops_w @ hr for hq, hr in hists))
    t0 = time.perf_counter()
    batch = GapblockBatch(jobs, autotune=True, stream=stream.cuda_stream)
    plan_s = time.perf_counter() - t0
    tiles = batch.tiles()

    def step():
        batch.run()

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    label, launches, avg_s = dominant(kernels)
    all_twin = batch.twin_jobs() == len(jobs)
    roofline = valu_roofline(label, launches, avg_s, ops, lookups_per_launch=lookups,
                             ops_per_position=5, lds_bytes_per_lookup=2 if all_twin else 4)
    cpu = None
    if args.cpu_rows > 0:
        # the reference has no batch call: its gap-block score_matrix on rows of job 0
        _, pen, _ = matrix_tables("blosum62")
        j0 = jobs[0]
        width = int(max(j0.queries.lengths().max(), j0.refs.lengths().max()))
        from seqtree_amd.gapblock import _prior_cube
        to_strs = lambda e: ["".join(AA_ORDER[c] for c in e.codes[int(e.offsets[i]):int(e.offsets[i + 1])])  # noqa: E731,E501
                             for i in range(len(e))]
        cpu = cpu_reference("gapblock", to_strs(j0.queries), to_strs(j0.refs),
                            min(args.cpu_rows, 200), seconds=args.cpu_seconds, pen=pen,
                            go=j0.gap_open,
                            cube=_prior_cube(j0.gap_prior, width), width=width)
    batch.close()
    return base_line(
        "all-vs-all pairs/s (ragged gap-block batch)", pairs / (ms_step / 1e3), args.steps,
        args.warmup, ms_step,
        {"workload": "gapblock_batch_64jobs_lognormal_cdr3_6-30_blosum62+structural",
         "jobs": BATCH_JOBS, "seed": seed, "pairs_per_step": pairs,
         "queries_total": int(sum(len(j.queries) for j in jobs)),
         "refs_total": int(sum(len(j.refs) for j in jobs)),
         "lookups_per_step": lookups, "tiles": tiles, "plan_and_autotune_s": round(plan_s, 2),
         "parallelism": "single GPU"},
        kernels, roofline, cpu, glookups=lookups / (ms_step / 1e3) / 1e9)


# Learned prefilter (BASELINE C5): 200,000 synthetic CDR3s (independent draws and clonal
# families of 4); 180,000 train on 2M pairs mined and labelled by the exact gap-block kernel
# (BLOSUM62, gap_open 28, central prior 21), 20,000 calibrate the screen's margin; then search 50,000 queries x 1,000,000
# references at a fixed score cutoff, on two reference/query sets: independent synthetic
# CDR3s (the generator alone; the headline value) and planted families (250,000 families of 4
# references, each query a fresh mutant of one family's root).
PF_LEN = (10, 20)
PF_TRAIN_IID, PF_TRAIN_FAM = 90_000, (22_500, 4)   # 180,000 training sequences
PF_CALIB_IID, PF_CALIB_FAM = 10_000, (2_500, 4)    # and 20,000 held out for the margin
PF_PAIRS, PF_NEAR = 2_000_000, 0.4
PF_REFS, PF_FAMILIES = 1_000_000, (250_000, 4)
PF_QUERIES = 50_000
PF_CUTOFF = 48
PF_EPOCHS = 16
PF_QUANTILE = 0.997  # seqtree_amd.prefilter.DEFAULT_QUANTILE
PF_AUDIT = 0.01  # queries searched exhaustively per search to calibrate the threshold


def pf_sets(rng, syn):
    ic, io, _ = syn.cdr3_csr(rng, PF_REFS, *PF_LEN)
    jc, jo, _ = syn.cdr3_csr(rng, PF_QUERIES, *PF_LEN)
    roots: list = []
    rc, ro, _ = syn.families(rng, *PF_FAMILIES, *PF_LEN, roots_out=roots)
    pick = rng.choice(PF_FAMILIES[0], size=PF_QUERIES, replace=False)
    qc, qo = syn.join(syn._mutate_many(rng, [roots[f] for f in pick], 1, *PF_LEN))
    return {"iid": (jc, jo, ic, io), "planted": (qc, qo, rc, ro)}


def pf_search_set(torch, lib, pf, model, gp, tau, sets, args):
    """One query / reference set: the exhaustive exact pass (best of two), then `steps`
    searches through the prefilter, each step the reference index build and one search, timed
    apart; recall against the exhaustive pass."""
    qc, qo, rc, ro = sets
    Q, R = pf.DeviceSet.upload(qc, qo), pf.DeviceSet.upload(rc, ro)
    ex_times = []
    for _ in range(2):
        t0 = time.perf_counter()
        eq, er, es = pf.exact_threshold(Q, R, args.pf_cutoff, gp)
        ex_times.append(time.perf_counter() - t0)
    ex_s = min(ex_times)
    searcher = pf.Searcher(model, gp, PF_LEN[1])
    for _ in range(args.warmup):
        searcher.index(R)
        res = searcher.query(Q, args.pf_cutoff, tau, audit=PF_AUDIT,
                             quantile=args.pf_quantile)
    torch.cuda.synchronize()
    build_times, search_times = [], []
    for step in range(args.steps):
        t0 = time.perf_counter()
        searcher.index(R)
        t1 = time.perf_counter()
        res = searcher.query(Q, args.pf_cutoff, tau, audit=PF_AUDIT, quantile=args.pf_quantile,
                             seed=step)
        search_times.append(time.perf_counter() - t1)
        build_times.append(t1 - t0)
    return {"Q": Q, "R": R, "exact": (eq, er, es), "ex_s": ex_s, "res": res,
            "ms_build": 1e3 * float(np.mean(build_times)),
            "ms_search": 1e3 * float(np.mean(search_times)), "searcher": searcher}


def run_prefilter(torch, dev, lib, stream, args) -> dict:
    from seqtree_amd import prefilter as pf
    from seqtree_amd import synthetic as syn

    seed = SEED + 4
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    tc, to = syn.training_set(rng, PF_TRAIN_IID, *PF_TRAIN_FAM, *PF_LEN)
    hc, ho = syn.training_set(rng, PF_CALIB_IID, *PF_CALIB_FAM, *PF_LEN)
    gp = pf.GapblockParams.blosum62_central(PF_LEN[1])
    pa, pb, n_close = pf.mine_pairs(rng, tc, to, gp, PF_PAIRS, PF_NEAR, args.pf_cutoff,
                                    2 * args.pf_cutoff)
    score = pf.label_pairs(tc, to, pa, pb, gp)
    sets = pf_sets(rng, syn)
    data_s = time.perf_counter() - t0

    model = pf.Prefilter(seed=seed)
    t0 = time.perf_counter()
    losses = pf.train(model, pf.TrainingData(tc, to, pa, pb, score), epochs=args.pf_epochs,
                      seed=seed, max_len=PF_LEN[1],
                      log=lambda m: print(m, file=sys.stderr, flush=True))
    train_s = time.perf_counter() - t0
    tau, n_cal = pf.calibrate_on_set(model, hc, ho, gp, args.pf_cutoff, args.pf_quantile,
                                      PF_LEN[1])
    calib = {"calibration_sequences": len(ho) - 1, "calibration_close_pairs": int(n_cal),
             "quantile": args.pf_quantile, "tau_calibrated": round(tau, 3)}

    per_set = {}
    for name in ("iid", "planted"):
        r = pf_search_set(torch, lib, pf, model, gp, tau, sets[name], args)
        eq, er, _ = r["exact"]
        res = r["res"]
        pairs = float(r["Q"].n) * float(r["R"].n)
        per_set[name] = {
            "pairs_per_step": pairs, "queries": r["Q"].n, "refs": r["R"].n,
            "effective_pairs_per_s": pairs / (r["ms_search"] / 1e3),
            "ms_per_search": round(r["ms_search"], 3), "index_build_ms": round(r["ms_build"], 3),
            "one_shot_ms": round(r["ms_build"] + r["ms_search"], 3),
            "exhaustive_ms": round(1e3 * r["ex_s"], 3),
            "exhaustive_pairs_per_s": pairs / r["ex_s"],
            "speedup_search_only": r["ex_s"] / (r["ms_search"] / 1e3),
            "speedup_one_shot": r["ex_s"] / ((r["ms_build"] + r["ms_search"]) / 1e3),
            "recall": pf.recall(res.query, res.ref, eq, er), "exact_hits": int(len(eq)),
            "found_hits": int(len(res.query)), "candidates": res.candidates,
            "tau_score": round(res.tau, 3), "audit_queries": res.audit_queries,
            "audit_hits": res.audit_hits,
            "stage_ms": {k: round(1e3 * v, 3) for k, v in res.seconds.items()}}
        if name == "iid":
            main_set = r
    # per-kernel breakdown of the headline set's steps, with the library's launch timing on
    searcher, Q, R = main_set["searcher"], main_set["Q"], main_set["R"]
    lib.lib.st_timing_enable(1)
    lib.timing()
    for step in range(args.steps):
        searcher.index(R)
        searcher.query(Q, args.pf_cutoff, tau, audit=PF_AUDIT, quantile=args.pf_quantile,
                       seed=step)
    torch.cuda.synchronize()
    lib.lib.st_timing_enable(0)
    kernels = lib.timing()
    head = per_set["iid"]
    pairs = head["pairs_per_step"]
    label, launches, avg_s = dominant(kernels)
    if label == "prefilter_screen":
        flops = 2.0 * 128 * pairs
    else:  # encoder: three convolution GEMMs over one set's rows per launch
        slot = PF_LEN[1] + 4
        rows = (Q.n + R.n) * slot
        flops = 2.0 * rows * 128 * (5 * 32 + 2 * 5 * 128) / max(launches / args.steps, 1)
    achieved = flops / avg_s / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": 2500.0,
                "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4),
                "traffic": traffic_for(label), "kernel": label, "launches": launches,
                "avg_launch_ms": round(avg_s * 1e3, 4), "flops_per_launch": flops}
    cpu = None
    if args.cpu_rows > 0:
        qc, qo, rc, ro = sets["iid"]
        qs, rs = syn.decode(qc, qo), syn.decode(rc, ro)
        cpu = cpu_reference("gapblock", qs, rs, min(args.cpu_rows, 20),
                            seconds=args.cpu_seconds, pen=gp.pen, go=gp.gap_open,
                            cube=gp.cube.tolist(), width=gp.width)
    line = base_line(
        "learned-prefilter effective pairs/s (exact gap-block hits <= cutoff, reference "
        "embeddings built once; independent synthetic CDR3s)",
        head["effective_pairs_per_s"], args.steps, args.warmup, head["ms_per_search"],
        {"workload": "prefilter_cdr3_50000x1000000_blosum62_go28_central21_cutoff48",
         "queries": Q.n, "refs": R.n, "cutoff": args.pf_cutoff, "seed": seed,
         "train_sequences": len(to) - 1, "train_pairs": len(pa),
         "mined_close_pairs": int(n_close), "epochs": args.pf_epochs,
         "pairs_per_step": pairs, "parallelism": "single GPU"},
        kernels, roofline, cpu,
        **{k: v for k, v in head.items() if k not in ("pairs_per_step", "queries", "refs")},
        planted=per_set["planted"], target="recall >= 0.995 and >= 5x the exhaustive exact pass",
        train_s=round(train_s, 2), data_s=round(data_s, 2), calibration=calib,
        train_loss=[round(x, 5) for x in losses])
    line["dtype"] = "bf16"
    line["arithmetic"] = ("BF16 MFMA with FP32 accumulation (encoder, screen); exact int32 / "
                          "packed int16 gap-block rescoring")
    return line


def run_distance(torch, dev, lib, stream, args, which: str) -> dict:
    rng = np.random.default_rng(SEED + 2)
    if which == "hamming":
        a = dna_set(rng, 50000, 12, 12)
        b = dna_set(rng, 50000, 12, 12)
        fn = lib.lib.st_hamming_matrix_device
    else:
        base = dna_set(rng, 10000, 40, 60)
        reads = base + [mutate(rng, s) for s in base]
        order = rng.permutation(len(reads))
        a = b = [reads[i] for i in order]
        fn = lib.lib.st_levenshtein_matrix_device
    A_, B_ = DeviceSet(torch, dev, a), DeviceSet(torch, dev, b)
    out = torch.empty((A_.n, B_.n), dtype=torch.int32, device=dev)
    st = Status()

    def step():
        rc = fn(A_.bytes.data_ptr(), A_.off.data_ptr(), A_.n, A_.max_len, B_.bytes.data_ptr(),
                B_.off.data_ptr(), B_.n, B_.max_len, out.data_ptr(),
                ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
        check(rc, st, which)

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    pairs = A_.n * B_.n
    label, launches, avg_s = dominant(kernels)
    extra = {}
    if which == "hamming":
        # HBM-write bound: the int32 matrix is the algorithmic traffic (4 B per pair)
        achieved = 4.0 * pairs / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                    "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                    "traffic": traffic_for(label),
                    "kernel": label, "launches": launches, "avg_launch_ms": round(avg_s * 1e3, 4),
                    "bytes_per_launch": 4.0 * pairs}
        workload = "hamming_matrix_acgt_umi12_50000x50000"
    else:
        cells = float(A_.lens.sum()) * float(B_.lens.sum())
        # Myers: 17 64-bit word operations per text column, 2 VALU lane-ops each
        cols = float(A_.n) * float(B_.lens.sum())
        roofline = valu_roofline(label, launches, avg_s, cols * 34, columns_per_launch=cols,
                                 ops_per_column=34) if label.startswith("levenshtein_myers") \
            else None
        extra["gcups"] = cells / (ms_step / 1e3) / 1e9
        workload = "levenshtein_matrix_acgt_reads40-60_20000x20000"
    cpu = None
    if args.cpu_rows > 0:
        rows = 500 if which == "hamming" else 200
        cpu = cpu_reference(which, [bytes(x) for x in a], [bytes(x) for x in b],
                            min(rows, args.cpu_rows), seconds=args.cpu_seconds)
    return base_line(
        f"all-vs-all pairs/s ({which}_matrix)", pairs / (ms_step / 1e3), args.steps,
        args.warmup, ms_step,
        {"workload": workload, "queries": A_.n, "refs": B_.n, "seed": SEED + 2,
         "pairs_per_step": pairs, "parallelism": "single GPU"},
        kernels, roofline, cpu, **extra)


def copy_calibration() -> int:
    """Three copies of a 1 GiB int32 tensor: 1 GiB read + 1 GiB written per dispatch."""
    import torch

    n = 256 * 1024 * 1024
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_dispatch": 4 * n, "dispatches": 3}))
    return 0


TB_PAIRS = 1_000_000


def run_traceback(torch, dev, lib, stream, args) -> dict:
    """Gotoh alignment with traceback of candidate pairs (the step after a thresholded search):
    1,000,000 random (query, reference) pairs of two 20,000-string CDR3-like sets, both modes,
    through st_align_pairs_device (scores, ops and the aligned sub-sequence starts)."""
    rng = np.random.default_rng(SEED + 7)
    qs = cdr3_set(rng, N_QUERIES, LEN_LO, LEN_HI)
    rs = cdr3_set(rng, N_REFS, LEN_LO, LEN_HI)
    Q, R = DeviceSet(torch, dev, qs), DeviceSet(torch, dev, rs)
    pq = rng.integers(0, Q.n, TB_PAIRS).astype(np.uint32)
    pr = rng.integers(0, R.n, TB_PAIRS).astype(np.uint32)
    d_pq = torch.from_numpy(pq.view(np.int32)).to(dev)
    d_pr = torch.from_numpy(pr.view(np.int32)).to(dev)
    sim, _, _ = matrix_tables("blosum62")
    stride = Q.max_len + R.max_len
    ops = torch.empty(TB_PAIRS * stride, dtype=torch.uint8, device=dev)
    score = torch.empty(TB_PAIRS, dtype=torch.int64, device=dev)
    meta = torch.empty(3 * TB_PAIRS, dtype=torch.int32, device=dev)
    st = Status()

    def step():
        for local in (0, 1):
            rc = lib.lib.st_align_pairs_device(
                Q.bytes.data_ptr(), Q.off.data_ptr(), Q.n, Q.max_len, R.bytes.data_ptr(),
                R.off.data_ptr(), R.n, R.max_len, d_pq.data_ptr(), d_pr.data_ptr(), TB_PAIRS,
                sim.ctypes.data, len(AA_ORDER), local, GAP_OPEN, GAP_EXTEND, ops.data_ptr(),
                stride, score.data_ptr(), meta.data_ptr(), meta.data_ptr() + 4 * TB_PAIRS,
                meta.data_ptr() + 8 * TB_PAIRS, ctypes.c_void_p(stream.cuda_stream),
                ctypes.byref(st))
            check(rc, st, "st_align_pairs_device")

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    cells = float((Q.lens[pq].astype(np.float64) * R.lens[pr]).sum())
    label, launches, avg_s = dominant(kernels)
    local = label.endswith("_local")
    # the score recurrences of the cell (K_MIN, as the score kernels); the back-pointers and
    # their byte come on top, so the fraction below understates the work done
    roofline = valu_roofline(label, launches, avg_s, cells * K_MIN["local" if local else "global"],
                             cells_per_launch=cells,
                             ops_per_cell=K_MIN["local" if local else "global"])
    cpu = None
    if args.cpu_rows > 0:
        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        cpu = cpu_traceback([to_str(s) for s in qs], [to_str(s) for s in rs], pq, pr,
                            seconds=args.cpu_seconds, sim=sim)
    return base_line(
        "candidate-pair alignments with traceback/s (Gotoh NW+SW align_pair)",
        2 * TB_PAIRS / (ms_step / 1e3), args.steps, args.warmup, ms_step,
        {"workload": "traceback_cdr3_1000000pairs_of_20000x20000_blosum62_go11_ge1_global+local",
         "queries": Q.n, "refs": R.n, "pairs": TB_PAIRS, "lengths": [LEN_LO, LEN_HI],
         "seed": SEED + 7, "alignments_per_step": 2 * TB_PAIRS, "cells_per_step": 2 * cells,
         "parallelism": "single GPU"},
        kernels, roofline, cpu, gcups=2 * cells / (ms_step / 1e3) / 1e9)


COUNTS_QUERIES, COUNTS_CONTROL, COUNTS_CEILING = 1000, 250_000, 64


def run_counts(torch, dev, lib, stream, args) -> dict:
    """Control-calibrated counts: per-query histograms of the exact gap-block scores of 1,000
    CDR3-like queries against a 250,000-string control (lengths 10..20, BLOSUM62, gap_open 28,
    central prior 21) up to the ceiling 64 -- the one control pass behind threshold_for_evalue
    and the control half of evalues. A step is st_gapblock_histogram_device: the full score
    rows scored by the gap-block kernels, each score added to its query's LDS histogram in the
    kernel (no score matrix is written)."""
    from seqtree_amd.gapblock import _prior_cube, central_prior

    rng = np.random.default_rng(SEED + 9)
    qs = cdr3_set(rng, COUNTS_QUERIES, 10, 20)
    rs = cdr3_set(rng, COUNTS_CONTROL, 10, 20)
    Q, R = DeviceSet(torch, dev, qs), DeviceSet(torch, dev, rs)
    _, pen, scale = matrix_tables("blosum62")
    go, lam = 2 * scale, int(1.5 * scale)
    width = max(Q.max_len, R.max_len)
    cube = np.array(_prior_cube(central_prior(lam), width), np.int32)
    bins = COUNTS_CEILING + 1
    hist = torch.empty((Q.n, bins + 2), dtype=torch.int64, device=dev)
    st = Status()

    def step():
        rc = lib.lib.st_gapblock_histogram_device(
            Q.bytes.data_ptr(), Q.off.data_ptr(), Q.n, Q.max_len, R.bytes.data_ptr(),
            R.off.data_ptr(), R.n, R.max_len, pen.ctypes.data, len(AA_ORDER), go, 1,
            cube.ctypes.data, width, bins, hist.data_ptr(), ctypes.c_void_p(stream.cuda_stream),
            ctypes.byref(st))
        check(rc, st, "st_gapblock_histogram_device")

    ms_step, kernels = time_steps(torch, lib, stream, step, args.steps, args.warmup)
    pairs = Q.n * R.n
    h = hist.cpu().numpy()
    if not (h.sum(axis=1) == R.n).all():
        raise RuntimeError("histogram rows do not sum to the control size")
    L = np.minimum(Q.lens[:, None], R.lens[None, :])
    d = np.abs(Q.lens[:, None] - R.lens[None, :])
    ops = float(np.where(d > 0, 5 * L, L).sum())
    label, launches, avg_s = dominant(kernels)
    # one launch per row block: the step's ops averaged over the launches
    roofline = valu_roofline(label, launches, avg_s, ops * args.steps / launches,
                             ops_per_position=5, row_blocks=launches // args.steps) \
        if label.startswith("gapblock") else None
    hist_label = "score_histogram"
    if hist_label in kernels:
        n_h, ms_h = kernels[hist_label]
        bytes_h = 4.0 * pairs / n_h  # every score of the block is read once
        ach = bytes_h / (ms_h / n_h / 1e3) / 1e9
        info = {"kernel": hist_label, "bound": "hbm", "achieved_gbs": round(ach, 1),
                "peak_gbs": 8000.0, "frac": round(ach / 8000.0, 4),
                "avg_launch_ms": round(ms_h / n_h, 4), "bytes_per_launch": bytes_h}
    else:
        info = None
    cpu = None
    if args.cpu_rows > 0:
        to_str = lambda v: "".join(AA_ORDER[c] for c in v)  # noqa: E731
        cpu = cpu_reference("gapblock", [to_str(s) for s in qs], [to_str(s) for s in rs],
                            max(1, min(args.cpu_rows, Q.n) // 25), seconds=args.cpu_seconds,
                            pen=pen, go=go, cube=cube.tolist(), width=width)
        cpu["sample"] += " (score rows only; the counting on top is not timed)"
    extra = {"histogram_roofline": info} if info else {}
    return base_line(
        "query-control pairs/s (control-calibrated gap-block score counts)",
        pairs / (ms_step / 1e3), args.steps, args.warmup, ms_step,
        {"workload": "counts_cdr3_1000x250000_blosum62_go28_ge1_central21_ceiling64",
         "queries": Q.n, "refs": R.n, "lengths": [10, 20], "seed": SEED + 9, "bins": bins,
         "pairs_per_step": pairs, "parallelism": "single GPU"},
        kernels, roofline, cpu, **extra)


def cpu_traceback(qs, rs, pq, pr, seconds: float = 3.0, **kw) -> dict:
    """The reference's own align_pair (oracle/_ref) over the first pairs of the same list, both
    modes, with 16 worker threads and with every usable core; the plain-C port (one thread)
    where the compiled reference is absent."""
    import oracle

    host = host_cpu()
    have_ref = oracle.reference_available()
    n = 20_000 if have_ref else 1_000

    def once(threads) -> int:
        for local in (False, True):
            if have_ref:
                oracle.Reference().align_pairs(qs, rs, pq[:n], pr[:n], "blosum62", local,
                                                GAP_OPEN, GAP_EXTEND, threads=threads)
            else:
                orc = oracle.Oracle()
                for t in range(n):
                    orc.align_pair(qs[pq[t]], rs[pr[t]], kw["sim"], local, GAP_OPEN, GAP_EXTEND)
        return 2 * n

    threads = CPU_THREADS if have_ref else 1
    rate, calls, pairs, dt = timed_rate(lambda: once(threads), seconds)
    out = {"value": rate, "unit": "alignments/s", "cores": threads,
           "kind": "reference" if have_ref else "port",
           "sample": f"the first {n} pairs of the same list, global + local, {calls} rounds "
                     f"after one warm-up: {pairs} alignments in {dt:.2f} s wall", "host": host}
    if have_ref and host["usable_cpus"] != CPU_THREADS:
        all_rate, _, _, _ = timed_rate(lambda: once(host["usable_cpus"]), seconds)
        out["all_cores"] = {"value": all_rate, "cores": host["usable_cpus"]}
    return out


RUNNERS = {
    "gotoh": run_gotoh,            # C2 (the headline)
    "gapblock": run_gapblock,      # C1
    "hamming": lambda *a: run_distance(*a, "hamming"),          # C3
    "levenshtein": lambda *a: run_distance(*a, "levenshtein"),  # C3
    "batch": run_batch,            # C4
    "prefilter": run_prefilter,    # C5
    "traceback": run_traceback,
    "counts": run_counts,
    "germline": run_germline,
    "longpairs": run_longpairs,
}

# The arithmetic each configuration's dominant kernel computes in (base_line's default int32 is
# the output type).
ARITHMETIC = {
    "gapblock": ("int16", "twin 16-bit rows (two queries per register, packed int16 adds and "
                          "mins; exact under the host's range proof, int32 past it)"),
    "batch": ("int16", "twin 16-bit rows for the short jobs, int32 otherwise (range-proven)"),
    "counts": ("int16", "twin 16-bit rows (range-proven), LDS uint32 histograms"),
    "hamming": ("u8", "byte compares as XOR / zero-byte tests on packed 32-bit words"),
    "levenshtein": ("u64", "Myers bit-vectors in 64-bit words (block carries up to 4 words)"),
    "traceback": ("int32", "int32 Gotoh cells with one direction byte per cell (range-proven; "
                           "int64 past it)"),
}

COMPACT_EXTRAS = ("gcups", "glookups", "end_to_end", "recall", "exact_hits", "found_hits",
                  "candidates", "tau_score", "speedup_search_only", "speedup_one_shot",
                  "index_build_ms", "one_shot_ms", "exhaustive_ms", "target",
                  "histogram_roofline")


def compact(line: dict) -> dict:
    """One configuration's entry of the all-configs line: its value, step time, the roofline
    of its dominant kernel and its CPU baseline, without the per-kernel breakdown."""
    out = {k: line[k] for k in ("metric", "value", "unit", "ms_per_step", "dtype",
                                "arithmetic") if k in line}
    out["workload"] = line["config"]["workload"]
    rf = line.get("roofline")
    if rf:
        out["roofline"] = {k: rf[k] for k in ("bound", "kernel", "achieved", "peak", "unit",
                                              "frac", "traffic", "avg_launch_ms", "lds")
                           if k in rf}
    cb = line.get("cpu_baseline")
    if cb:
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                  "all_cores") if k in cb}
    for k in COMPACT_EXTRAS:
        if line.get(k) is not None:
            out[k] = line[k]
    if "planted" in line:
        out["planted"] = {k: line["planted"][k] for k in (
            "effective_pairs_per_s", "ms_per_search", "recall", "exact_hits", "candidates",
            "speedup_search_only", "speedup_one_shot", "exhaustive_ms", "one_shot_ms")}
    return out


def finish_roofline(line: dict, lib) -> None:
    """The measured peak beside the published one for the dominant kernel's bound."""
    rf = line.get("roofline")
    if not rf:
        return
    if rf["bound"] == "valu":
        rf["probe_tops"] = valu_probe(lib)
    if "lds_bytes_per_lookup" in rf:
        # the co-bound of the table-lookup kernels: one LDS word per lookup (4 bytes; with
        # twin rows one word serves two pairs, 2 bytes per pair-lookup), prior reads not
        # counted, against the measured conflict-free 4-byte LDS read rate
        rate = ctypes.c_double()
        st = Status()
        if lib.lib.st_probe_lds(1, ctypes.byref(rate), ctypes.byref(st)) == 0:
            peak = rate.value / 1e12
            achieved = (rf["lookups_per_launch"] * rf["lds_bytes_per_lookup"]
                        / (rf["avg_launch_ms"] / 1e3) / 1e12)
            rf["lds"] = {"achieved": round(achieved, 2), "peak": round(peak, 2),
                         "unit": "TB/s", "frac": round(achieved / peak, 4),
                         "peak_basis": "st_probe_lds, 4-byte conflict-free reads"}
    elif rf["bound"] == "mfma":
        rate = ctypes.c_double()
        st = Status()
        if lib.lib.st_probe_mfma(ctypes.byref(rate), ctypes.byref(st)) == 0:
            rf["probe_tflops_bf16_32x32x16"] = round(rate.value / 1e12, 1)
    elif rf["bound"] == "hbm":
        # the attainable write rate: a plain 16-byte-store fill of 4 GiB
        rate = ctypes.c_double()
        st = Status()
        if lib.lib.st_probe_hbm_write(4 << 30, ctypes.byref(rate), ctypes.byref(st)) == 0:
            rf["probe_write_gbs"] = round(rate.value / 1e9, 1)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-rows", type=int, default=1000,
                    help="query rows of the CPU baseline sample (0 skips it)")
    ap.add_argument("--cpu-seconds", type=float, default=3.0,
                    help="wall time each CPU baseline sample is repeated for")
    ap.add_argument("--workload", default="all", choices=["all", *RUNNERS],
                    help="'all' (default): every configuration, the Gotoh dist_matrix (C2) as "
                         "the headline value and the rest under 'configs'")
    ap.add_argument("--no-end-to-end", dest="end_to_end", action="store_false",
                    help="skip the end-to-end (Python call, host output) timings")
    ap.add_argument("--pf-epochs", type=int, default=PF_EPOCHS)
    ap.add_argument("--pf-cutoff", type=int, default=PF_CUTOFF)
    ap.add_argument("--pf-quantile", type=float, default=PF_QUANTILE,
                    help="quantile of the calibration hits' model distances that sets the "
                         "screening threshold")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale for rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the all-vs-all matrix shards by query rows with no collective between GPUs)",
              file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    lib = Lib()
    stream = torch.cuda.Stream(device=dev)
    if args.workload != "all":
        line = RUNNERS[args.workload](torch, dev, lib, stream, args)
        if args.workload in ARITHMETIC:
            line["dtype"], line["arithmetic"] = ARITHMETIC[args.workload]
        finish_roofline(line, lib)
        print(json.dumps(line))
        return 0
    t_all = time.perf_counter()
    line = run_gotoh(torch, dev, lib, stream, args)
    finish_roofline(line, lib)
    line.pop("kernels_ms", None)
    configs = {}
    for name, fn in RUNNERS.items():
        if name == "gotoh":
            continue
        t0 = time.perf_counter()
        print(f"bench.py: {name} ...", file=sys.stderr, flush=True)
        try:
            sub = fn(torch, dev, lib, stream, args)
            if name in ARITHMETIC:
                sub["dtype"], sub["arithmetic"] = ARITHMETIC[name]
            finish_roofline(sub, lib)
            configs[name] = compact(sub)
        except Exception as e:  # noqa: BLE001 -- one configuration's failure is reported
            configs[name] = {"error": f"{type(e).__name__}: {e}"}
        configs.setdefault(name, {})["wall_s"] = round(time.perf_counter() - t0, 1)
    try:
        configs["scalar_latency"] = scalar_latency(args)
    except Exception as e:  # noqa: BLE001
        configs["scalar_latency"] = {"error": f"{type(e).__name__}: {e}"}
    line["configs"] = configs
    line["bench_wall_s"] = round(time.perf_counter() - t_all, 1)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
