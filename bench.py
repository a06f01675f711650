#!/usr/bin/env python
"""All-vs-all Gotoh scoring throughput on one MI355X.

Workload (one "step"): the 20,000 x 20,000 all-vs-all Gotoh affine score matrix of synthetic
CDR3-like amino-acid strings ('C' + body + 'F', total length uniform in 8..25, body residues
i.i.d. at Robinson & Robinson background frequencies, numpy PCG64 seed 2), BLOSUM62, gap_open 11,
gap_extend 1, computed twice -- global (Needleman-Wunsch) and local (Smith-Waterman) -- into
two device-resident int32 matrices. Inputs (codec-encoded CSR strings) are already in HBM when
the timed region starts; each step runs the full device path (reference packing, query
bucketing, scoring kernels) through the C-ABI entry point st_align_matrix_device on one HIP
stream.

Prints one JSON line: pairs/s (value), GCUPS, the roofline of the dominant kernel measured
with HIP events on its own stream, and the reference's multithreaded CPU path timed on a
bounded sample on this host.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
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

# VALU lane-instructions per DP cell of the register kernel's cell update (see
# seqtree_amd/csrc/gotoh.hip, "H form"): M = Hg_diag + s', X = max(Hg_up, X_up - ge),
# Y = max(Hg_left, Y_left - ge), Hg = max(M, X, Y) - go -> 8 instructions (7 with v_max3 in
# int32), local adds the 0 floor and the running best (+2). A packed int16 instruction
# advances two cells, so the i16x2 kernels spend half as many per cell.
K_MIN = {("i16x2", "global"): 4.0, ("i16x2", "local"): 5.0,
         ("i32", "global"): 7.0, ("i32", "local"): 9.0}
# VALU lane-instruction issue peak of MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12


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


def csr(seqs: list[np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return np.concatenate(seqs).astype(np.uint8), off


def blosum62_sim() -> np.ndarray:
    import seqtree_amd as st

    m = st.SubstitutionMatrix.blosum62()
    return np.array([[m.similarity(a, b) for b in AA_ORDER] for a in AA_ORDER], np.int32)


class Lib:
    def __init__(self):
        self.lib = ctypes.CDLL(str(ROOT / "seqtree_amd" / "libseqtree_hip.so"))
        L = self.lib
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
        L.st_align_matrix_device.restype = i32
        L.st_align_matrix_device.argtypes = [vp, vp, u64, u32, vp, vp, u64, u32, vp, u32, i32,
                                             i32, i32, i32, vp, vp, vp]
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


def cpu_baseline(qs, rs, sim, rows: int) -> dict:
    """The reference's own align_score_matrix (oracle/_ref, compiled -O3 from its sources) on
    the first `rows` queries against all references, global and local, on every host core."""
    import oracle

    threads = oracle.host_threads()
    q = ["".join(AA_ORDER[c] for c in s) for s in qs[:rows]]
    r = ["".join(AA_ORDER[c] for c in s) for s in rs]
    if oracle.reference_available():
        ref = oracle.Reference()
        kind = "reference"
        run = lambda local: ref.align_matrix(q, r, "blosum62", local, False, GAP_OPEN,  # noqa
                                             GAP_EXTEND, threads=threads)
    else:
        orc = oracle.Oracle()
        kind = "port"
        run = lambda local: orc.align_matrix(q, r, sim, local, False, GAP_OPEN,  # noqa
                                             GAP_EXTEND, threads=threads)[1]
    t0 = time.perf_counter()
    for local in (False, True):
        run(local)
    dt = time.perf_counter() - t0
    pairs = 2 * len(q) * len(r)
    return {"value": pairs / dt, "unit": "pairs/s", "cores": threads, "kind": kind,
            "sample": f"first {len(q)} queries x all {len(r)} references of the same workload, "
                      f"global + local ({pairs} pairs, {dt:.1f} s wall)"}


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-rows", type=int, default=1000,
                    help="query rows of the CPU baseline sample (0 skips it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale for rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the all-vs-all matrix shards by query rows with no collective, see DESIGN.md)",
              file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED)
    qs = cdr3_set(rng, N_QUERIES, LEN_LO, LEN_HI)
    rs = cdr3_set(rng, N_REFS, LEN_LO, LEN_HI)
    qb, qo = csr(qs)
    rb, ro = csr(rs)
    q_len = np.diff(qo.astype(np.int64))
    r_len = np.diff(ro.astype(np.int64))
    sim = blosum62_sim()
    A = len(AA_ORDER)

    tq_b = torch.from_numpy(qb).to(dev)
    tq_o = torch.from_numpy(qo.view(np.int64)).to(dev)
    tr_b = torch.from_numpy(rb).to(dev)
    tr_o = torch.from_numpy(ro.view(np.int64)).to(dev)
    outs = {mode: torch.empty((N_QUERIES, N_REFS), dtype=torch.int32, device=dev)
            for mode in ("global", "local")}
    lib = Lib()
    stream = torch.cuda.Stream(device=dev)
    st = Status()

    def step():
        for mode in ("global", "local"):
            rc = lib.lib.st_align_matrix_device(
                tq_b.data_ptr(), tq_o.data_ptr(), N_QUERIES, int(q_len.max()), tr_b.data_ptr(),
                tr_o.data_ptr(), N_REFS, int(r_len.max()), sim.ctypes.data, A,
                int(mode == "local"), 0, GAP_OPEN, GAP_EXTEND, outs[mode].data_ptr(),
                ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
            if rc != 0:
                raise RuntimeError(f"st_align_matrix_device: {rc} {st.message.decode()}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.lib.st_timing_enable(1)
    lib.timing()  # drop anything recorded before the timed region
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    lib.lib.st_timing_enable(0)
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    kernels = lib.timing()

    pairs_per_step = 2 * N_QUERIES * N_REFS
    cells_per_mode = float(q_len.sum()) * float(r_len.sum())
    value = pairs_per_step / (ms_step / 1e3)
    gcups = 2 * cells_per_mode / (ms_step / 1e3) / 1e9

    # dominant kernel: the label with the largest total time; its algorithmic work per launch
    # is sum(m) * sum(n) over the queries of its length class, times K_MIN ops per cell
    label, (launches, k_ms) = max(kernels.items(), key=lambda kv: kv[1][1])
    roofline = None
    if label.startswith("gotoh_i"):
        _, cell, rtag, mode = label.split("_")
        rows = int(rtag[1:])
        in_class = (q_len <= rows) & (q_len > rows - 4)
        cells = float(q_len[in_class].sum()) * float(r_len.sum())
        avg_s = k_ms / launches / 1e3
        achieved = cells * K_MIN[(cell, mode)] / avg_s / 1e12
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
                    "unit": "Tops/s", "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
                    "kernel": label, "launches": launches,
                    "avg_launch_ms": round(k_ms / launches, 4),
                    "cells_per_launch": cells, "ops_per_cell": K_MIN[(cell, mode)]}

    cpu = None
    if args.cpu_rows > 0:
        cpu = cpu_baseline(qs, rs, sim, args.cpu_rows)

    line = {
        "metric": "all-vs-all pairs/s (Gotoh NW+SW score_matrix)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "gcups": gcups,
        "config": {"workload": "gotoh_score_matrix_cdr3_20000x20000_blosum62_go11_ge1_global+local",
                   "queries": N_QUERIES, "refs": N_REFS, "lengths": [LEN_LO, LEN_HI],
                   "seed": SEED, "pairs_per_step": pairs_per_step,
                   "cells_per_step": 2 * cells_per_mode, "parallelism": "single GPU"},
        "kernels_ms": {k: round(v[1] / args.steps, 4) for k, v in sorted(kernels.items())},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
