#!/usr/bin/env python
"""Writes tests/golden/golden.json: input/output vectors produced by the reference's own C++
(oracle/_ref/libseqtree_ref.so, compiled from the reference sources by oracle/Makefile).

The fixtures are data only -- sequences, parameters, the prior cube's generating rule and the
reference's output matrices -- so parity can be checked where the reference tree is absent
(the GPU host). Regenerate with:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import random
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from seqtree_amd.gapblock import _prior_cube  # noqa: E402

AA = "ACDEFGHIKLMNPQRSTVWY"

# Named priors the fixtures use; tests rebuild the cube from these rules.
PRIORS = {
    "none": None,
    "central21": lambda i, d, m: 0 if d == 0 else int(21 * abs(2 * i + d - m) // 2),
    "frame21_5": lambda i, d, m: 0 if d == 0 else 21 * abs(i - 5),
    "positions3443": None,  # filled below
}


def _positions(starts):
    def prior(i, d, m):
        if d == 0:
            return 0
        shorter = m - d
        for p in starts:
            pos = min(p, shorter) if p >= 0 else max(0, shorter + p)
            if pos == i:
                return 0
        return 1 << 20

    return prior


PRIORS["positions3443"] = _positions((3, 4, -4, -3))


def seqs(rng, n, lo, hi, alphabet=AA):
    return ["".join(rng.choice(alphabet) for _ in range(rng.randint(lo, hi))) for _ in range(n)]


def main() -> None:
    ref = oracle.Reference()
    rng = random.Random(20241203)
    cases = []

    q, r = seqs(rng, 24, 0, 22), seqs(rng, 30, 0, 22)
    w = max(len(s) for s in q + r)
    for matrix, go, ge in (("blosum62", 28, 1), ("structural", 14, 1), ("pam100", 0, 3)):
        for pname, prior in PRIORS.items():
            cube = _prior_cube(prior, w) if prior else []
            out = ref.gapblock_matrix(q, r, matrix, go, ge, cube, w)
            cases.append({"kind": "gapblock", "matrix": matrix, "gap_open": go, "gap_extend": ge,
                          "prior": pname, "alphabet": "aa", "queries": q, "refs": r,
                          "expected": out.tolist()})
    qn, rn = seqs(rng, 12, 0, 16, "ACGT"), seqs(rng, 15, 0, 16, "ACGT")
    cases.append({"kind": "gapblock", "matrix": None, "gap_open": 1, "gap_extend": 1,
                  "prior": "none", "alphabet": "nt", "queries": qn, "refs": rn,
                  "expected": ref.gapblock_matrix(qn, rn, None, 1, 1, [], 0, "nt").tolist()})

    q, r = seqs(rng, 20, 0, 26), seqs(rng, 25, 0, 26)
    for matrix, go, ge in (("blosum62", 11, 1), ("blosum45", 5, 5), ("blosum80", 1, 3),
                           ("pam250", 0, 0), ("blosum62", 1000, 3), ("blosum62", 10**8, 1),
                           ("blosum62", 2**31 - 1, 2**31 - 1)):
        for local in (False, True):
            for dist in (False, True):
                try:
                    out = ref.align_matrix(q, r, matrix, local, dist, go, ge).tolist()
                    err = None
                except oracle.ReferenceError_ as e:
                    out, err = None, str(e)
                cases.append({"kind": "gotoh", "matrix": matrix, "gap_open": go,
                              "gap_extend": ge, "mode": "local" if local else "global",
                              "dist": dist, "queries": q, "refs": r, "expected": out,
                              "error": err})
    long_q = seqs(rng, 4, 60, 110)
    out = ref.align_matrix(long_q, r[:10], "blosum62", False, True, 12, 1).tolist()
    cases.append({"kind": "gotoh", "matrix": "blosum62", "gap_open": 12, "gap_extend": 1,
                  "mode": "global", "dist": True, "queries": long_q, "refs": r[:10],
                  "expected": out, "error": None})

    a = seqs(rng, 20, 12, 12, "ACGT")
    cases.append({"kind": "hamming", "queries": a, "refs": a[:15],
                  "expected": ref.distance_matrix("hamming", a, a[:15]).tolist()})
    a = seqs(rng, 16, 0, 70, "ACGT") + seqs(rng, 4, 100, 300, "ACGT")
    b = seqs(rng, 18, 0, 70, "ACGT")
    cases.append({"kind": "levenshtein", "queries": a, "refs": b,
                  "expected": ref.distance_matrix("levenshtein", a, b).tolist()})

    (HERE / "golden.json").write_text(json.dumps({"source": "seqtree 0.6.0 C++ (oracle/_ref)",
                                                  "cases": cases}, separators=(",", ":")))
    print(f"wrote {len(cases)} cases to {HERE / 'golden.json'}")


if __name__ == "__main__":
    main()
