"""Seeded synthetic CDR3-like amino-acid sets (the benchmark and prefilter workloads).

The reference ships no generator for these workloads, so the generator is this project's:
'C' + body + 'F' with body residues drawn i.i.d. from the 20 standard amino acids at the
Robinson & Robinson background frequencies (PNAS 88:8880, 1991, as tabulated in NCBI BLAST),
and "clonal families" of sequences derived from a root by a few point edits.

Sets are returned in codec codes (``AA_ORDER``, the reference's amino-acid codec order) as
CSR ``(codes uint8, offsets uint64)``.
"""
from __future__ import annotations

import numpy as np

AA_ORDER = "ARNDCQEGHILKMFPSTWYVBZX*"
BG_RESIDUES = "ARNDCQEGHILKMFPSTWYV"
BG_FREQ = np.array([0.07805, 0.05129, 0.04487, 0.05364, 0.01925, 0.04264, 0.06295, 0.07377,
                    0.02199, 0.05142, 0.09019, 0.05744, 0.02243, 0.03856, 0.05203, 0.07120,
                    0.05841, 0.01330, 0.03216, 0.06441])
BG_FREQ = BG_FREQ / BG_FREQ.sum()
_BG_CODES = np.array([AA_ORDER.index(c) for c in BG_RESIDUES], np.uint8)
_C, _F = AA_ORDER.index("C"), AA_ORDER.index("F")


def cdr3_csr(rng: np.random.Generator, n: int, lo: int, hi: int):
    """n strings 'C' + body + 'F' with total length uniform in [lo, hi]: (codes, offsets,
    lengths)."""
    lens = rng.integers(lo, hi + 1, size=n).astype(np.int64)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    codes = np.empty(int(off[-1]), np.uint8)
    body = np.ones(len(codes), bool)
    starts = off[:-1].astype(np.int64)
    body[starts] = False
    body[starts + lens - 1] = False
    codes[body] = _BG_CODES[rng.choice(20, size=int(body.sum()), p=BG_FREQ)]
    codes[starts] = _C
    codes[starts + lens - 1] = _F
    return codes, off, lens


def split(codes: np.ndarray, off: np.ndarray) -> list[np.ndarray]:
    return [codes[int(off[i]):int(off[i + 1])] for i in range(len(off) - 1)]


def join(seqs: list[np.ndarray]):
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    codes = np.concatenate(seqs).astype(np.uint8) if seqs else np.zeros(0, np.uint8)
    return codes, off


def mutate(rng: np.random.Generator, s: np.ndarray, lo: int, hi: int) -> np.ndarray:
    """1-3 point edits inside the body (substitution 70%, insertion 15%, deletion 15%),
    keeping the length within [lo, hi] and the 'C' / 'F' ends."""
    t = list(s)
    for _ in range(int(rng.integers(1, 4))):
        op = rng.random()
        body = len(t) - 2
        if op < 0.15 and len(t) < hi:
            p = 1 + int(rng.integers(0, body + 1))
            t.insert(p, int(_BG_CODES[rng.choice(20, p=BG_FREQ)]))
        elif op < 0.30 and len(t) > lo and body > 0:
            del t[1 + int(rng.integers(0, body))]
        elif body > 0:
            t[1 + int(rng.integers(0, body))] = int(_BG_CODES[rng.choice(20, p=BG_FREQ)])
    return np.array(t, np.uint8)


def families(rng: np.random.Generator, n_families: int, per_family: int, lo: int, hi: int):
    """n_families roots, each with per_family members derived by mutate(): (codes, offsets,
    family id per sequence). Members of a family are consecutive."""
    rc, ro, _ = cdr3_csr(rng, n_families, lo, hi)
    roots = split(rc, ro)
    seqs = [mutate(rng, r, lo, hi) for r in roots for _ in range(per_family)]
    fam = np.repeat(np.arange(n_families, dtype=np.int64), per_family)
    codes, off = join(seqs)
    return codes, off, fam


def decode(codes: np.ndarray, off: np.ndarray) -> list[str]:
    lut = np.frombuffer(AA_ORDER.encode(), np.uint8)
    return [lut[codes[int(off[i]):int(off[i + 1])]].tobytes().decode()
            for i in range(len(off) - 1)]
