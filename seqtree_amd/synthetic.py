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
    return _mutate_many(rng, [s], 1, lo, hi)[0]


def _mutate_many(rng: np.random.Generator, roots: list[np.ndarray], per: int, lo: int,
                 hi: int) -> list[np.ndarray]:
    """`per` mutants of every root (see mutate), with the random draws made in bulk."""
    n = len(roots) * per
    k = rng.integers(1, 4, size=n)
    op = rng.random((n, 3))
    pos = rng.random((n, 3))
    res = _BG_CODES[rng.choice(20, size=(n, 3), p=BG_FREQ)].tolist()
    k, op, pos = k.tolist(), op.tolist(), pos.tolist()
    out = []
    j = 0
    for r in roots:
        base = r.tolist()
        for _ in range(per):
            t = list(base)
            for e in range(k[j]):
                body = len(t) - 2
                o = op[j][e]
                if o < 0.15 and len(t) < hi:
                    t.insert(1 + int(pos[j][e] * (body + 1)), res[j][e])
                elif o < 0.30 and len(t) > lo and body > 0:
                    del t[1 + int(pos[j][e] * body)]
                elif body > 0:
                    t[1 + int(pos[j][e] * body)] = res[j][e]
            out.append(np.array(t, np.uint8))
            j += 1
    return out


def families(rng: np.random.Generator, n_families: int, per_family: int, lo: int, hi: int,
             roots_out: list | None = None):
    """n_families roots, each with per_family members derived by mutate(): (codes, offsets,
    family id per sequence). Members of a family are consecutive; the roots themselves are
    appended to roots_out when given."""
    rc, ro, _ = cdr3_csr(rng, n_families, lo, hi)
    roots = split(rc, ro)
    seqs = _mutate_many(rng, roots, per_family, lo, hi)
    fam = np.repeat(np.arange(n_families, dtype=np.int64), per_family)
    codes, off = join(seqs)
    if roots_out is not None:
        roots_out.extend(roots)
    return codes, off, fam


def padded(codes: np.ndarray, off: np.ndarray, width: int) -> np.ndarray:
    """Sequences as rows of a [n][width] array (zero past each length)."""
    lens = np.diff(off.astype(np.int64))
    out = np.zeros((len(lens), width), np.uint8)
    mask = np.arange(width)[None, :] < lens[:, None]
    out[mask] = codes[:int(off[-1])]
    return out


def decode(codes: np.ndarray, off: np.ndarray) -> list[str]:
    lut = np.frombuffer(AA_ORDER.encode(), np.uint8)
    return [lut[codes[int(off[i]):int(off[i + 1])]].tobytes().decode()
            for i in range(len(off) - 1)]


def training_set(rng: np.random.Generator, n_iid: int, n_families: int, per_family: int,
                 lo: int, hi: int):
    """A prefilter training set of CDR3-like sequences: n_iid independent draws of the
    generator followed by n_families clonal families of per_family members, as (codes,
    offsets). Independent draws supply the chance similarities of unrelated sequences, the
    families the close pairs that differ by point edits, insertions and deletions included."""
    c1, o1, _ = cdr3_csr(rng, n_iid, lo, hi)
    if n_families == 0:
        return c1, o1
    c2, o2, _ = families(rng, n_families, per_family, lo, hi)
    return np.concatenate([c1, c2]), np.concatenate([o1, o1[-1] + o2[1:]])
