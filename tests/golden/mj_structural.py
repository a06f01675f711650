"""Regenerates the "structural" similarity grid from the Miyazawa-Jernigan contact energies.

Test data, not product code: tests/test_host.py checks that the product's built-in structural
matrix equals what this construction yields, so the table is pinned by its published
definition rather than by a copy of the reference's .inc file.

Construction (seqtree 0.6.0, bench/gen_matrices.py:134-165, after the A-N correction of
CHANGELOG.md:10-27):
  * e(a, b): MJ residue-residue contact energies (Miyazawa & Jernigan, J Mol Biol 1996, as
    tabulated by Keskin et al., Protein Sci 1998), symmetric;
  * interaction strength q(a) = mean over the 20 residues b of e(a, b), min-max normalised;
  * sim(a, b) = 10 * (1 - |q(a) - q(b)|) for the 20 standard residues;
  * B, Z, X are bilinear combinations: B = (N + D) / 2, Z = (Q + E) / 2, X = mean of all 20;
  * every entry is rounded half-to-even; the stop '*' scores 0 against everything but itself
    (10).
"""
from __future__ import annotations

ORDER = "ARNDCQEGHILKMFPSTWYVBZX*"
RESIDUES = "ACDEFGHIKLMNPQRSTVWY"

# Lower triangle of e(a, b) in RESIDUES order: row a lists b = RESIDUES[0..a].
MJ_LOWER = [
    -0.12,
    -0.33, -1.19,
     0.27,  0.12,  0.29,
     0.38,  0.20,  0.44,  0.46,
    -0.36, -0.67,  0.18,  0.14, -0.88,
    -0.08, -0.31,  0.11,  0.32, -0.19, -0.29,
     0.07, -0.36, -0.10,  0.00, -0.34,  0.00, -0.40,
    -0.37, -0.64,  0.22,  0.17, -0.73, -0.13, -0.13, -0.74,
     0.41,  0.33, -0.01, -0.06,  0.19,  0.29,  0.38,  0.24,  0.76,
    -0.38, -0.65,  0.27,  0.17, -0.80, -0.16, -0.18, -0.81,  0.22, -0.84,
    -0.27, -0.61,  0.30,  0.12, -0.83, -0.17, -0.29, -0.66,  0.29, -0.70, -0.70,
     0.15, -0.01,  0.02,  0.12, -0.01, -0.01,  0.00,  0.14,  0.22,  0.04,  0.04, -0.06,
     0.15, -0.18,  0.33,  0.37, -0.19,  0.02,  0.01, -0.05,  0.47, -0.12, -0.13,  0.18,  0.11,
     0.22, -0.07,  0.24,  0.27, -0.11,  0.13,  0.15, -0.01,  0.28, -0.04, -0.06,  0.06,  0.17,  0.20,
     0.24,  0.08, -0.24, -0.22, -0.05,  0.09,  0.05,  0.00,  0.66, -0.04,  0.03,  0.10,  0.17,  0.09,  0.19,
     0.10, -0.13,  0.10,  0.18, -0.12, -0.01,  0.04,  0.03,  0.36, -0.02,  0.05,  0.09,  0.20,  0.22,  0.16,  0.05,
     0.04, -0.15,  0.11,  0.16, -0.15, -0.04, -0.03, -0.15,  0.33, -0.15, -0.11,  0.04,  0.13,  0.12,  0.11,  0.04,  0.03,
    -0.32, -0.59,  0.36,  0.26, -0.67, -0.15, -0.06, -0.67,  0.29, -0.74, -0.51,  0.12, -0.05,  0.08,  0.08,  0.04, -0.07, -0.65,
    -0.27, -0.66,  0.07,  0.00, -0.68, -0.25, -0.37, -0.60,  0.09, -0.62, -0.73, -0.10, -0.37, -0.02, -0.21, -0.01, -0.02, -0.51, -0.64,
    -0.20, -0.39, -0.07, -0.08, -0.58, -0.22, -0.30, -0.49, -0.05, -0.55, -0.56, -0.11, -0.25, -0.14, -0.25, -0.08, -0.09, -0.38, -0.49, -0.45,
]


def contact_energy(a: str, b: str) -> float:
    i, j = RESIDUES.index(a), RESIDUES.index(b)
    if j > i:
        i, j = j, i
    return MJ_LOWER[i * (i + 1) // 2 + j]


def structural_grid() -> list[list[int]]:
    n = len(RESIDUES)
    strength = {a: sum(contact_energy(a, b) for b in RESIDUES) / n for a in RESIDUES}
    lo, hi = min(strength.values()), max(strength.values())
    norm = {a: (strength[a] - lo) / (hi - lo) for a in RESIDUES}
    base = {(a, b): 10 * (1 - abs(norm[a] - norm[b])) for a in RESIDUES for b in RESIDUES}
    weights = {"B": {"N": 0.5, "D": 0.5}, "Z": {"Q": 0.5, "E": 0.5},
               "X": {a: 1 / n for a in RESIDUES}}

    def sim(a: str, b: str) -> float:
        wa = weights.get(a, {a: 1.0})
        wb = weights.get(b, {b: 1.0})
        return sum(x * y * base[(p, r)] for p, x in wa.items() for r, y in wb.items())

    grid = [[0] * len(ORDER) for _ in ORDER]
    for i, a in enumerate(ORDER):
        for j, b in enumerate(ORDER):
            if a == "*" or b == "*":
                grid[i][j] = 10 if a == b else 0
            else:
                grid[i][j] = round(sim(a, b))
    return grid
