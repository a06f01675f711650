"""Needleman-Wunsch and Smith-Waterman with affine gaps, all-vs-all on the GPU.

These maximise the signed similarity view of a :class:`SubstitutionMatrix`. A gap run of
length ``L`` costs ``gap_open + (L - 1) * gap_extend`` (``gap_open == gap_extend`` is linear);
``mode="global"`` charges end gaps, ``mode="local"`` floors at zero and takes the best cell.
Gap costs are positive magnitudes. Results are exact: the kernels reproduce the reference's
int64 dynamic programme and raise ``OverflowError`` where it does.

Reference behaviour: ``python/seqtree/pairwise.py`` of seqtree 0.6.0 -- ``score`` :55-90,
``align`` :93-116, ``score_matrix`` :119-136, ``dist_matrix`` :139-167, over
``src/pairwise.cpp``. ``align_pairs`` (the traceback of many candidate pairs in one device
batch) has no reference counterpart.
"""
from __future__ import annotations

from collections.abc import Sequence

from ._core import Alignment, ScoreMatrix, SubstitutionMatrix
from ._core import align_dist_matrix as _dist_matrix
from ._core import align_pair as _align_pair
from ._core import align_pairs as _align_pairs
from ._core import align_score as _align_score
from ._core import align_score_matrix as _score_matrix

__all__ = ["score", "align", "align_pairs", "score_matrix", "dist_matrix"]


def score(
    query: str,
    ref: str,
    matrix: SubstitutionMatrix,
    mode: str = "global",
    gap_open: int = 11,
    gap_extend: int = 1,
    alphabet: str = "aa",
) -> int:
    """Optimal alignment score of one pair (signed; higher is more similar)."""
    return _align_score(query, ref, matrix, mode=mode, gap_open=gap_open,
                        gap_extend=gap_extend, alphabet=alphabet)


def align(
    query: str,
    ref: str,
    matrix: SubstitutionMatrix,
    mode: str = "global",
    gap_open: int = 11,
    gap_extend: int = 1,
    alphabet: str = "aa",
) -> Alignment:
    """As :func:`score`, but also returns the aligned strings and the edit ops.

    ``Alignment.score`` is the similarity (signed, higher is better). In local mode the
    aligned strings are the matched sub-sequences only.

    >>> a = align("CASSLGQAYEQYF", "CASSPGQAYEQF", SubstitutionMatrix.blosum62())
    >>> a.aligned_query, a.aligned_ref
    ('CASSLGQAYEQYF', 'CASSPGQAYEQ-F')
    """
    return _align_pair(query, ref, matrix, mode=mode, gap_open=gap_open,
                       gap_extend=gap_extend, alphabet=alphabet)


def align_pairs(
    queries: Sequence[str],
    refs: Sequence[str],
    pair_query: Sequence[int],
    pair_ref: Sequence[int],
    matrix: SubstitutionMatrix,
    mode: str = "global",
    gap_open: int = 11,
    gap_extend: int = 1,
    alphabet: str = "aa",
) -> list[Alignment]:
    """:func:`align` for every ``(queries[pair_query[p]], refs[pair_ref[p]])``, traced back in
    one device batch (for example the candidate pairs of a thresholded search)."""
    return _align_pairs(list(queries), list(refs), [int(x) for x in pair_query],
                        [int(x) for x in pair_ref], matrix, mode=mode, gap_open=gap_open,
                        gap_extend=gap_extend, alphabet=alphabet)


def score_matrix(
    queries: Sequence[str],
    refs: Sequence[str],
    matrix: SubstitutionMatrix,
    mode: str = "global",
    gap_open: int = 11,
    gap_extend: int = 1,
    alphabet: str = "aa",
    threads: int = 0,
) -> ScoreMatrix:
    """Every query against every reference: a ``(len(queries), len(refs))`` similarity matrix."""
    return _score_matrix(list(queries), list(refs), matrix, mode=mode, gap_open=gap_open,
                         gap_extend=gap_extend, alphabet=alphabet, threads=threads)


def dist_matrix(
    queries: Sequence[str],
    refs: Sequence[str],
    matrix: SubstitutionMatrix,
    mode: str = "global",
    gap_open: int = 11,
    gap_extend: int = 1,
    alphabet: str = "aa",
    threads: int = 0,
) -> ScoreMatrix:
    """Alignment distances ``d(a, b) = s(a,a) + s(b,b) - 2 s(a,b)`` for every pair."""
    return _dist_matrix(list(queries), list(refs), matrix, mode=mode, gap_open=gap_open,
                        gap_extend=gap_extend, alphabet=alphabet, threads=threads)
