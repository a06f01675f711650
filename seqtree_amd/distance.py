"""Unit-cost Hamming and Levenshtein distances on raw, case-sensitive bytes, on the GPU.

Reference behaviour: ``python/seqtree/distance.py`` of seqtree 0.6.0 (``hamming`` :39-59,
``levenshtein`` :62-78, ``hamming_matrix`` :81-102, ``levenshtein_matrix`` :105-119) over
``src/distance.cpp``. Hamming requires equal lengths and raises ``ValueError`` naming the
first mismatched pair otherwise.
"""
from __future__ import annotations

from collections.abc import Sequence

from ._core import ScoreMatrix
from ._core import hamming as _hamming
from ._core import hamming_matrix as _hamming_matrix
from ._core import levenshtein as _levenshtein
from ._core import levenshtein_matrix as _levenshtein_matrix

__all__ = ["hamming", "levenshtein", "hamming_matrix", "levenshtein_matrix"]


def hamming(a: str, b: str) -> int:
    """Number of differing positions of two equal-length strings."""
    return _hamming(a, b)


def levenshtein(a: str, b: str) -> int:
    """Fewest single-character insertions, deletions and substitutions from ``a`` to ``b``."""
    return _levenshtein(a, b)


def hamming_matrix(a: Sequence[str], b: Sequence[str], threads: int = 0) -> ScoreMatrix:
    """Hamming distance of every ``a`` against every ``b`` as an int32 ``ScoreMatrix``."""
    return _hamming_matrix(list(a), list(b), threads=threads)


def levenshtein_matrix(a: Sequence[str], b: Sequence[str], threads: int = 0) -> ScoreMatrix:
    """Levenshtein distance of every ``a`` against every ``b`` as an int32 ``ScoreMatrix``."""
    return _levenshtein_matrix(list(a), list(b), threads=threads)
