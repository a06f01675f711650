"""seqtree_amd: seqtree's all-vs-all scoring path on AMD Instinct MI355X (gfx950).

A drop-in for the reference's dense-matrix API (seqtree 0.6.0, ``python/seqtree/__init__.py``
hot-path re-exports): ``score_matrix`` and the gap priors (:mod:`seqtree_amd.gapblock`), Gotoh
``score_matrix`` / ``dist_matrix`` (:mod:`seqtree_amd.pairwise`) and ``hamming_matrix`` /
``levenshtein_matrix`` (:mod:`seqtree_amd.distance`), plus Gotoh alignment with traceback
(``pairwise.align``, ``pairwise.align_pairs``; ``Alignment``), the penalty-space
``Index.align`` of a query to a stored reference (``Index``, ``SearchParams``; batched as
``Index.align_batch``) and control-calibrated E-values
over exhaustive gap-block balls (:mod:`seqtree_amd.evalue`). Results come back as the same
zero-copy ``ScoreMatrix``. The scoring runs in hand-written HIP kernels; importing this package loads
the native library and fails if it has not been built.
"""
from ._core import (
    Alignment, Index, ScoreMatrix, SearchParams, SubstitutionMatrix, alphabet_symbols,
    amino_acids,
)
from ._core import device_count, version
from . import distance, evalue, gapblock, pairwise
from .gapblock import (
    UNREACHABLE, central_prior, frame_prior, gap_cost, gapblock_score, positions_prior,
    profile_prior, score_matrix,
)

__all__ = [
    "distance",
    "evalue",
    "gapblock",
    "pairwise",
    "Alignment",
    "Index",
    "ScoreMatrix",
    "SearchParams",
    "SubstitutionMatrix",
    "alphabet_symbols",
    "amino_acids",
    "UNREACHABLE",
    "central_prior",
    "frame_prior",
    "gap_cost",
    "gapblock_score",
    "positions_prior",
    "profile_prior",
    "score_matrix",
    "device_count",
    "version",
]
