"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on the CPU."""
from __future__ import annotations

import random
import sys
import zlib
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

AA = "ACDEFGHIKLMNPQRSTVWY"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    return oracle.Oracle()


@pytest.fixture(scope="session")
def reference():
    import oracle

    if not oracle.reference_available():
        pytest.skip("compiled reference (oracle/_ref) not built on this machine")
    return oracle.Reference()


@pytest.fixture(scope="session")
def st():
    """The product package; GPU tests require the device and fail loudly without it."""
    import seqtree_amd

    return seqtree_amd


@pytest.fixture(scope="session")
def gpu(st):
    if st.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return st


def seed_of(*key) -> int:
    """A seed fixed by the test's parameters (``hash`` of a string changes per process)."""
    return zlib.crc32(repr(key).encode()) & 0xFFFF


def rand_seq(rng: random.Random, lo: int, hi: int, alphabet: str = AA) -> str:
    return "".join(rng.choice(alphabet) for _ in range(rng.randint(lo, hi)))


def cdr3_like(rng: random.Random, n: int, lo: int, hi: int) -> list[str]:
    return ["C" + rand_seq(rng, lo - 2, hi - 2) + "F" for _ in range(n)]
