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


def launch_labels(fn):
    """Runs fn with the library's launch timing on; returns (fn's result, the set of launch
    labels recorded meanwhile), which names the kernels that ran."""
    import ctypes

    from seqtree_amd import _device

    L = _device.lib()
    L.st_timing_enable.argtypes = [ctypes.c_int32]
    L.st_timing_collect.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
    L.st_timing_collect.restype = ctypes.c_int32
    L.st_timing_enable(1)
    try:
        out = fn()
    finally:
        buf = ctypes.create_string_buffer(1 << 16)
        n = L.st_timing_collect(buf, len(buf))
        L.st_timing_enable(0)
    assert n >= 0
    return out, {line.split()[0] for line in buf.value.decode().splitlines()}
