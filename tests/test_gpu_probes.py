"""The measurement probes of the C-ABI (issue rates and bandwidth the rooflines are quoted
against): each returns a positive rate on the device and rejects arguments it cannot take."""
from __future__ import annotations

import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from seqtree_amd import _device

    L = _device.lib()
    dbl = ctypes.POINTER(ctypes.c_double)
    L.st_probe_valu.argtypes = [ctypes.c_int32, ctypes.c_char_p, ctypes.c_uint64, dbl,
                                ctypes.c_void_p]
    L.st_probe_mfma.argtypes = [dbl, ctypes.c_void_p]
    L.st_probe_hbm_write.argtypes = [ctypes.c_uint64, dbl, ctypes.c_void_p]
    L.st_probe_lds.argtypes = [ctypes.c_int32, dbl, ctypes.c_void_p]
    return L, _device.Status


def test_valu_probe_kinds(gpu):
    L, Status = _lib()
    names = []
    for kind in range(6):
        name = ctypes.create_string_buffer(32)
        rate, st = ctypes.c_double(), Status()
        assert L.st_probe_valu(kind, name, 32, ctypes.byref(rate), ctypes.byref(st)) == 0
        assert rate.value > 1e12
        names.append(name.value.decode())
    assert names[0] == "v_add_u32" and len(set(names)) == 6
    rate, st = ctypes.c_double(), Status()
    assert L.st_probe_valu(99, ctypes.create_string_buffer(32), 32, ctypes.byref(rate),
                           ctypes.byref(st)) == 1


def test_mfma_and_write_probes(gpu):
    L, Status = _lib()
    rate, st = ctypes.c_double(), Status()
    assert L.st_probe_mfma(ctypes.byref(rate), ctypes.byref(st)) == 0
    assert rate.value > 1e14
    assert L.st_probe_hbm_write(256 << 20, ctypes.byref(rate), ctypes.byref(st)) == 0
    assert 1e11 < rate.value < 2e13
    assert L.st_probe_hbm_write(4096, ctypes.byref(rate), ctypes.byref(st)) == 1
    assert b"at least 1 MiB" in st.message


def test_lds_probe(gpu):
    L, Status = _lib()
    rates = []
    for dwords in (1, 2, 4):
        rate, st = ctypes.c_double(), Status()
        assert L.st_probe_lds(dwords, ctypes.byref(rate), ctypes.byref(st)) == 0, st.message
        # above the HBM rate, below 256 CUs x 512 bytes per clock at 3 GHz
        assert 1e13 < rate.value < 4e14
        rates.append(rate.value)
    rate, st = ctypes.c_double(), Status()
    assert L.st_probe_lds(3, ctypes.byref(rate), ctypes.byref(st)) == 1
    assert b"1, 2 or 4" in st.message
