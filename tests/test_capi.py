"""The C ABI library without a GPU: it loads, exports every entry point the
header declares, fails cleanly (no crash) when no device is present, and its
host-side key generation is deterministic."""
import ctypes
import os
import re

import numpy as np
import pytest

import yatfhe_amd as yg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "yatfhe_gpu.h")).read()
    return sorted(set(re.findall(r"\b(ygpu_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(yg.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(yg.exported_symbols())


def test_struct_mirrors_match_the_library():
    # every ctypes mirror has the size the library was compiled with
    assert yg.abi_struct_mismatches() == []


def test_no_device_is_an_error_not_a_crash():
    if yg.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = yg.load_library()
    h = ctypes.c_void_p()
    cp = yg.PARAMS_T32_N1024.c()
    rc = lib.ygpu_init(ctypes.byref(cp), 0, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert lib.ygpu_last_error()


def test_parameter_validation_messages():
    lib = yg.load_library()
    h = ctypes.c_void_p()
    bad = yg.replace(yg.PARAMS_T32_N1024, level=5, radix_bits=7).c()
    assert lib.ygpu_init(ctypes.byref(bad), 0, ctypes.byref(h)) != 0
    bad = yg.replace(yg.PARAMS_T32_N1024, level=4, radix_bits=9).c()   # 36 bits > 32-bit torus
    assert lib.ygpu_init(ctypes.byref(bad), 0, ctypes.byref(h)) != 0
    assert b"do not fit" in lib.ygpu_last_error() or yg.device_count() == 0
    # a 2^16 gadget has the digit +2^15, which a 16-bit digit field cannot hold
    bad = yg.replace(yg.PARAMS_T56_N2048, level=3, radix_bits=16).c()
    assert lib.ygpu_init(ctypes.byref(bad), 0, ctypes.byref(h)) != 0
    assert b"radix_bits" in lib.ygpu_last_error() or yg.device_count() == 0


def test_keygen_deterministic_and_seed_sensitive(p32):
    a = yg.keygen(p32, seed=7, threads=3)
    b = yg.keygen(p32, seed=7, threads=8)
    c = yg.keygen(p32, seed=8)
    assert np.array_equal(a.bsk, b.bsk) and np.array_equal(a.ksk, b.ksk)
    assert not np.array_equal(a.bsk, c.bsk)
    assert int(a.bsk.max()) < p32.q_ntt
