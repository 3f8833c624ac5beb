"""The engine's device digit decompositions, compiled for the host, against the
oracle's zeroMeanDigits restatement (gadget_decomposition.h:75-95), which
test_oracle_digits.py pins against the reference's own decomposeRow.

- digits_packed_add32 (the batched and latency TORUS32 kernels): the carry
  chain as one 32-bit addition; every 32-bit input for the 2^7 x 3 gadget,
  sampled inputs for other shapes;
- digits_packed_wide (the TORUS56 kernel): the bias added at full width and
  the top 32 bits through the same carry chain; sampled 56-bit inputs.

Equal order-dependent hashes of every digit of every input (the oracle's
or_digits_hash and the helper's eng_digits_hash mix identically).
"""
import ctypes
import os
import subprocess

import pytest

from tests.oracle_py import lib

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "engine_digits.cpp")
CSRC = os.path.join(os.path.dirname(HERE), "yatfhe_amd", "csrc")
OUT = os.path.join(HERE, "native", "_build", "libengdigits.so")
THREADS = 8
_eng = None


def eng():
    global _eng
    if _eng is None:
        deps = [SRC, os.path.join(CSRC, "ygpu_device.h")]
        if not os.path.exists(OUT) or any(os.path.getmtime(d) > os.path.getmtime(OUT) for d in deps):
            os.makedirs(os.path.dirname(OUT), exist_ok=True)
            subprocess.run([os.environ.get("HIPCC", "/opt/rocm/bin/hipcc"), "-O2", "-std=c++17", "-fPIC", "-shared",
                            f"-I{CSRC}", "-o", OUT + ".tmp", SRC], check=True, capture_output=True)
            os.replace(OUT + ".tmp", OUT)
        L = ctypes.CDLL(OUT)
        L.eng_digits_hash.restype = ctypes.c_uint64
        L.eng_digits_hash.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.c_uint64, ctypes.c_int32]
        _eng = L
    return _eng


def _both(l, b, w, lo, stride, count):
    mine = eng().eng_digits_hash(l, b, w, lo, stride, count, THREADS)
    assert mine != 0, "shape not instantiated in the helper"
    return mine, lib().or_digits_hash(l, b, w, lo, stride, count, THREADS)


def test_add32_exhaustive_2p7x3():
    mine, oracle = _both(3, 7, 32, 0, 1, 1 << 32)
    assert mine == oracle


@pytest.mark.parametrize("l,b", [(3, 10), (4, 7), (2, 8)])
def test_add32_sampled(l, b):
    stride = (1 << 32) // (1 << 24) + 7
    mine, oracle = _both(l, b, 32, 3, stride, 1 << 24)
    assert mine == oracle


@pytest.mark.parametrize("l,b", [(3, 10), (2, 14)])
def test_wide_sampled_t56(l, b):
    stride = (1 << 56) // (1 << 26) + 12345
    mine, oracle = _both(l, b, 56, 11, stride, 1 << 26)
    assert mine == oracle


def test_wide_edges_t56():
    # every top-30-bit pattern at the rounding edges of the dropped 26 bits
    # is covered by a stride of 2^26 - 1 from -2^55 (low parts cycle through
    # all residues); the extremes too
    for lo, stride, count in [((1 << 64) - (1 << 55), (1 << 26) - 1, 1 << 24),
                              ((1 << 25) - 2, 1 << 26, 1 << 24),
                              ((1 << 64) - (1 << 25) - 2, (1 << 64) - (1 << 26), 1 << 24)]:
        mine, oracle = _both(3, 10, 56, lo, stride, count)
        assert mine == oracle


def test_planted_error_is_seen():
    # a different shape must give a different hash (the comparison can fail)
    a = eng().eng_digits_hash(3, 10, 32, 3, 257, 1 << 16, THREADS)
    b = lib().or_digits_hash(3, 7, 32, 3, 257, 1 << 16, THREADS)
    assert a != b
