"""Gadget decomposition of the oracle, pinned against the reference.

- exhaustively against the reference's own decomposeRow (gadget_decomposition.h:
  189-236), compiled from its sources into oracle/_ref (order-dependent hash of
  every digit of every 32-bit input);
- exhaustively between the oracle's two restatements (scalar zeroMeanDigits
  form, gadget_decomposition.h:75-95, and level-major row form, :110-152);
- the GADGET_DECOMP.ZERO_MEAN invariants (decomposition_test.cpp:48-129).
"""
import ctypes

import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import P, lib, ref_digits_lib

THREADS = 8


@pytest.mark.parametrize("l,b", [(3, 7)])
def test_digits_match_reference_exhaustive_t32(l, b):
    ref = ref_digits_lib(32)
    if ref is None:
        pytest.skip("reference sources not built into oracle/_ref")
    assert ref.ref_torus_bits() == 32
    mine = lib().or_digits_hash(l, b, 32, 0, 1, 1 << 32, THREADS)
    theirs = ref.ref_digits_hash(l, b, 0, 1, 1 << 32, THREADS)
    assert mine == theirs


@pytest.mark.parametrize("l,b", [(8, 4), (2, 16), (4, 7), (3, 10), (1, 7)])
def test_digits_match_reference_sampled_t32(l, b):
    ref = ref_digits_lib(32)
    if ref is None:
        pytest.skip("reference sources not built into oracle/_ref")
    stride = (1 << 32) // (1 << 24) + 7
    assert lib().or_digits_hash(l, b, 32, 3, stride, 1 << 24, THREADS) == ref.ref_digits_hash(l, b, 3, stride, 1 << 24, THREADS)


@pytest.mark.parametrize("l,b", [(3, 10), (4, 14), (1, 14)])
def test_digits_match_reference_t56(l, b):
    ref = ref_digits_lib(56)
    if ref is None:
        pytest.skip("reference sources not built into oracle/_ref")
    assert ref.ref_torus_bits() == 56
    stride = (1 << 56) // (1 << 26) + 12345
    assert lib().or_digits_hash(l, b, 56, 11, stride, 1 << 26, THREADS) == ref.ref_digits_hash(l, b, 11, stride, 1 << 26, THREADS)


def test_ks_digits_match_reference():
    ref = ref_digits_lib(32)
    if ref is None:
        pytest.skip("reference sources not built into oracle/_ref")
    rng = np.random.default_rng(5)
    xs = list(rng.integers(-(1 << 31), 1 << 31, 4000)) + [0, -1, 1, -(1 << 31), (1 << 31) - 1, 1 << 15, -(1 << 15), 3 << 14]
    a = np.zeros(8, np.int64)
    bb = np.zeros(8, np.int64)
    for x in xs:
        lib().or_digits_scalar(int(x), 8, 2, 32, P(a))
        ref.ref_decompose_ks(int(x), 8, 2, 32, P(bb))
        assert np.array_equal(a, bb), x


@pytest.mark.parametrize("l,b,w", [(3, 7, 32), (8, 2, 32), (4, 8, 32)])
def test_two_oracle_forms_agree_exhaustive(l, b, w):
    assert lib().or_digits_cross_check(l, b, w, 0, 1, 1 << 32, THREADS) == 0


def test_two_oracle_forms_agree_t56():
    stride = (1 << 56) // (1 << 26) + 999
    assert lib().or_digits_cross_check(3, 10, 56, 5, stride, 1 << 26, THREADS) == 0


@pytest.mark.parametrize("l,b", [(3, 7), (2, 16), (4, 8), (7, 4), (3, 5), (3, 10)])
def test_zero_mean_invariants(l, b):
    """|d| <= B/2, recomposition error <= half the dropped range, d(-x) = -d(x)."""
    rng = np.random.default_rng(l * 100 + b)
    w = 32
    xs = rng.integers(-(1 << 31), 1 << 31, 20000).astype(np.int64)
    out = np.zeros((l, xs.size), np.int64)
    lib().or_digits_row(P(np.ascontiguousarray(xs)), xs.size, l, b, w, P(out))
    assert np.abs(out).max() <= (1 << (b - 1))
    rec = np.zeros(xs.size, dtype=object)
    for j in range(l):
        rec = rec + out[j].astype(object) * (1 << (w - (j + 1) * b))
    dropped = w - l * b
    half = (1 << (dropped - 1)) if dropped > 0 else 0
    for x, r in zip(xs.tolist(), rec.tolist()):
        d = (x - r) % (1 << w)
        d = d - (1 << w) if d >= (1 << (w - 1)) else d
        assert abs(d) <= half
    neg = np.zeros_like(out)
    mask = xs != -(1 << 31)
    lib().or_digits_row(P(np.ascontiguousarray(-xs)), xs.size, l, b, w, P(neg))
    assert np.array_equal(neg[:, mask], -out[:, mask])


def test_ks_digits_base8_match_reference_and_reach_plus_128():
    """At keyswitch base 2^8 (setKsRadixBits(8), legal in the reference) a
    balanced digit can be +128, which no signed byte holds: the oracle must
    agree with the reference's own decomposition there, and the engine, whose
    keyswitch GEMM takes signed-byte digits, must refuse the base."""
    ref = ref_digits_lib(32)
    xs = [(1 << 31) - 1, (1 << 31) - 2, 0x7F800000, 0x7FFFFFFF - (1 << 20), 0x00800000, -(1 << 31), -1, 12345]
    rng = np.random.default_rng(6)
    xs += list(rng.integers(-(1 << 31), 1 << 31, 2000))
    a = np.zeros(4, np.int64)
    bb = np.zeros(4, np.int64)
    seen = False
    for x in xs:
        lib().or_digits_scalar(int(x), 4, 8, 32, P(a))
        seen |= bool((a == 128).any())
        assert np.abs(a).max() <= 128
        if ref is not None:
            ref.ref_decompose_ks(int(x), 4, 8, 32, P(bb))
            assert np.array_equal(a, bb), x
    assert seen, "no input produced the digit +128"
    h = ctypes.c_void_p()
    cp = yg.replace(yg.PARAMS_T32_N1024, ks_radix_bits=8, ks_level=4).c()
    L = yg.load_library()
    assert L.ygpu_init(ctypes.byref(cp), 0, ctypes.byref(h)) == -2
    assert b"ks_radix_bits" in L.ygpu_last_error()
