"""The oracle's transform route against brute force.

The reference computes external products with Intel HEXL's NTT (ntt_hexl.h:35-44,
ntt_hexl.cpp:153-190); the result it depends on is the negacyclic product mod q,
which is unique, so the oracle's NTT is pinned by an O(N^2) schoolbook product
(the HEXL.POLY_MULT equivalence of hexl_test.cpp:42-102) and the whole external
product by the schoolbook route, bit for bit.
"""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle, P, lib

PRIMES = [yg.Q_49P_T32, yg.Q_61P_T56, 562949952700417]


@pytest.mark.parametrize("q", PRIMES)
@pytest.mark.parametrize("N", [8, 64, 1024])
def test_negacyclic_product_matches_schoolbook(q, N):
    rng = np.random.default_rng(N)
    a = rng.integers(-512, 513, N).astype(np.int64)
    b = rng.integers(0, q, N, dtype=np.uint64)
    x = np.zeros(N, np.uint64)
    y = np.zeros(N, np.uint64)
    lib().or_negacyclic_mul_modq(q, N, P(a), P(b), P(x), 0)
    lib().or_negacyclic_mul_modq(q, N, P(a), P(b), P(y), 1)
    assert np.array_equal(x, y)


def test_schoolbook_small_by_hand():
    # (1 + X) * (1 + X) = 1 + 2X + X^2 ;  X^3 * X = X^4 = -1 in Z_q[X]/(X^4+1)
    q = yg.Q_49P_T32
    out = np.zeros(4, np.uint64)
    lib().or_negacyclic_mul_modq(q, 4, P(np.array([1, 1, 0, 0], np.int64)), P(np.array([1, 1, 0, 0], np.uint64)), P(out), 1)
    assert out.tolist() == [1, 2, 1, 0]
    lib().or_negacyclic_mul_modq(q, 4, P(np.array([0, 0, 0, 1], np.int64)), P(np.array([0, 1, 0, 0], np.uint64)), P(out), 1)
    assert out.tolist() == [q - 1, 0, 0, 0]


@pytest.mark.parametrize("q", PRIMES)
def test_ntt_roundtrip(q):
    rng = np.random.default_rng(7)
    a = rng.integers(0, q, 2048, dtype=np.uint64)
    assert lib().or_ntt_roundtrip(q, 2048, P(a)) == 1


def test_intt_exit_centring():
    q = yg.Q_49P_T32
    h = (q + 1) // 2
    assert lib().or_intt_exit(0, q, 32) == 0
    assert lib().or_intt_exit(h - 1, q, 32) == np.int64(np.int32(np.uint32((h - 1) & 0xFFFFFFFF)))
    assert lib().or_intt_exit(h, q, 32) == np.int64(np.int32(np.uint32((h - q) & 0xFFFFFFFF)))
    assert lib().or_intt_exit(q - 1, q, 32) == -1
    q56 = yg.Q_61P_T56
    assert lib().or_intt_exit(q56 - 5, q56, 56) == -5


@pytest.mark.parametrize("which", ["t32", "t56"])
def test_external_product_ntt_equals_schoolbook(which, p32, keys32, p56, keys56):
    p, keys = (p32, keys32) if which == "t32" else (p56, keys56)
    orc = Oracle(p, keys)
    rng = np.random.default_rng(3)
    w = p.torus_bits
    acc = rng.integers(-(1 << (w - 1)), 1 << (w - 1), (p.k + 1, p.N), dtype=np.int64)
    acc[0, 0] = -(1 << (w - 1))
    for i in (0, p.n - 1):
        a = orc.external_product(i, acc, schoolbook=False)
        b = orc.external_product(i, acc, schoolbook=True)
        assert np.array_equal(a, b)
    orc.close()
