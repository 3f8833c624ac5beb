"""Monomial rotation, pinned by the reference's POLY.ROT known-answer table
(polynomial_rotation_test.cpp:27-65) and by an explicit monomial product."""
import numpy as np
import pytest

from tests.oracle_py import P, lib

# the reference's table for a = (1,2,3,4), N = 4, indexed by r + 2N
M = {
    "zo": [1, 2, 3, 4], "m7": [-4, 1, 2, 3], "m6": [-3, -4, 1, 2], "m5": [-2, -3, -4, 1], "m4": [-1, -2, -3, -4],
    "m3": [4, -1, -2, -3], "m2": [3, 4, -1, -2], "m1": [2, 3, 4, -1],
}
TABLE = {0: "zo", 8: "zo", 16: "zo", 1: "m7", 9: "m7", 2: "m6", 10: "m6", 3: "m5", 11: "m5", 4: "m4", 12: "m4",
         5: "m3", 13: "m3", 6: "m2", 14: "m2", 7: "m1", 15: "m1"}


@pytest.mark.parametrize("w", [32, 56])
def test_poly_rot_kat(w):
    N = 4
    a = np.array([1, 2, 3, 4], np.int64)
    out = np.zeros(N, np.int64)
    for r in range(-2 * N, 2 * N + 1):
        lib().or_rotate(w, N, r, P(a), P(out))
        assert out.tolist() == M[TABLE[r + 2 * N]], r
        lib().or_rotate_minus_one(w, N, r, P(a), P(out))
        assert out.tolist() == (np.array(M[TABLE[r + 2 * N]]) - a).tolist(), r


def monomial_product(a, r, N, w):
    """X^r * a in Z_{2^w}[X]/(X^N+1) by definition (coefficient-wise placement)."""
    out = [0] * N
    for i, c in enumerate(a):
        e = (i + r) % (2 * N)
        if e < N:
            out[e] += c
        else:
            out[e - N] -= c
    return [((x + (1 << (w - 1))) % (1 << w)) - (1 << (w - 1)) for x in out]


@pytest.mark.parametrize("w,N", [(32, 1024), (56, 2048), (32, 16)])
def test_rotation_matches_monomial_product(w, N):
    rng = np.random.default_rng(N + w)
    a = rng.integers(-(1 << (w - 1)), 1 << (w - 1), N, dtype=np.int64)
    a[0] = -(1 << (w - 1))  # the most negative torus value wraps on negation
    out = np.zeros(N, np.int64)
    for r in [0, 1, -1, N - 1, N, N + 1, 2 * N - 1, -N, 3 * N + 5, int(rng.integers(-4 * N, 4 * N))]:
        lib().or_rotate(w, N, r, P(a), P(out))
        assert out.tolist() == monomial_product(a.tolist(), r, N, w), r
        lib().or_rotate_minus_one(w, N, r, P(a), P(out))
        want = [((x - y + (1 << (w - 1))) % (1 << w)) - (1 << (w - 1)) for x, y in zip(monomial_product(a.tolist(), r, N, w), a.tolist())]
        assert out.tolist() == want, r
