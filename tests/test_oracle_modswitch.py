"""Modulus switching (numeric.cpp:189-208), pinned by exact rational rounding:
round half away from zero of x * m / 2^32 (or 2^w), centred mod m."""
from fractions import Fraction

import numpy as np
import pytest

from tests.oracle_py import lib


def exact(x, new_mod, torus_q):
    f = Fraction(x * new_mod, torus_q)
    r = int(abs(f) + Fraction(1, 2)) * (1 if f >= 0 else -1)
    r %= new_mod
    return r - new_mod if r >= new_mod // 2 else r


@pytest.mark.parametrize("log2m", [11, 12])
def test_modswitch_to_2n(log2m):
    m = 1 << log2m
    scale = (1 << 32) // m
    rng = np.random.default_rng(log2m)
    xs = set(rng.integers(-(1 << 31), 1 << 31, 20000).tolist())
    for k in range(-m // 2 - 2, m // 2 + 2, 37):
        for d in (-1, 0, 1):
            xs.add(max(-(1 << 31), min((1 << 31) - 1, k * scale + scale // 2 + d)))
            xs.add(max(-(1 << 31), min((1 << 31) - 1, -k * scale - scale // 2 + d)))
    xs |= {-(1 << 31), (1 << 31) - 1, 0, -1, 1}
    for x in xs:
        assert lib().or_mod_switch_from_torus(x, m, 1 << 32) == exact(x, m, 1 << 32), x


def test_modswitch_exhaustive_closed_form():
    """All 2^32 inputs for 2N = 2048 against the rounding written as |x| + scale/2."""
    assert lib().or_modswitch_closed_form_check(11, 0, 1 << 32) == 0


def test_torus56_to_lwe32():
    rng = np.random.default_rng(3)
    xs = rng.integers(-(1 << 55), 1 << 55, 20000).tolist() + [-(1 << 55), (1 << 55) - 1, 1 << 55, (1 << 23), -(1 << 23), 3 << 23, -(3 << 23)]
    for x in xs:
        assert lib().or_mod_switch_from_torus(x, 1 << 32, 1 << 56) == exact(x, 1 << 32, 1 << 56), x
