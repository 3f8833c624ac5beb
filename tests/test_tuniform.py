"""The reference's TUniform noise law (genTUniformNoise, numeric.cpp:47-55) in
the seeded key generation and encryption.

TUniform(b) takes integers in [-2^b, 2^b], the two end points with probability
2^-(b+2) and every other point with probability 2^-(b+1) (the reference's own
doc comment, numeric.cpp:42-46), so its mean is 0 and its variance
(2^(2b+1) + 1) / 6.  These are checked on the noise itself: under an all-zero
secret key an encryption of 0 is (a, e).  Key rows are checked against the
hard bound |e| <= 2^b, which a Gaussian of the same scale would break."""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import glwe_phase_modq


def wrap(x, w):
    x = int(x) % (1 << w)
    return x - (1 << w) if x >= (1 << (w - 1)) else x


@pytest.mark.parametrize("b", [0, 1, 3, 18])
def test_tuniform_law(b):
    count = 400000
    zero = np.zeros(4, np.int32)
    e = yg.lwe_encrypt_tuniform(zero, np.zeros(count, np.int32), seed=11 + b, bound_bits=b)[:, -1].astype(np.int64)
    lo, hi = -(1 << b), 1 << b
    assert e.min() >= lo and e.max() <= hi
    var = ((1 << (2 * b + 1)) + 1) / 6
    assert abs(e.mean()) < 6 * np.sqrt(var / count)
    assert abs(e.var() / var - 1) < 0.02
    if b <= 3:
        # exact point probabilities: every count within 6 sigma of its expectation
        vals, cnt = np.unique(e, return_counts=True)
        assert list(vals) == list(range(lo, hi + 1))
        for v, c in zip(vals, cnt):
            pr = 2.0 ** -(b + 2) if v in (lo, hi) else 2.0 ** -(b + 1)
            assert abs(c - count * pr) < 6 * np.sqrt(count * pr * (1 - pr)), (v, c, count * pr)


def test_tuniform_keys_rows_within_bound(p32):
    """BSK rows decrypt (mod q) to their gadget message plus noise of at most
    2^rlweNoiseB; KSK rows to theirs plus at most 2^lweNoiseB; the secret keys
    are the Gaussian keygen's for the same seed."""
    lwe_b, glwe_b = 18, 3   # the reference's TORUS32 lweNoiseB / rlweNoiseB (yatfhe_parameters.h:31-44)
    keys = yg.keygen_tuniform(p32, seed=5, lwe_bound_bits=lwe_b, glwe_bound_bits=glwe_b)
    g = yg.keygen(p32, seed=5)
    assert np.array_equal(keys.lwe_key, g.lwe_key) and np.array_equal(keys.glwe_key, g.glwe_key)
    q, N, w = p32.q_ntt, p32.N, 32
    worst = 0
    for i in (0, 100, p32.n - 1):
        s = int(keys.lwe_key[i])
        for lvl in range(p32.level):
            mu = s << (w - (lvl + 1) * p32.radix_bits)
            # row k (cPrime): the body carries mu at coefficient 0
            ph = glwe_phase_modq(q, keys.bsk[i, lvl, p32.k], keys.glwe_key)
            want = np.zeros(N, np.int64)
            want[0] = mu
            err = ph - want
            assert np.abs(err).max() <= (1 << glwe_b)
            worst = max(worst, int(np.abs(err).max()))
    assert worst > 0
    errs = []
    for i in (0, 7, 1023):
        for j in range(p32.ks_level):
            row = keys.ksk[i, j].astype(np.int64)
            ph = wrap(row[-1] - int((row[:-1] * keys.lwe_key).sum()), 32)
            m = int(keys.glwe_key[i]) << (32 - (j + 1) * p32.ks_radix_bits)
            errs.append(wrap(ph - m, 32))
    assert max(abs(x) for x in errs) <= (1 << lwe_b)


@pytest.mark.parametrize("form", ["classic", "succinct"])
def test_tuniform_other_key_forms(p32, form):
    keys = yg.keygen_tuniform(p32, seed=6, lwe_bound_bits=18, glwe_bound_bits=3, form=form, s2_level=4)
    g = yg.keygen(p32, seed=6)
    assert np.array_equal(keys.glwe_key, g.glwe_key)
    q = p32.q_ntt
    if form == "classic":
        # row k of a classic key: body = torus value of the mask product + mu + e
        s = int(keys.lwe_key[3])
        ph = glwe_phase_modq(q, keys.bsk[3, 0, p32.k], keys.glwe_key)
        ph = np.array([wrap(x, 32) for x in ph])
        ph[0] = wrap(ph[0] - (s << (32 - p32.radix_bits)), 32)
        assert np.abs(ph).max() <= 8 + 1
    else:
        ph = glwe_phase_modq(q, keys.bsk_half[3, 0], keys.glwe_key)
        ph = np.array([wrap(x, 32) for x in ph])
        ph[0] = wrap(ph[0] - (int(keys.lwe_key[3]) << (32 - p32.radix_bits)), 32)
        assert np.abs(ph).max() <= 8


def test_tuniform_rejects_bad_bounds(p32):
    with pytest.raises(yg.YgpuError):
        yg.keygen_tuniform(p32, seed=1, lwe_bound_bits=31, glwe_bound_bits=3)
