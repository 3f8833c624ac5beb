"""Semantic gates of the reference's tests, on the oracle with seeded keys:
TRLWE.NOISELESS_SAMPLE_CLEARS_MASK (trlwe_test.cpp:579-621), BLIND_ROT_MP21_NTT
(blindrot_test.cpp:218-276), extraction (trlwe.cpp:241-250), KSK.KSK
(keyswitch_test.cpp:12-40), BOOTSTRAPPING.MP21_FR (bootstrappping_test.cpp:133-180).
"""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle, glwe_phase, lib, P


def wrap(x, w):
    x = int(x) % (1 << w)
    return x - (1 << w) if x >= (1 << (w - 1)) else x


def test_noiseless_sample_clears_mask(p32, keys32):
    orc = Oracle(p32, keys32)
    rng = np.random.default_rng(0)
    v = rng.integers(-(1 << 31), 1 << 31, p32.N).astype(np.int32)
    N = p32.N
    for b in [0, 1, N - 1, N, N + 3, 2 * N - 1, -1, -N]:
        acc = orc.init_acc(v, b)
        assert not acc[:N].any()
        want = np.zeros(N, np.int64)
        lib().or_rotate(32, N, -b, P(v.astype(np.int64)), P(want))
        assert np.array_equal(glwe_phase(p32, acc, keys32.glwe_key), want)
    orc.close()


def test_blind_rotation_phase(p32, keys32):
    """decrypt(acc) = X^{-bbar + sum abar_i s_i} v, up to the external-product noise."""
    orc = Oracle(p32, keys32)
    rng = np.random.default_rng(1)
    N = p32.N
    v = (rng.integers(0, 8, N) * (1 << 29)).astype(np.int32)
    mu = np.array([3 << 29], np.int32)
    ct = yg.lwe_encrypt(keys32.lwe_key, mu, seed=5)[0]
    sc = orc.rescale(ct)
    rot = (-int(sc[-1]) + int((sc[:-1].astype(np.int64) * keys32.lwe_key).sum())) % (2 * N)
    acc = orc.blind_rotate(ct, v)
    ph = glwe_phase(p32, acc, keys32.glwe_key)
    want = np.zeros(N, np.int64)
    lib().or_rotate(32, N, rot, P(v.astype(np.int64)), P(want))
    err = np.array([wrap(a - b, 32) for a, b in zip(ph, want)])
    assert np.abs(err).max() < (1 << 26)
    orc.close()


def test_extract_is_coefficient_zero_phase(p32, keys32):
    orc = Oracle(p32, keys32)
    rng = np.random.default_rng(2)
    acc = rng.integers(-(1 << 31), 1 << 31, (2, p32.N)).astype(np.int64)
    acc[0, 5] = -(1 << 31)
    ext = orc.extract(acc)
    gp = glwe_phase(p32, acc, keys32.glwe_key)
    assert yg.lwe_decrypt_phase(ext, keys32.glwe_key)[0] == np.int32(gp[0])
    orc.close()


@pytest.mark.parametrize("index", [0, 1, 37, 511, 1023])
def test_extract_at_index_is_coefficient_phase(p32, keys32, index):
    """extractTlweFromTrlwe(out, in, index) (trlwe.cpp:241-250): the extracted
    LWE decrypts under the flattened GLWE key to coefficient `index` of the
    GLWE phase, exactly at 32 bits (an index or sign slip changes it by a
    torus-uniform amount)."""
    from tests.oracle_py import or_params
    import ctypes
    rng = np.random.default_rng(40 + index)
    acc = rng.integers(-(1 << 31), 1 << 31, (2, p32.N)).astype(np.int64)
    acc[0, (index + 3) % p32.N] = -(1 << 31)   # the wrapping negation
    cp = or_params(p32, index)
    ext = np.zeros(p32.N + 1, np.int32)
    lib().or_extract(ctypes.byref(cp), P(np.ascontiguousarray(acc.reshape(-1))), P(ext))
    gp = glwe_phase(p32, acc, keys32.glwe_key)
    assert yg.lwe_decrypt_phase(ext, keys32.glwe_key)[0] == np.int32(gp[index])
    if index == 0:
        orc = Oracle(p32, keys32)
        assert np.array_equal(ext, orc.extract(acc))
        orc.close()


def test_extract_at_index_t56(p56, keys56):
    """At 56 bits each extracted coefficient is modswitched to 2^32 (round half
    away from zero), so the phase matches coefficient `index` of the GLWE phase
    up to the rounding of kN + 1 terms."""
    from tests.oracle_py import or_params
    import ctypes
    rng = np.random.default_rng(9)
    N = p56.N
    acc = rng.integers(-(1 << 55), 1 << 55, (2, N)).astype(np.int64)
    for index in (0, 5, N - 1):
        cp = or_params(p56, index)
        ext = np.zeros(N + 1, np.int32)
        lib().or_extract(ctypes.byref(cp), P(np.ascontiguousarray(acc.reshape(-1))), P(ext))
        gp = int(glwe_phase(p56, acc, keys56.glwe_key)[index])
        want = wrap(round(gp / (1 << 24)), 32)
        got = int(yg.lwe_decrypt_phase(ext, keys56.glwe_key)[0])
        assert abs(wrap(got - want, 32)) <= (N + 1) // 2 + 1


def test_keyswitch_every_message(p32, keys32):
    orc = Oracle(p32, keys32)
    for m in range(-4, 4):
        mu = np.array([m * (1 << 29)], np.int32)
        ct = yg.lwe_encrypt(keys32.glwe_key, mu, seed=100 + m)[0]
        out = orc.keyswitch(ct)
        assert np.array_equal(out, orc.keyswitch(ct, u32=True))  # R11: int64 == uint32 wrap
        ph = int(yg.lwe_decrypt_phase(out, keys32.lwe_key)[0])
        assert round(ph / (1 << 29)) % 8 == m % 8
    orc.close()


def test_bsk_rows_decrypt(p32, keys32):
    """Every RGSW row of the seeded BSK decrypts (mod q) to its gadget message."""
    q = p32.q_ntt
    N = p32.N
    out = np.zeros(N, np.int64)
    for i in (0, 7, p32.n - 1):
        s = int(keys32.lwe_key[i])
        for lvl in range(p32.level):
            mu = s << (32 - (lvl + 1) * p32.radix_bits)
            for row in range(2):
                r = np.ascontiguousarray(keys32.bsk[i, lvl, row])
                lib().or_glwe_phase_modq(q, 1, N, P(r), P(np.ascontiguousarray(keys32.glwe_key)), P(out))
                want = np.zeros(N, np.int64)
                if row == 1:
                    want[0] = mu
                else:
                    # mask row: phase = e - mu * z (the message sits in a_0)
                    z = keys32.glwe_key.astype(np.int64)
                    want = -mu * z
                assert np.abs(out - want).max() < (1 << 12), (i, lvl, row)


def test_pbs_full_range_lut_t32(p32, keys32):
    """MP21_FR pattern: every slot of a random 8-entry LUT decrypts to the table."""
    p = 8
    rng = np.random.default_rng(9)
    lut = rng.integers(0, p, p)
    v = yg.lut_test_polynomial(p32, lut, p)
    orc = Oracle(p32, keys32)
    msgs = np.arange(p)
    cts = yg.lwe_encrypt(keys32.lwe_key, (msgs * (1 << 32) // (2 * p)).astype(np.int64).astype(np.int32), seed=77)
    outs = orc.pbs_batch(cts, v)
    ph = yg.lwe_decrypt_phase(outs, keys32.lwe_key).astype(np.int64)
    dec = np.round(ph / (1 << 32) * p).astype(np.int64) % p
    assert np.array_equal(dec, lut[msgs] % p)
    orc.close()


def test_pbs_lut_t56(p56, keys56):
    """4-bit programmable bootstrap at TORUS56 / N = 2048 decrypts to the LUT."""
    p = 16
    rng = np.random.default_rng(2)
    lut = rng.integers(0, p, p)
    v = yg.lut_test_polynomial(p56, lut, p)
    orc = Oracle(p56, keys56)
    msgs = np.array([0, 1, 7, 8, 15, 3])
    cts = yg.lwe_encrypt(keys56.lwe_key, (msgs * (1 << 32) // (2 * p)).astype(np.int64).astype(np.int32), seed=3)
    outs = orc.pbs_batch(cts, v)
    ph = yg.lwe_decrypt_phase(outs, keys56.lwe_key).astype(np.int64)
    dec = np.round(ph / (1 << 32) * p).astype(np.int64) % p
    assert np.array_equal(dec, lut[msgs] % p)
    orc.close()


def test_level_parallel_blind_rotation_is_identical(p32, keys32):
    """The reference's threading of the external product (levels on a pool,
    trgsw.cpp:725-760) changes nothing in the result."""
    orc = Oracle(p32, keys32)
    v = yg.sign_test_polynomial(p32)
    cts = yg.lwe_encrypt(keys32.lwe_key, np.array([1 << 29, -(1 << 29), 7], np.int32), seed=12)
    for ct in cts:
        _, _, want = orc.pbs_batch_all(ct[None, :], v)
        got, secs = orc.pbs_level_parallel(ct, v, reps=2)
        assert np.array_equal(got, want[0]) and (secs > 0).all()
    orc.close()
