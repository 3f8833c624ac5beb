"""WWL+24 succinct keys on the CPU: key generation, the oracle's derivation of
the full RGSW rows, and blind rotation with the derived key.

The reference forms the succinct key in genBootstrappingKeyWWL24
(bootstrapping.cpp:106-126) and derives the mask-digit rows at every step of
blindRotateWWL24Ntt (blind_rotate.cpp:1383-1412) with
switchTrlweToSecretEmbeddingNtt (trgsw.cpp:861-909).  The oracle's derivation is
pinned here by (1) a schoolbook route that shares no transform code with the
NTT route, bit for bit; (2) what the mathematics fixes for every derived row:
its torus phase is -s_i 2^(w-(lvl+1)b) z, the phase of the reference's own
mask-digit RGSW row, within the noise bound; (3) the gate of
BLIND_ROT.BLIND_ROT_WWL24_NTT (blindrot_test.cpp:279-333): blind rotation with
the derived key decrypts to X^{sum abar_i s_i} v.
"""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import (Oracle, P, derive_succinct, derive_succinct_row, glwe_phase, lib)

# The reference's TORUS32 defaults (yatfhe_parameters.h:31-47): N=2048, gadget
# 2^4 with l=8 (the s^2 digits) and lApprox=4 (the blind-rotation gadget),
# keyswitch 2^2 x 16; noise log2 sigma 2.21 (GLWE) and 17.21 (LWE) in torus units.
P_REF32 = yg.Params(torus_bits=32, n=680, k=1, N=2048, radix_bits=4, level=4, ks_radix_bits=2, ks_level=16,
                    q_ntt=yg.Q_49P_T32)
SIG_REF32 = (17.21 - 32, 2.21 - 32)


def wrap(x, w):
    m = 1 << w
    x = np.asarray(x, dtype=object) % m
    return np.array([int(v) - m if v >= (m >> 1) else int(v) for v in x], dtype=np.int64)


def torus_of(params, residues):
    """applyIntt's exit of coefficient residues: centre mod q, low w bits."""
    q = params.q_ntt
    return np.array([lib().or_intt_exit(int(r), q, params.torus_bits) for r in np.asarray(residues).reshape(-1)],
                    dtype=np.int64).reshape(np.asarray(residues).shape)


def negacyclic(a, b, N):
    """Exact integer negacyclic product of two integer polynomials."""
    full = np.convolve(np.asarray(a, dtype=object), np.asarray(b, dtype=object))
    out = np.array(full[:N], dtype=object)
    out[: len(full) - N] -= full[N:]
    return out


@pytest.fixture(scope="module")
def sk32(p32):
    return yg.keygen_succinct(p32, seed=11, s2_level=4)


@pytest.fixture(scope="module")
def skref():
    return yg.keygen_succinct(yg.replace(P_REF32, n=48), seed=12, sigma_log2_lwe=SIG_REF32[0],
                              sigma_log2_glwe=SIG_REF32[1])


def test_keygen_shapes_and_shared_secrets(p32, sk32):
    full = yg.keygen(p32, seed=11)
    assert sk32.bsk_half.shape == (p32.n, p32.level, 2, p32.N)
    assert sk32.s2.shape == (4, 2, p32.N)
    assert np.array_equal(full.lwe_key, sk32.lwe_key) and np.array_equal(full.glwe_key, sk32.glwe_key)
    assert np.array_equal(full.ksk, sk32.ksk)
    assert (sk32.bsk_half < p32.q_ntt).all() and (sk32.s2 < p32.q_ntt).all()
    # the cPrime masks are torus-uniform (symEncTrlweMultiSampleNtt), lifted mod q
    a = sk32.bsk_half[:, :, 0].reshape(-1)
    half_t = 1 << (p32.torus_bits - 1)
    assert ((a < half_t) | (a >= p32.q_ntt - half_t)).all()
    assert yg.max_s2_level(p32) == 4
    with pytest.raises(yg.YgpuError):
        yg.keygen_succinct(p32, seed=0, s2_level=5)  # 5 digits of 7 bits exceed the 32-bit torus


def test_keygen_rows_decrypt(p32, sk32):
    """cPrime[lvl] of step i decrypts to s_i 2^(w-(lvl+1)b) at coefficient 0;
    s2[l2] decrypts to s^2 2^(w-(l2+1)b) (the TRLEV of symEncTrlevWithKeyNtt)."""
    w, N, rb = p32.torus_bits, p32.N, p32.radix_bits
    z = sk32.glwe_key
    for i, lvl in [(0, 0), (1, 2), (p32.n - 1, 1), (5, 0)]:
        row = torus_of(p32, sk32.bsk_half[i, lvl])
        ph = glwe_phase(p32, row.reshape(-1), z)
        want = np.zeros(N, np.int64)
        want[0] = wrap([int(sk32.lwe_key[i]) << (w - (lvl + 1) * rb)], w)[0]
        assert np.abs(wrap(ph - want, w)).max() < (1 << 12)
    zz = negacyclic(z, z, N)
    for l2 in range(sk32.s2_level):
        row = torus_of(p32, sk32.s2[l2])
        ph = glwe_phase(p32, row.reshape(-1), z)
        want = wrap(zz * (1 << (w - (l2 + 1) * rb)), w)
        assert np.abs(wrap(ph - want, w)).max() < (1 << 12)


@pytest.mark.parametrize("which", ["t32", "t56"])
def test_derivation_ntt_route_equals_schoolbook(which, p32, p56, sk32):
    if which == "t32":
        p, sk, rows = p32, sk32, [(0, 0), (3, 2), (p32.n - 1, 1)]
    else:
        p = yg.replace(p56, n=4)
        sk = yg.keygen_succinct(p, seed=13)
        rows = [(2, 1)]
    for i, lvl in rows:
        a = derive_succinct_row(p, sk, i, lvl, schoolbook=False)
        b = derive_succinct_row(p, sk, i, lvl, schoolbook=True)
        assert np.array_equal(a, b), (which, i, lvl)
        assert np.array_equal(a[1], sk.bsk_half[i, lvl])  # the body-digit row is cPrime itself


def test_derive_all_rows_matches_per_row(p32, sk32):
    full = derive_succinct(p32, sk32)
    assert full.shape == (p32.n, p32.level, 2, 2, p32.N)
    for i, lvl in [(0, 0), (17, 1), (p32.n - 1, 2)]:
        assert np.array_equal(full[i, lvl], derive_succinct_row(p32, sk32, i, lvl))


def test_derived_rows_are_mask_digit_rgsw_rows(skref):
    """Torus phase of derived row c[lvl][0] of step i: -s_i 2^(w-(lvl+1)b) z,
    within the derivation noise (s^2 digits exact at l=8, so the bound is
    sum_l2 |D| e + e z, a few 2^12 at the reference's default noise)."""
    p = skref.params
    w, rb = p.torus_bits, p.radix_bits
    full = derive_succinct(p, skref)
    z = skref.glwe_key.astype(np.int64)
    worst = 0
    for i in range(p.n):
        for lvl in range(p.level):
            ph = glwe_phase(p, torus_of(p, full[i, lvl, 0]).reshape(-1), skref.glwe_key)
            want = wrap(-int(skref.lwe_key[i]) * (1 << (w - (lvl + 1) * rb)) * z, w)
            worst = max(worst, int(np.abs(wrap(ph - want, w)).max()))
    # smallest gadget value 2^(32-16) = 2^16: a sign or level slip exceeds the bound
    assert worst < (1 << 15), worst


def test_blind_rotation_with_derived_key_decrypts():
    """BLIND_ROT.BLIND_ROT_WWL24_NTT (blindrot_test.cpp:279-333) at the reference's
    TORUS32 defaults: decrypt(acc) = X^{-bbar + sum abar_i s_i} v."""
    p = P_REF32
    sk = yg.keygen_succinct(p, seed=14, sigma_log2_lwe=SIG_REF32[0], sigma_log2_glwe=SIG_REF32[1])
    full = derive_succinct(p, sk)
    keys = yg.KeySet(p, sk.lwe_key, sk.glwe_key, full, sk.ksk)
    orc = Oracle(p, keys)
    rng = np.random.default_rng(3)
    N = p.N
    v = (rng.integers(0, 8, N) * (1 << 29)).astype(np.int32)
    ct = yg.lwe_encrypt(sk.lwe_key, np.array([5 << 28], np.int32), seed=6, sigma_log2=SIG_REF32[0])[0]
    sc = orc.rescale(ct)
    rot = (-int(sc[-1]) + int((sc[:-1].astype(np.int64) * sk.lwe_key).sum())) % (2 * N)
    acc = orc.blind_rotate(ct, v)
    ph = glwe_phase(p, acc, sk.glwe_key)
    want = np.zeros(N, np.int64)
    lib().or_rotate(32, N, rot, P(v.astype(np.int64)), P(want))
    err = wrap(ph - want, 32)
    assert np.abs(err).max() < (1 << 27)
    orc.close()


def test_keygen_succinct_deterministic(p32):
    p = yg.replace(p32, n=20)
    a = yg.keygen_succinct(p, seed=7, threads=2)
    b = yg.keygen_succinct(p, seed=7, threads=8)
    c = yg.keygen_succinct(p, seed=8)
    assert np.array_equal(a.bsk_half, b.bsk_half) and np.array_equal(a.s2, b.s2)
    assert not np.array_equal(a.bsk_half, c.bsk_half)
