"""Classic bootstrapping keys (genBootstrappingKey -> encryptTrgswNtt,
bootstrapping.cpp:40-44, 98-104; trgsw.cpp:185-191, 227-246, 281-284) on the CPU.

encryptTrgswNtt adds mu_lvl = s_i 2^(w-(lvl+1)b) in the torus: to a_r[0] of
row r < k after the body was formed (so the row's phase is -mu z_r), and to the
torus value of row k's body (so its residues are lifted torus values and its
phase is +mu at coefficient 0).  Pinned here by those phases, by the body's
form, and by the bootstrap through functionalBootstrappingNtt's arithmetic
(the MP21 external product with l = lApprox) decrypting every sign."""
import numpy as np

import yatfhe_amd as yg
from tests.oracle_py import Oracle, glwe_phase, lib


def wrap(x, w):
    m = 1 << w
    x = np.asarray(x, dtype=object) % m
    return np.array([int(v) - m if v >= (m >> 1) else int(v) for v in x], dtype=np.int64)


def torus_of(params, residues):
    q = params.q_ntt
    return np.array([lib().or_intt_exit(int(r), q, params.torus_bits) for r in np.asarray(residues).reshape(-1)],
                    dtype=np.int64).reshape(np.asarray(residues).shape)


def test_classic_rows(p32):
    p = yg.replace(p32, n=12)
    keys = yg.keygen_classic(p, seed=41)
    full = yg.keygen(p, seed=41)
    assert np.array_equal(keys.lwe_key, full.lwe_key) and np.array_equal(keys.ksk, full.ksk)
    w, N, rb, q = p.torus_bits, p.N, p.radix_bits, p.q_ntt
    half_t = 1 << (w - 1)
    z = keys.glwe_key.astype(np.int64)
    for i in range(p.n):
        for lvl in range(p.level):
            mu = int(wrap([int(keys.lwe_key[i]) << (w - (lvl + 1) * rb)], w)[0])
            rows = keys.bsk[i, lvl]
            # masks: lifted torus values
            for row in range(2):
                a = rows[row, 0]
                assert ((a < half_t) | (a >= q - half_t)).all()
            # row k's body: lifted torus values too (reduced before mu was added)
            b1 = rows[1, 1]
            assert ((b1 < half_t) | (b1 >= q - half_t)).all()
            ph0 = glwe_phase(p, torus_of(p, rows[0]).reshape(-1), keys.glwe_key)
            ph1 = glwe_phase(p, torus_of(p, rows[1]).reshape(-1), keys.glwe_key)
            want1 = np.zeros(N, np.int64)
            want1[0] = mu
            assert np.abs(wrap(ph0 + mu * z, w)).max() < (1 << 12), (i, lvl)
            assert np.abs(wrap(ph1 - want1, w)).max() < (1 << 12), (i, lvl)


def test_classic_key_bootstraps(p32):
    keys = yg.keygen_classic(p32, seed=42)
    orc = Oracle(p32, keys)
    bits = np.random.default_rng(4).integers(0, 2, 6)
    cts = yg.lwe_encrypt(keys.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=4)
    out = orc.pbs_batch(cts, yg.sign_test_polynomial(p32), threads=6)
    orc.close()
    ph = yg.lwe_decrypt_phase(out, keys.lwe_key).astype(np.int64)
    assert np.array_equal((ph > 0).astype(np.int64), bits)
