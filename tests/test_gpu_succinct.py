"""WWL+24 succinct keys on the GPU: the device derivation of the full rows
(k_bsk_derive) against the oracle's derivation, through every use of a key id.

A succinct upload must give the same bits as uploading the oracle-derived full
key through the ordinary path, and both must equal the oracle's bootstrap with
the derived key; at the reference's TORUS32 defaults the results also decrypt
to the sign (the gate of BLIND_ROT.BLIND_ROT_WWL24_NTT, blindrot_test.cpp:279-333,
carried through extraction and keyswitch)."""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle, derive_succinct
from tests.test_succinct_key import P_REF32, SIG_REF32

pytestmark = pytest.mark.gpu


def sign_inputs(keys, count, seed, sigma=yg.SIGMA_LOG2_LWE):
    bits = np.random.default_rng(seed).integers(0, 2, count)
    return bits, yg.lwe_encrypt(keys.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=seed, sigma_log2=sigma)


def check_against_full(params, sk, cts, v, n_oracle):
    print(f"succinct: torus {params.torus_bits} N={params.N} s2_level={sk.s2_level} batch {len(cts)}", flush=True)
    full = derive_succinct(params, sk)
    keys = yg.KeySet(params, sk.lwe_key, sk.glwe_key, full, sk.ksk)
    with yg.Context(params, 0) as ctx:
        ks = ctx.upload_key_succinct(sk)
        kf = ctx.upload_key(keys)
        assert ctx.key_bytes(ks) == ctx.key_bytes(kf)
        out_s, acc_s, _ = ctx.pbs_batch(ks, cts, v, want_acc=True)
        out_f, acc_f, _ = ctx.pbs_batch(kf, cts, v, want_acc=True)
        assert np.array_equal(acc_s, acc_f), "succinct-key accumulators differ from the derived full key's"
        assert np.array_equal(out_s, out_f)
        one, acc1, _ = ctx.pbs(ks, cts[0], v, want_acc=True)
        assert np.array_equal(one, out_s[0]) and np.array_equal(acc1.reshape(-1), acc_s[0].reshape(-1))
    orc = Oracle(params, keys)
    ref_acc, _, ref_out = orc.pbs_batch_all(cts[:n_oracle], v)
    orc.close()
    assert np.array_equal(ref_acc.reshape(n_oracle, -1).astype(np.int64),
                          acc_s[:n_oracle].reshape(n_oracle, -1).astype(np.int64))
    assert np.array_equal(ref_out, out_s[:n_oracle])
    return out_s


def test_succinct_t32_matches_oracle(p32):
    sk = yg.keygen_succinct(p32, seed=21, s2_level=4)
    _, cts = sign_inputs(sk, 64, 5)
    check_against_full(p32, sk, cts, yg.sign_test_polynomial(p32), 8)


def test_succinct_t56_matches_oracle(p56):
    sk = yg.keygen_succinct(p56, seed=22)
    assert sk.s2_level == 5
    msgs = np.random.default_rng(6).integers(0, 16, 24)
    cts = yg.lwe_encrypt(sk.lwe_key, (msgs * (1 << 32) // 32).astype(np.int64).astype(np.int32), seed=6)
    lut = np.random.default_rng(7).integers(0, 16, 16)
    check_against_full(p56, sk, cts, yg.lut_test_polynomial(p56, lut, 16), 3)


def test_succinct_reference_defaults_decrypt():
    p = P_REF32
    sk = yg.keygen_succinct(p, seed=23, sigma_log2_lwe=SIG_REF32[0], sigma_log2_glwe=SIG_REF32[1])
    bits, cts = sign_inputs(sk, 32, 8, sigma=SIG_REF32[0])
    out = check_against_full(p, sk, cts, yg.sign_test_polynomial(p), 2)
    ph = yg.lwe_decrypt_phase(out, sk.lwe_key).astype(np.int64)
    assert np.array_equal((ph > 0).astype(np.int64), bits), "succinct-key bootstrap does not decrypt to the sign"


def test_succinct_key_file_upload(tmp_path, p32):
    sk = yg.keygen_succinct(p32, seed=24, s2_level=4)
    path = str(tmp_path / "succinct.ygkey")
    yg.save_keys(path, sk)
    _, cts = sign_inputs(sk, 16, 9)
    v = yg.sign_test_polynomial(p32)
    print("succinct: key file upload", flush=True)
    with yg.Context(p32, 0) as ctx:
        kf = ctx.upload_key_file(path)
        ks = ctx.upload_key_succinct(sk)
        a, acc_a, _ = ctx.pbs_batch(kf, cts, v, want_acc=True)
        b, acc_b, _ = ctx.pbs_batch(ks, cts, v, want_acc=True)
    assert np.array_equal(a, b) and np.array_equal(acc_a, acc_b)
    # the file-loaded key against the oracle with the oracle-derived full key
    keys = yg.KeySet(p32, sk.lwe_key, sk.glwe_key, derive_succinct(p32, sk), sk.ksk)
    orc = Oracle(p32, keys)
    ref_acc, _, ref_out = orc.pbs_batch_all(cts[:4], v)
    orc.close()
    assert np.array_equal(ref_out, a[:4])
    assert np.array_equal(ref_acc.reshape(4, -1).astype(np.int64), acc_a[:4].reshape(4, -1).astype(np.int64))
