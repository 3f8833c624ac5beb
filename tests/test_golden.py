"""Committed golden fixtures (tests/golden/make_golden.py): the seeded key
material, the oracle's outputs, and -- on the GPU -- the engine's outputs."""
import hashlib
import json
import os

import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(HERE, "golden.json")))
G = np.load(os.path.join(HERE, "golden.npz"), allow_pickle=False)


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_key_material_is_reproducible(p32, keys32, p56, keys56):
    assert sha(keys32.lwe_key, keys32.glwe_key, keys32.bsk, keys32.ksk) == META["sha256"]["t32_keys"]
    assert sha(keys56.lwe_key, keys56.glwe_key, keys56.bsk, keys56.ksk) == META["sha256"]["t56_keys"]
    # thread count does not change the keys
    k1 = yg.keygen(p32, seed=0, threads=1)
    assert sha(k1.lwe_key, k1.glwe_key, k1.bsk, k1.ksk) == META["sha256"]["t32_keys"]


def test_oracle_matches_golden_chain(p32, keys32):
    orc = Oracle(p32, keys32)
    acc, ext, out = orc.pbs(G["t32_ct0"], yg.sign_test_polynomial(p32))
    assert np.array_equal(acc, G["t32_acc0"])
    assert np.array_equal(ext, G["t32_ext0"])
    assert np.array_equal(out, G["t32_out0"])
    ph = yg.lwe_decrypt_phase(out, keys32.lwe_key)
    assert int(ph[0] > 0) == int(G["t32_bits0"][0])
    orc.close()


def test_golden_outputs_decrypt(p32, keys32, p56, keys56):
    ph = yg.lwe_decrypt_phase(G["t32_out1"], keys32.lwe_key)
    assert np.array_equal((ph > 0).astype(int), G["t32_bits1"])
    ph = yg.lwe_decrypt_phase(G["t56_out2"], keys56.lwe_key).astype(np.int64)
    assert np.array_equal(np.round(ph / (1 << 32) * 16).astype(np.int64) % 16, G["t56_lut"][G["t56_msgs"]] % 16)
    assert sha(G["t32_out1"]) == META["sha256"]["t32_batch_outputs"]
    assert sha(G["t56_out2"]) == META["sha256"]["t56_batch_outputs"]


@pytest.mark.gpu
def test_engine_matches_golden(p32, keys32, p56, keys56):
    bits = np.random.default_rng(1).integers(0, 2, 64)
    ct1 = yg.lwe_encrypt(keys32.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=1)
    assert sha(ct1) == META["sha256"]["t32_batch_inputs"]
    v32 = yg.sign_test_polynomial(p32)
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys32)
        out, acc, ext = ctx.pbs(kid, G["t32_ct0"], v32, want_acc=True, want_extract=True)
        assert np.array_equal(acc, G["t32_acc0"])
        assert np.array_equal(ext, G["t32_ext0"])
        assert np.array_equal(out, G["t32_out0"])
        out1, _, _ = ctx.pbs_batch(kid, ct1, v32)
        assert np.array_equal(out1, G["t32_out1"])
    rng = np.random.default_rng(2)
    lut = rng.integers(0, 16, 16)
    msgs = rng.integers(0, 16, 16)
    ct2 = yg.lwe_encrypt(keys56.lwe_key, (msgs * (1 << 32) // 32).astype(np.int64).astype(np.int32), seed=3)
    assert sha(ct2) == META["sha256"]["t56_batch_inputs"]
    with yg.Context(p56, 0) as ctx:
        kid = ctx.upload_key(keys56)
        out2, _, _ = ctx.pbs_batch(kid, ct2, yg.lut_test_polynomial(p56, lut, 16))
        assert np.array_equal(out2, G["t56_out2"])
