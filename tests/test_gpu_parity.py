"""GPU parity: the HIP engine against the CPU oracle, bit for bit, through the C ABI."""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle

pytestmark = pytest.mark.gpu


def sign_cts(keys, count, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2, count)
    mu = ((2 * bits - 1) * (1 << 29)).astype(np.int32)
    return bits, yg.lwe_encrypt(keys.lwe_key, mu, seed=seed)


@pytest.fixture(scope="module")
def ctx32(p32, keys32):
    with yg.Context(p32, 0) as c:
        kid = c.upload_key(keys32)
        yield c, kid


@pytest.fixture(scope="module")
def orc32(p32, keys32):
    o = Oracle(p32, keys32)
    yield o
    o.close()


def test_plain_pbs_t32_bit_exact(p32, keys32, ctx32, orc32):
    ctx, kid = ctx32
    bits, cts = sign_cts(keys32, 2, seed=11)
    v = yg.sign_test_polynomial(p32)
    for c in range(len(cts)):
        out, acc, ext = ctx.pbs(kid, cts[c], v, want_acc=True, want_extract=True)
        o_acc, o_ext, o_out = orc32.pbs(cts[c], v)
        assert np.array_equal(acc, o_acc), f"accumulator differs, ct {c}"
        assert np.array_equal(ext, o_ext), f"extracted LWE differs, ct {c}"
        assert np.array_equal(out, o_out), f"keyswitched LWE differs, ct {c}"


def test_batched_pbs_t32_bit_exact(p32, keys32, ctx32, orc32):
    ctx, kid = ctx32
    bits, cts = sign_cts(keys32, 64, seed=12)
    v = yg.sign_test_polynomial(p32)
    out, acc, ext = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
    o_out = orc32.pbs_batch(cts, v)
    assert np.array_equal(out, o_out)
    for c in range(0, 64, 16):
        o_acc, o_ext, _ = orc32.pbs(cts[c], v)
        assert np.array_equal(acc[c], o_acc), f"accumulator differs, ct {c}"
        assert np.array_equal(ext[c], o_ext), f"extracted LWE differs, ct {c}"
    ph = yg.lwe_decrypt_phase(out, keys32.lwe_key)
    assert np.array_equal((ph > 0).astype(int), bits)


@pytest.fixture(scope="module")
def ctx56(p56, keys56):
    with yg.Context(p56, 0) as c:
        kid = c.upload_key(keys56)
        yield c, kid


def test_batched_pbs_t56_lut_bit_exact(p56, keys56, ctx56):
    ctx, kid = ctx56
    p = 16
    rng = np.random.default_rng(21)
    lut = rng.integers(0, p, p)
    v = yg.lut_test_polynomial(p56, lut, p)
    msgs = rng.integers(0, p, 8)
    cts = yg.lwe_encrypt(keys56.lwe_key, (msgs * (1 << 32) // (2 * p)).astype(np.int64).astype(np.int32), seed=22)
    out, acc, ext = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
    orc = Oracle(p56, keys56)
    for c in range(len(cts)):
        o_acc, o_ext, o_out = orc.pbs(cts[c], v)
        assert np.array_equal(acc[c], o_acc), f"accumulator differs, ct {c}"
        assert np.array_equal(ext[c], o_ext), f"extracted LWE differs, ct {c}"
        assert np.array_equal(out[c], o_out), f"keyswitched LWE differs, ct {c}"
    orc.close()
    ph = yg.lwe_decrypt_phase(out, keys56.lwe_key).astype(np.int64)
    assert np.array_equal(np.round(ph / (1 << 32) * p).astype(np.int64) % p, lut[msgs] % p)
