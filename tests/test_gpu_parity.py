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
