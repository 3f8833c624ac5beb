"""A blind-rotation team barrier that times out must reach the caller as
YGPU_ERR_DEVICE, not as a plausible-looking result with YGPU_OK.

The condition is forced from the host: ygpu_set_barrier_polls(0) makes every
team barrier give up at once (no wait, every kernel still finishes), the
workgroup marks its accumulators invalid and stores its failure flag into the
context's status word, and the next synchronising call reports it.  Restoring
the bound gives results bit-identical to the oracle again."""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle

pytestmark = pytest.mark.gpu

DEFAULT_POLLS = 1 << 16


def test_plain_call_reports_barrier_failure_t32(p32, keys32):
    # the plain TORUS32 call runs the latency kernel, whose row barrier is a team barrier
    bits = np.array([1, 0])
    cts = yg.lwe_encrypt(keys32.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=31)
    v = yg.sign_test_polynomial(p32)
    orc = Oracle(p32, keys32)
    want = [orc.pbs(c, v)[2] for c in cts]
    orc.close()
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys32)
        assert np.array_equal(ctx.pbs(kid, cts[0], v)[0], want[0])
        ctx.set_barrier_polls(0)
        with pytest.raises(yg.YgpuError) as e:
            ctx.pbs(kid, cts[1], v)
        assert e.value.code == yg.YGPU_ERR_DEVICE
        assert "barrier" in str(e.value)
        ctx.set_barrier_polls(DEFAULT_POLLS)
        # the report was cleared with the error; results are exact again
        for c in range(2):
            assert np.array_equal(ctx.pbs(kid, cts[c], v)[0], want[c])


def test_batched_call_reports_barrier_failure_t56(p56, keys56):
    # the N = 2048 TORUS56 kernel synchronises its two-wave teams with team barriers
    msgs = np.array([3, 9, 14])
    cts = yg.lwe_encrypt(keys56.lwe_key, (msgs * (1 << 32) // 32).astype(np.int64).astype(np.int32), seed=41)
    lut = np.arange(16)[::-1].copy()
    v = yg.lut_test_polynomial(p56, lut, 16)
    orc = Oracle(p56, keys56)
    want = np.stack([orc.pbs(c, v)[2] for c in cts])
    orc.close()
    with yg.Context(p56, 0) as ctx:
        kid = ctx.upload_key(keys56)
        ctx.set_barrier_polls(0)
        with pytest.raises(yg.YgpuError) as e:
            ctx.pbs_batch(kid, cts, v)   # asynchronous launch; the sync inside reports it
        assert e.value.code == yg.YGPU_ERR_DEVICE
        ctx.set_barrier_polls(DEFAULT_POLLS)
        out, _, _ = ctx.pbs_batch(kid, cts, v)
    assert np.array_equal(out, want)


def test_batched_call_reports_handover_failure_t32(p32, keys32):
    # the TORUS32 batched tile hands its partial sums between a ciphertext's
    # two one-wave teams with a bounded wait on a count in LDS
    bits = np.array([1, 0, 1, 1, 0])
    cts = yg.lwe_encrypt(keys32.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=37)
    v = yg.sign_test_polynomial(p32)
    orc = Oracle(p32, keys32)
    want = orc.pbs_batch(cts, v)
    orc.close()
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys32)
        ctx.set_tile_ex(2, 1, 2)   # the throughput tile even for a short batch
        ctx.set_barrier_polls(0)
        with pytest.raises(yg.YgpuError) as e:
            ctx.pbs_batch(kid, cts, v)
        assert e.value.code == yg.YGPU_ERR_DEVICE
        ctx.set_barrier_polls(DEFAULT_POLLS)
        out, _, _ = ctx.pbs_batch(kid, cts, v)
    assert np.array_equal(out, want)
