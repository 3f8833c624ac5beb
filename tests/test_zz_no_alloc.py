"""ygpu_pbs_batch_device_noalloc: the batched PBS with its scratch in the
library's static device array must give exactly what ygpu_pbs_batch_device
gives, call after call (the array keeps the previous call's contents; a
smaller second batch must not see them), and refuse batches beyond the array."""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle

pytestmark = pytest.mark.gpu


def run_both(ctx, kid, p, cts, v, want_acc):
    count = len(cts)
    tb = np.dtype(p.torus_dtype).itemsize
    d_lwe = ctx.alloc(cts.nbytes); d_lwe.upload(cts)
    d_v = ctx.alloc(v.nbytes); d_v.upload(v)
    outs = []
    for noalloc in (False, True):
        d_out = ctx.alloc(count * (p.n + 1) * 4)
        d_acc = ctx.alloc(count * (p.k + 1) * p.N * tb) if want_acc else None
        d_ext = ctx.alloc(count * (p.k * p.N + 1) * 4)
        if noalloc:
            ctx.pbs_batch_device_noalloc(kid, count, d_lwe, d_v, False, d_out, d_acc, d_ext)
        else:
            ctx.pbs_batch_device(kid, count, d_lwe, d_v, False, d_out, d_acc, d_ext)
        ctx.sync()
        out = d_out.download(np.zeros((count, p.n + 1), np.int32))
        ext = d_ext.download(np.zeros((count, p.k * p.N + 1), np.int32))
        acc = d_acc.download(np.zeros((count, p.k + 1, p.N), p.torus_dtype)) if d_acc else None
        outs.append((out, ext, acc))
        for b in (d_out, d_acc, d_ext):
            if b:
                b.free()
    d_lwe.free()
    d_v.free()
    return outs


@pytest.mark.parametrize("width", [32, 56])
def test_noalloc_equals_batch_two_calls(width, p32, keys32, p56, keys56):
    p, keys = (p32, keys32) if width == 32 else (p56, keys56)
    rng = np.random.default_rng(70 + width)
    if width == 32:
        v = yg.sign_test_polynomial(p)
        sizes = (300, 37)   # the throughput tile, then the per-ciphertext tile with sliced keyswitch sums
    else:
        v = yg.lut_test_polynomial(p, rng.integers(0, 16, 16), 16)
        sizes = (260, 6)
    with yg.Context(p, 0) as ctx:
        kid = ctx.upload_key(keys)
        first = None
        for call, count in enumerate(sizes):
            bits = rng.integers(0, 2, count)
            cts = yg.lwe_encrypt(keys.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=100 * width + call)
            # the library's own accumulators on the first call, the caller's on the second
            (o1, e1, _), (o2, e2, a2) = run_both(ctx, kid, p, cts, v, want_acc=(call == 1))
            assert np.array_equal(o1, o2), f"call {call}: keyswitched LWEs differ"
            assert np.array_equal(e1, e2), f"call {call}: extracted LWEs differ"
            if first is None:
                first = cts
            else:
                orc = Oracle(p, keys)
                o_acc, o_ext, o_out = orc.pbs_batch_all(cts[:2], v)
                orc.close()
                assert np.array_equal(o_out, o2[:2]) and np.array_equal(o_ext, e2[:2])
                assert np.array_equal(o_acc, a2[:2])
        # a batch beyond the static array is refused, naming the limit
        big = 6000 if width == 32 else 2000
        d_lwe = ctx.alloc(big * (p.n + 1) * 4)
        d_v = ctx.alloc(v.nbytes)
        d_out = ctx.alloc(big * (p.n + 1) * 4)
        with pytest.raises(yg.YgpuError) as e:
            ctx.pbs_batch_device_noalloc(kid, big, d_lwe, d_v, False, d_out)
        assert e.value.code == -1 and "static scratch holds" in str(e.value)
        for b in (d_lwe, d_v, d_out):
            b.free()
