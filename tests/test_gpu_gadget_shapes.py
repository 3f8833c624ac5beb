"""GPU parity beyond the two headline parameter sets.

The headline shapes (TORUS32 N=1024 2^7 x 3, TORUS56 N=2048 2^10 x 3) run
specialised kernels.  Every other shape the engine accepts runs a kernel that
reads the gadget at run time: TORUS32 at N=2048 (two-wave teams), TORUS56 at
N=1024 (one-wave teams) and other gadget bases and depths at either width, up
to 8 levels for bases up to 2^7.
Each shape here is checked bit for bit against the oracle, batched and plain,
with a short LWE dimension so the oracle stays fast.
"""
import dataclasses

import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle

pytestmark = pytest.mark.gpu

SHAPES = [
    # (name, base parameter set, changes)
    ("t32_n1024_b8x2", yg.PARAMS_T32_N1024, dict(n=64, radix_bits=8, level=2)),
    ("t32_n1024_b6x4", yg.PARAMS_T32_N1024, dict(n=64, radix_bits=6, level=4)),
    ("t32_n2048_b7x3", yg.PARAMS_T32_N1024, dict(n=64, N=2048)),
    ("t56_n1024_b10x3", yg.PARAMS_T56_N2048, dict(n=64, N=1024)),
    ("t56_n2048_b8x4", yg.PARAMS_T56_N2048, dict(n=64, radix_bits=8, level=4)),
    ("t56_n2048_b15x3", yg.PARAMS_T56_N2048, dict(n=48, radix_bits=15, level=3)),
    # more than four levels (8-bit digit fields, partial sums reduced after four):
    # the reference's TORUS32 default gadget 2^4 x 8 at N=2048, a rounded
    # 2^3 x 5 (17 dropped bits), and a full 56-bit 2^7 x 8
    ("t32_n2048_b4x8", yg.PARAMS_T32_N1024, dict(n=40, N=2048, radix_bits=4, level=8)),
    ("t32_n1024_b3x5", yg.PARAMS_T32_N1024, dict(n=48, radix_bits=3, level=5)),
    ("t56_n2048_b7x8", yg.PARAMS_T56_N2048, dict(n=32, radix_bits=7, level=8)),
]


def encrypt(params, keys, count, seed):
    rng = np.random.default_rng(seed)
    mu = rng.integers(-(1 << 31), 1 << 31, count).astype(np.int64).astype(np.int32)
    return yg.lwe_encrypt(keys.lwe_key, mu, seed=seed)


def lut_polynomial(params, seed):
    if params.torus_bits == 32:
        return yg.sign_test_polynomial(params)
    lut = np.random.default_rng(seed).integers(0, 16, 16)
    return yg.lut_test_polynomial(params, lut, 16)


@pytest.mark.parametrize("name,base,changes", SHAPES, ids=[s[0] for s in SHAPES])
def test_runtime_gadget_shape_bit_exact(name, base, changes):
    params = dataclasses.replace(base, **changes)
    keys = yg.keygen(params, seed=21)
    cts = encrypt(params, keys, 24, seed=22)
    v = lut_polynomial(params, 23)
    orc = Oracle(params, keys)
    try:
        print(f"shape {name}: batch {len(cts)} + 2 plain calls", flush=True)
        with yg.Context(params, 0) as ctx:
            kid = ctx.upload_key(keys)
            out, acc, ext = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
            plain = [ctx.pbs(kid, cts[c], v, want_acc=True, want_extract=True) for c in (0, 7)]
        assert np.array_equal(out, orc.pbs_batch(cts, v)), f"{name}: keyswitched LWE differs"
        for c in (0, 7, 23):
            o_acc, o_ext, o_out = orc.pbs(cts[c], v)
            assert np.array_equal(acc[c], o_acc), f"{name}: accumulator differs, ct {c}"
            assert np.array_equal(ext[c], o_ext), f"{name}: extracted LWE differs, ct {c}"
        for (p_out, p_acc, p_ext), c in zip(plain, (0, 7)):
            assert np.array_equal(p_acc, acc[c]) and np.array_equal(p_ext, ext[c]) and np.array_equal(p_out, out[c])
    finally:
        orc.close()
