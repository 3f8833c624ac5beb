"""The lazy-pipe key format on the GPU against the oracle's blindRotateLazyPipeNtt
(blind_rotate.cpp:822-902), extraction and keyswitch, bit for bit.

Two shapes at N = 1024: the headline gadget 2^7 x 3 (its own kernel form) and
the reference's TORUS32 lazy-pipe shape, base 2^4 with 8 digit levels and 4
row levels (the runtime-gadget form).  Inputs include the edge cases of the
step loop: rotation amounts of 0 (skipped steps, a_0 = 0 for the bskFirst
term), b = 0 and extreme torus values; windows of one step, a few steps and
the automatic window give the same results."""
import ctypes
import dataclasses

import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import LazyOracle, P, lib, or_params

pytestmark = pytest.mark.gpu

SHAPES = {
    "b7x3": (dataclasses.replace(yg.PARAMS_T32_N1024, n=64), None, None),
    "b4x4_d8": (dataclasses.replace(yg.PARAMS_T32_N1024, n=64, radix_bits=4, level=4), 8, 2.21 - 32),
}


def lazy_keys(name, seed):
    p, digits, sigma = SHAPES[name]
    v = yg.sign_test_polynomial(p)
    return p, yg.keygen_lazy(p, seed=seed, v=v, digit_levels=digits, sigma_log2_glwe=sigma)


def inputs(p, lk, count, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2, count)
    cts = yg.lwe_encrypt(lk.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=seed)
    cts[0, 1::3] = 0              # steps with a_i = 0 (mod-switch to 0)
    cts[1, 0] = 0                 # a_0 = 0: no bskFirst term
    cts[2, :] = 0                 # every a_i = 0 and b = 0
    cts[3, :] = np.int32(-(1 << 31))
    cts[4, 5:9] = (1 << 21) - 1   # values just below a mod-switch tie
    return bits, cts


@pytest.mark.parametrize("name", list(SHAPES))
def test_lazy_batch_bit_exact(name):
    p, lk = lazy_keys(name, 61)
    bits, cts = inputs(p, lk, 24, 62)
    orc = LazyOracle(p, lk)
    ref_acc, ref_ext, ref_out = orc.pbs_batch_all(cts, lk.v)
    orc.close()
    with yg.Context(p, 0) as ctx:
        kid = ctx.upload_key_lazy(lk)
        for window in (0, 1, 5):
            print(f"lazy {name}: window {window}", flush=True)
            ctx.set_lazy_window(window)
            out, acc, ext = ctx.pbs_lazy_batch(kid, cts, lk.v, want_acc=True, want_extract=True)
            assert np.array_equal(acc, ref_acc), f"{name} window {window}: accumulators differ"
            assert np.array_equal(ext, ref_ext), f"{name} window {window}: extracted LWE differs"
            assert np.array_equal(out, ref_out), f"{name} window {window}: keyswitched LWE differs"
        # the plain form and blindRotateLazyPipeNtt on a rescaled input
        o1, a1, e1 = ctx.pbs_lazy(kid, cts[5], lk.v, want_acc=True, want_extract=True)
        assert np.array_equal(o1, ref_out[5]) and np.array_equal(a1, ref_acc[5]) and np.array_equal(e1, ref_ext[5])
        orc = LazyOracle(p, lk)
        for c in (0, 6):
            sc = np.zeros(p.n + 1, np.int32)  # rescaleTlweToNewMod: the ScaledTlwe
            lib().or_rescale(ctypes.byref(or_params(p)), P(np.ascontiguousarray(cts[c])), P(sc))
            assert np.array_equal(ctx.blind_rotate_lazy(kid, sc, lk.v), orc.blind_rotate_scaled(sc, lk.v)), c
        orc.close()
    if name == "b4x4_d8":
        # exact digits bootstrap correctly: the sign of every message (rows 2-4 are crafted)
        ph = yg.lwe_decrypt_phase(out, lk.lwe_key)
        assert np.array_equal((ph[5:] > 0).astype(int), bits[5:])


def test_lazy_full_dimension():
    """n = 630 (the headline LWE dimension), 2^4 x 4 rows on 8 digit levels:
    every result decrypts, a sample equals the oracle's."""
    p = dataclasses.replace(yg.PARAMS_T32_N1024, radix_bits=4, level=4)
    v = yg.sign_test_polynomial(p)
    lk = yg.keygen_lazy(p, seed=63, v=v, digit_levels=8, sigma_log2_glwe=2.21 - 32)
    rng = np.random.default_rng(64)
    bits = rng.integers(0, 2, 96)
    cts = yg.lwe_encrypt(lk.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=65)
    print("lazy n=630: batch of 96", flush=True)
    with yg.Context(p, 0) as ctx:
        kid = ctx.upload_key_lazy(lk)
        out, acc, _ = ctx.pbs_lazy_batch(kid, cts, v, want_acc=True)
    ph = yg.lwe_decrypt_phase(out, lk.lwe_key)
    assert np.array_equal((ph > 0).astype(int), bits)
    sample = [0, 47, 95]
    orc = LazyOracle(p, lk)
    ref_acc, _, ref_out = orc.pbs_batch_all(cts[sample], v)
    orc.close()
    assert np.array_equal(ref_out, out[sample]) and np.array_equal(ref_acc, acc[sample])


def test_lazy_and_engine_keys_stay_apart():
    p, lk = lazy_keys("b7x3", 66)
    keys = yg.keygen(p, seed=67)
    _, cts = inputs(p, lk, 6, 68)
    with yg.Context(p, 0) as ctx:
        lazy_id = ctx.upload_key_lazy(lk)
        mp_id = ctx.upload_key(keys)
        with pytest.raises(yg.YgpuError) as e:
            ctx.pbs_batch(lazy_id, cts, lk.v)
        assert e.value.code == -2
        with pytest.raises(yg.YgpuError) as e:
            ctx.pbs_lazy_batch(mp_id, cts, lk.v)
        assert e.value.code == -2
        # both still work after the refusals, and a released lazy key is gone
        ctx.pbs_batch(mp_id, cts, lk.v)
        ctx.pbs_lazy_batch(lazy_id, cts, lk.v)
        ctx.release_key(lazy_id)
        with pytest.raises(yg.YgpuError):
            ctx.pbs_lazy_batch(lazy_id, cts, lk.v)


# ------------------------------------------------------------------ Alt form
from tests.oracle_py import AltOracle  # noqa: E402


@pytest.mark.parametrize("name", list(SHAPES))
def test_lazy_alt_batch_bit_exact(name):
    """BootstrappingKeyMPLazyPipeAlt (the serving benchmark's OURS) against the
    oracle's blindRotateLazyPipeAltNtt, windows 0/1/5, the plain and scaled
    forms, and a test polynomial per ciphertext (one first component each)."""
    p, digits, sigma = SHAPES[name]
    ak = yg.keygen_lazy_alt(p, seed=71, digit_levels=digits, sigma_log2_glwe=sigma)
    v = yg.sign_test_polynomial(p)
    bits, cts = inputs(p, ak, 24, 72)
    orc = AltOracle(p, ak)
    ref_acc, ref_ext, ref_out = orc.pbs_batch_all(cts, v)
    with yg.Context(p, 0) as ctx:
        kid = ctx.upload_key_lazy(ak)
        for window in (0, 1, 5):
            print(f"lazy alt {name}: window {window}", flush=True)
            ctx.set_lazy_window(window)
            out, acc, ext = ctx.pbs_lazy_batch(kid, cts, v, want_acc=True, want_extract=True)
            assert np.array_equal(acc, ref_acc), f"{name} window {window}: accumulators differ"
            assert np.array_equal(ext, ref_ext) and np.array_equal(out, ref_out), f"{name} window {window}"
        o1, a1, _ = ctx.pbs_lazy(kid, cts[7], v, want_acc=True)
        assert np.array_equal(o1, ref_out[7]) and np.array_equal(a1, ref_acc[7])
        sc = np.zeros(p.n + 1, np.int32)
        lib().or_rescale(ctypes.byref(or_params(p)), P(np.ascontiguousarray(cts[9])), P(sc))
        assert np.array_equal(ctx.blind_rotate_lazy(kid, sc, v), orc.blind_rotate_scaled(sc, v))
        # a different test polynomial per ciphertext
        rng = np.random.default_rng(73)
        vs = np.stack([yg.lut_test_polynomial(p, rng.integers(0, 4, 4), 4) for _ in range(6)])
        out6, acc6, _ = ctx.pbs_lazy_batch(kid, cts[:6], vs, want_acc=True)
        for c in range(6):
            ra, _, ro = orc.pbs_batch_all(cts[c:c + 1], vs[c])
            assert np.array_equal(acc6[c], ra[0]) and np.array_equal(out6[c], ro[0]), c
    orc.close()
    if name == "b4x4_d8":
        ph = yg.lwe_decrypt_phase(out, ak.lwe_key)
        assert np.array_equal((ph[5:] > 0).astype(int), bits[5:])


def test_lazy_alt_full_dimension():
    p = dataclasses.replace(yg.PARAMS_T32_N1024, radix_bits=4, level=4)
    ak = yg.keygen_lazy_alt(p, seed=74, digit_levels=8, sigma_log2_glwe=2.21 - 32)
    v = yg.sign_test_polynomial(p)
    rng = np.random.default_rng(75)
    bits = rng.integers(0, 2, 96)
    cts = yg.lwe_encrypt(ak.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=76)
    print("lazy alt n=630: batch of 96", flush=True)
    with yg.Context(p, 0) as ctx:
        kid = ctx.upload_key_lazy(ak)
        out, acc, _ = ctx.pbs_lazy_batch(kid, cts, v, want_acc=True)
    assert np.array_equal((yg.lwe_decrypt_phase(out, ak.lwe_key) > 0).astype(int), bits)
    orc = AltOracle(p, ak)
    ref_acc, _, ref_out = orc.pbs_batch_all(cts[[0, 50, 95]], v)
    orc.close()
    assert np.array_equal(ref_out, out[[0, 50, 95]]) and np.array_equal(ref_acc, acc[[0, 50, 95]])
