"""GPU parity: the HIP engine against the CPU oracle, bit for bit, through the C ABI."""
import os

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


def edge_ciphertexts(p, keys):
    """Inputs at the edges of the modswitch and rotation rules."""
    n = p.n
    scale = (1 << 32) // (2 * p.N)
    rows = []
    z = np.zeros(n + 1, np.int64)
    rows.append(z.copy())                                    # every abar = 0: all steps skipped
    r = z.copy(); r[-1] = -(1 << 31); rows.append(r)         # bbar = -N
    r = np.full(n + 1, -(1 << 31), np.int64); rows.append(r) # INT32_MIN everywhere
    r = np.full(n + 1, (1 << 31) - 1, np.int64); rows.append(r)
    r = np.zeros(n + 1, np.int64); r[::2] = scale // 2; r[1::2] = -(scale // 2); rows.append(r)   # ties round away
    r = np.zeros(n + 1, np.int64); r[::3] = scale // 2 - 1; r[1::3] = p.N * scale; rows.append(r)
    rng = np.random.default_rng(99)
    r = rng.integers(-(1 << 31), 1 << 31, n + 1); rows.append(r)
    return np.array(rows).astype(np.int64).astype(np.int32)


def test_edge_inputs_t32(p32, keys32, ctx32, orc32):
    ctx, kid = ctx32
    cts = edge_ciphertexts(p32, keys32)
    rng = np.random.default_rng(5)
    vs = [yg.sign_test_polynomial(p32),
          rng.integers(-(1 << 31), 1 << 31, p32.N).astype(np.int32)]
    vs[1][:3] = [-(1 << 31), (1 << 31) - 1, 0]
    for v in vs:
        out, acc, ext = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
        for c in range(len(cts)):
            o_acc, o_ext, o_out = orc32.pbs(cts[c], v)
            assert np.array_equal(acc[c], o_acc), f"accumulator differs, edge ct {c}"
            assert np.array_equal(ext[c], o_ext), f"extract differs, edge ct {c}"
            assert np.array_equal(out[c], o_out), f"keyswitch differs, edge ct {c}"


@pytest.fixture(scope="module")
def orc56(p56, keys56):
    o = Oracle(p56, keys56)
    yield o
    o.close()


def test_edge_inputs_t56(p56, keys56, ctx56, orc56):
    ctx, kid = ctx56
    cts = edge_ciphertexts(p56, keys56)
    rng = np.random.default_rng(15)
    half = 1 << 55
    v = rng.integers(-half, half, p56.N).astype(np.int64)
    v[:4] = [-half, half - 1, 0, -1]
    out, acc, ext = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
    for c in range(len(cts)):
        o_acc, o_ext, o_out = orc56.pbs(cts[c], v)
        assert np.array_equal(acc[c], o_acc), f"accumulator differs, edge ct {c}"
        assert np.array_equal(ext[c], o_ext), f"extract differs, edge ct {c}"
        assert np.array_equal(out[c], o_out), f"keyswitch differs, edge ct {c}"


def test_extract_keyswitch_adversarial_t56(p56, keys56, ctx56, orc56):
    # the 2^56 -> 2^32 modswitch of the extracted sample rounds half away from
    # zero: accumulators at exact ties (k 2^24 + 2^23), at the torus extremes
    # and random
    ctx, kid = ctx56
    rng = np.random.default_rng(16)
    half = 1 << 55
    for trial in range(4):
        acc = rng.integers(-half, half, (2, p56.N)).astype(np.int64)
        if trial == 1:
            acc = (rng.integers(-(1 << 31), 1 << 31, (2, p56.N)) * (1 << 24) + (1 << 23)).astype(np.int64)
            acc = ((acc + half) % (1 << 56)) - half
        if trial == 2:
            acc[:] = -half
            acc[:, 1::2] = half - 1
        if trial == 3:
            acc[:] = 0
        ext, out = ctx.extract_keyswitch(kid, acc)
        o_ext = orc56.extract(acc)
        assert np.array_equal(ext, o_ext), f"extract differs, trial {trial}"
        assert np.array_equal(out, orc56.keyswitch(o_ext)), f"keyswitch differs, trial {trial}"


def test_extract_keyswitch_adversarial_t32(p32, keys32, ctx32, orc32):
    ctx, kid = ctx32
    rng = np.random.default_rng(6)
    for trial in range(4):
        acc = rng.integers(-(1 << 31), 1 << 31, (2, p32.N)).astype(np.int32)
        if trial == 1:
            acc[:] = -(1 << 31)
        if trial == 2:
            acc[:] = 0
        ext, out = ctx.extract_keyswitch(kid, acc)
        o_ext = orc32.extract(acc)
        assert np.array_equal(ext, o_ext)
        assert np.array_equal(out, orc32.keyswitch(o_ext))


def test_multi_key_jobs_t32(p32):
    keysets = [yg.keygen(p32, seed=100 + t) for t in range(3)]
    counts = [5, 70, 1]
    v = yg.sign_test_polynomial(p32)
    cts, bits = [], []
    for t, ks in enumerate(keysets):
        b = np.random.default_rng(t).integers(0, 2, counts[t])
        bits.append(b)
        cts.append(yg.lwe_encrypt(ks.lwe_key, ((2 * b - 1) * (1 << 29)).astype(np.int32), seed=200 + t))
    allc = np.concatenate(cts)
    with yg.Context(p32, 0) as ctx:
        kids = [ctx.upload_key(ks) for ks in keysets]
        jobs, first = [], 0
        for t in range(3):
            jobs.append((kids[t], counts[t], first))
            first += counts[t]
        d_lwe = ctx.alloc(allc.nbytes); d_lwe.upload(allc)
        d_v = ctx.alloc(v.nbytes); d_v.upload(v)
        d_out = ctx.alloc(allc.shape[0] * (p32.n + 1) * 4)
        ctx.pbs_jobs_device(jobs, allc.shape[0], d_lwe, d_v, False, d_out)
        ctx.sync()
        mixed = d_out.download(np.zeros((allc.shape[0], p32.n + 1), np.int32))
        for b in (d_lwe, d_v, d_out):
            b.free()
        first = 0
        for t in range(3):
            alone, _, _ = ctx.pbs_batch(kids[t], cts[t], v)
            assert np.array_equal(mixed[first:first + counts[t]], alone), f"tenant {t} differs alone vs in the mix"
            ph = yg.lwe_decrypt_phase(alone, keysets[t].lwe_key)
            assert np.array_equal((ph > 0).astype(int), bits[t])
            first += counts[t]
    orc = Oracle(p32, keysets[1])
    o = orc.pbs(cts[1][7], v)[2]
    assert np.array_equal(mixed[5 + 7], o)
    orc.close()


def test_full_batch_4096_t32(p32, keys32, ctx32, orc32):
    """The benchmark's batch size: a sample against the oracle, every output
    decrypted, determinism, and each result equal alone and in the batch."""
    ctx, kid = ctx32
    bits, cts = sign_cts(keys32, 4096, seed=1)
    v = yg.sign_test_polynomial(p32)
    out1, _, _ = ctx.pbs_batch(kid, cts, v)
    out2, _, _ = ctx.pbs_batch(kid, cts, v)
    assert np.array_equal(out1, out2)
    ph = yg.lwe_decrypt_phase(out1, keys32.lwe_key)
    assert np.array_equal((ph > 0).astype(int), bits)
    idx = np.random.default_rng(3).choice(4096, 24, replace=False)
    assert np.array_equal(out1[idx], orc32.pbs_batch(cts[idx], v))
    for c in idx[:3]:
        alone, _, _ = ctx.pbs(kid, cts[c], v)
        assert np.array_equal(alone, out1[c])


def test_key_file_upload_matches_direct_upload(tmp_path, p32, keys32):
    """A key read back from the interchange file bootstraps bit-identically."""
    path = str(tmp_path / "keys32.ygkey")
    yg.save_keys(path, keys32)
    bits = np.random.default_rng(31).integers(0, 2, 5)
    cts = yg.lwe_encrypt(keys32.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=31)
    v = yg.sign_test_polynomial(p32)
    with yg.Context(p32, 0) as ctx:
        a = ctx.upload_key(keys32)
        b = ctx.upload_key_file(path)
        out_a, _, _ = ctx.pbs_batch(a, cts, v)
        out_b, acc_b, _ = ctx.pbs_batch(b, cts, v, want_acc=True)
    assert np.array_equal(out_a, out_b)
    # and the file-loaded key against the oracle, accumulators too
    orc = Oracle(p32, keys32)
    ref_acc, _, ref_out = orc.pbs_batch_all(cts, v)
    orc.close()
    assert np.array_equal(ref_out, out_b)
    assert np.array_equal(ref_acc, acc_b)
    assert np.array_equal((yg.lwe_decrypt_phase(out_b, keys32.lwe_key) > 0).astype(int), bits)


def test_all_4096_outputs_match_oracle(p32, keys32, ctx32, orc32):
    """Config 1's full batch: all 4096 x 3 outputs (accumulator, extracted LWE,
    keyswitched LWE) against the oracle, the oracle on all host threads."""
    ctx, kid = ctx32
    bits, cts = sign_cts(keys32, 4096, seed=1)
    v = yg.sign_test_polynomial(p32)
    out, acc, ext = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
    print("oracle: 4096 ciphertexts", flush=True)
    o_acc, o_ext, o_out = orc32.pbs_batch_all(cts, v, threads=min(16, os.cpu_count() or 1))
    assert np.array_equal(acc, o_acc), "accumulators differ"
    assert np.array_equal(ext, o_ext), "extracted LWEs differ"
    assert np.array_equal(out, o_out), "keyswitched LWEs differ"


def test_classic_key_bit_exact(p32):
    """A classic key (genBootstrappingKey) through the batched and plain forms:
    bit-exact against the oracle, decrypting every sign."""
    keys = yg.keygen_classic(p32, seed=43)
    bits = np.random.default_rng(8).integers(0, 2, 40)
    cts = yg.lwe_encrypt(keys.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=8)
    v = yg.sign_test_polynomial(p32)
    print("classic key: batch 40", flush=True)
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys)
        out, acc, _ = ctx.pbs_batch(kid, cts, v, want_acc=True)
        one, _, _ = ctx.pbs(kid, cts[3], v)
    orc = Oracle(p32, keys)
    ref_acc, _, ref_out = orc.pbs_batch_all(cts[:4], v)
    orc.close()
    assert np.array_equal(ref_out, out[:4]) and np.array_equal(one, out[3])
    assert np.array_equal(ref_acc.reshape(4, -1).astype(np.int64), acc[:4].reshape(4, -1).astype(np.int64))
    ph = yg.lwe_decrypt_phase(out, keys.lwe_key).astype(np.int64)
    assert np.array_equal((ph > 0).astype(np.int64), bits)


@pytest.mark.parametrize("index", [1, 64, 1023])
def test_extract_index_t32(p32, keys32, orc32, index):
    """The bootstrap extracting at param.driftPhase != 0 (as compare_test.cpp:22
    sets it): batched, plain and the standalone extraction against the oracle."""
    bits, cts = sign_cts(keys32, 6, seed=60 + index)
    v = yg.sign_test_polynomial(p32)
    orc = Oracle(p32, keys32, extract_index=index)
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys32)
        ctx.set_extract_index(index)
        assert ctx.extract_index() == index
        out, acc, ext = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
        one, _, one_ext = ctx.pbs(kid, cts[2], v, want_extract=True)
        alone = ctx.sample_extract(acc[1], index)
        at0 = ctx.sample_extract(acc[1], 0)
    o_acc, o_ext, o_out = orc.pbs_batch_all(cts, v)
    orc.close()
    assert np.array_equal(acc, o_acc)
    assert np.array_equal(ext, o_ext), "extracted LWEs differ"
    assert np.array_equal(out, o_out), "keyswitched LWEs differ"
    assert np.array_equal(one, o_out[2]) and np.array_equal(one_ext, o_ext[2])
    assert np.array_equal(alone, o_ext[1])
    assert np.array_equal(at0, orc32.extract(acc[1]))


def test_extract_index_t56_and_server(p56, keys56):
    index = 300
    msgs = np.random.default_rng(61).integers(0, 16, 4)
    cts = yg.lwe_encrypt(keys56.lwe_key, (msgs * (1 << 32) // 32).astype(np.int64).astype(np.int32), seed=61)
    v = yg.lut_test_polynomial(p56, np.random.default_rng(62).integers(0, 16, 16), 16)
    orc = Oracle(p56, keys56, extract_index=index)
    o_acc, o_ext, o_out = orc.pbs_batch_all(cts, v)
    orc.close()
    with yg.Context(p56, 0) as ctx:
        kid = ctx.upload_key(keys56)
        ctx.set_extract_index(index)
        out, _, ext = ctx.pbs_batch(kid, cts, v, want_extract=True)
    assert np.array_equal(ext, o_ext) and np.array_equal(out, o_out)
    with yg.Server(0, lanes=1, max_batch=64) as srv:
        ps = srv.add_params(p56)
        srv.set_extract_index(ps, index)
        t = srv.add_tenant(ps, keys56)
        reqs = [srv.request(t, c, v) for c in cts]
        for r in reqs:
            srv.submit(r)
        for r, want in zip(reqs, o_out):
            assert np.array_equal(srv.wait(r, timeout=60), want)


def test_concurrent_plain_calls_one_context(p32, keys32, orc32):
    """Synchronous calls from several threads on one context run on lanes of
    their own (the reference's serving workers call the bootstrap at once):
    every result equals the oracle's, and a key release waits for them."""
    import threading
    bits, cts = sign_cts(keys32, 24, seed=63)
    v = yg.sign_test_polynomial(p32)
    want = orc32.pbs_batch(cts, v)
    got = [None] * len(cts)
    errors = []
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys32)

        def work(idx):
            try:
                for c in idx:
                    got[c] = ctx.pbs(kid, cts[c], v)[0]
            except Exception as e:  # reported below
                errors.append(e)

        ths = [threading.Thread(target=work, args=(range(t, len(cts), 6),)) for t in range(6)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        ctx.release_key(kid)
    assert not errors, errors
    for c in range(len(cts)):
        assert np.array_equal(got[c], want[c]), f"ciphertext {c} differs from the oracle"
