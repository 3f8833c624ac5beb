"""The benchmark workloads at their full sizes on the GPU, checked against the
oracle.

* TORUS56 n=742 N=2048 gadget 2^10 x 3, 8192 ciphertexts of uniform 4-bit
  messages under one key with a random 16-entry LUT, on the autotuned tile:
  determinism, every output LUT-decrypts, a seeded sample of 16 bit-exact
  against the oracle (accumulator, extracted and keyswitched LWE), and a
  result alone equals the same ciphertext in the batch.
* 64 tenants at TORUS32 n=630 N=1024 with Zipf(1.1) ciphertext counts (8,224
  in all) in one key-grouped hipGraph launch: the first and last ciphertext
  of every tenant against the oracle under that tenant's key, every output
  decrypts to its sign, graph replays equal the eager pipeline.

The gate is BOOTSTRAPPING.MP21_FR's (bootstrappping_test.cpp:133-180): every
programmable-bootstrap output decrypts to its table entry."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle

pytestmark = pytest.mark.gpu

P16 = 16


def test_t56_full_batch_8192(p56, keys56):
    count = 8192
    lut = np.random.default_rng(2).integers(0, P16, P16)
    v = yg.lut_test_polynomial(p56, lut, P16)
    msgs = np.random.default_rng(3).integers(0, P16, count)
    cts = yg.lwe_encrypt(keys56.lwe_key, (msgs * (1 << 32) // (2 * P16)).astype(np.int64).astype(np.int32), seed=3)
    idx = np.random.default_rng(4).choice(count, 16, replace=False)
    with yg.Context(p56, 0) as ctx:
        kid = ctx.upload_key(keys56)
        print("t56 8192: autotune", flush=True)
        cpw, twl, ms = ctx.autotune(kid, count)
        print(f"t56 8192: tile ({cpw}, {twl}) {ms:.1f} ms; two batches", flush=True)
        out1, acc1, ext1 = ctx.pbs_batch(kid, cts, v, want_acc=True, want_extract=True)
        out2, _, _ = ctx.pbs_batch(kid, cts, v)
        alone = [ctx.pbs(kid, cts[c], v)[0] for c in idx[:2]]
    assert np.array_equal(out1, out2), "two runs of the same batch differ"
    for c, a in zip(idx[:2], alone):
        assert np.array_equal(a, out1[c]), f"ciphertext {c} alone differs from the batch"
    ph = yg.lwe_decrypt_phase(out1, keys56.lwe_key).astype(np.int64)
    dec = np.round(ph / (1 << 32) * P16).astype(np.int64) % P16
    assert np.array_equal(dec, lut[msgs] % P16), f"{int((dec != lut[msgs] % P16).sum())} outputs decrypt wrongly"
    print("t56 8192: oracle on 16", flush=True)
    orc = Oracle(p56, keys56)
    o_acc, o_ext, o_out = orc.pbs_batch_all(cts[idx], v)
    orc.close()
    assert np.array_equal(o_acc, acc1[idx]), "accumulators differ from the oracle"
    assert np.array_equal(o_ext, ext1[idx]), "extracted LWEs differ from the oracle"
    assert np.array_equal(o_out, out1[idx]), "keyswitched LWEs differ from the oracle"


def multitenant_counts():
    """c_t = clamp(round(12288 r_t^-1.1 / H_64), 1, 512), ranks from a seed-4 permutation."""
    r = np.random.default_rng(4).permutation(64) + 1
    h = sum(k ** -1.1 for k in range(1, 65))
    return np.clip(np.round(12288 * r.astype(np.float64) ** -1.1 / h), 1, 512).astype(int)


def _oracle_ends(params, keys, cts, v):
    orc = Oracle(params, keys)
    out = orc.pbs_batch(cts[[0, -1]], v, threads=2)
    orc.close()
    return out


def test_multitenant_64_graph_full(p32):
    counts = multitenant_counts()
    total = int(counts.sum())
    assert total == 8224
    v = yg.sign_test_polynomial(p32)
    cts, bits, lwe_keys, futs = [], [], [], []
    with yg.Context(p32, 0) as ctx, ThreadPoolExecutor(8) as pool:
        kids = []
        print("multitenant: 64 keys", flush=True)
        for t in range(64):
            k = yg.keygen(p32, seed=1000 + t)
            kids.append(ctx.upload_key(k))
            b = np.random.default_rng(2000 + t).integers(0, 2, int(counts[t]))
            c = yg.lwe_encrypt(k.lwe_key, ((2 * b - 1) * (1 << 29)).astype(np.int32), seed=2000 + t)
            bits.append(b)
            cts.append(c)
            lwe_keys.append(k.lwe_key)
            futs.append(pool.submit(_oracle_ends, p32, k, c, v))
            del k
        all_cts = np.concatenate(cts)
        jobs, first = [], 0
        for t in range(64):
            jobs.append((kids[t], int(counts[t]), first))
            first += int(counts[t])
        d_lwe = ctx.alloc(all_cts.nbytes); d_lwe.upload(all_cts)
        d_v = ctx.alloc(v.nbytes); d_v.upload(v)
        d_out = ctx.alloc(total * (p32.n + 1) * 4)
        d_eager = ctx.alloc(total * (p32.n + 1) * 4)
        print("multitenant: graph", flush=True)
        gid = ctx.graph_create(jobs, total, d_lwe, d_v, False, d_out)
        replays = []
        for _ in range(2):
            ctx.graph_launch(gid)
            ctx.sync()
            replays.append(d_out.download(np.zeros((total, p32.n + 1), np.int32)))
            ctx.pbs_jobs_device(jobs, total, d_lwe, d_v, False, d_eager)  # eager work between replays
            ctx.sync()
        eager = d_eager.download(np.zeros((total, p32.n + 1), np.int32))
        ctx.graph_destroy(gid)
        for bf in (d_lwe, d_v, d_out, d_eager):
            bf.free()
        print("multitenant: oracle", flush=True)
        ends = [f.result() for f in futs]
    out = replays[0]
    assert np.array_equal(out, replays[1]) and np.array_equal(out, eager), "graph replays differ from the eager pipeline"
    for t in range(64):
        s, e = jobs[t][2], jobs[t][2] + jobs[t][1]
        assert np.array_equal(out[s], ends[t][0]), f"tenant {t}: first ciphertext differs from the oracle"
        assert np.array_equal(out[e - 1], ends[t][1]), f"tenant {t}: last ciphertext differs from the oracle"
        ph = yg.lwe_decrypt_phase(out[s:e], lwe_keys[t])
        assert np.array_equal((ph > 0).astype(int), bits[t]), f"tenant {t}: a sign decrypts wrongly"
