"""hipGraph replays and tile choices must not change a single output bit."""
import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle

pytestmark = pytest.mark.gpu


def sign_batch(keys, count, seed):
    bits = np.random.default_rng(seed).integers(0, 2, count)
    return bits, yg.lwe_encrypt(keys.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=seed)


def test_graph_replays_match_eager(p32):
    keysets = [yg.keygen(p32, seed=300 + t) for t in range(2)]
    orcs = [Oracle(p32, k) for k in keysets]
    counts = [70, 130]
    v = yg.sign_test_polynomial(p32)
    with yg.Context(p32, 0) as ctx:
        kids = [ctx.upload_key(k) for k in keysets]
        jobs = [(kids[0], counts[0], 0), (kids[1], counts[1], counts[0])]
        total = sum(counts)
        d_lwe = ctx.alloc(total * (p32.n + 1) * 4)
        d_v = ctx.alloc(v.nbytes); d_v.upload(v)
        d_out = ctx.alloc(total * (p32.n + 1) * 4)
        d_out2 = ctx.alloc(total * (p32.n + 1) * 4)
        gid = ctx.graph_create(jobs, total, d_lwe, d_v, False, d_out)
        for rep in range(3):
            cts = np.concatenate([sign_batch(keysets[0], counts[0], 10 * rep)[1], sign_batch(keysets[1], counts[1], 10 * rep + 1)[1]])
            d_lwe.upload(cts)
            ctx.graph_launch(gid)
            ctx.sync()
            got = d_out.download(np.zeros((total, p32.n + 1), np.int32))
            # eager work between replays
            ctx.pbs_jobs_device(jobs, total, d_lwe, d_v, False, d_out2)
            ctx.sync()
            want = d_out2.download(np.zeros((total, p32.n + 1), np.int32))
            assert np.array_equal(got, want), f"replay {rep} differs from the eager pipeline"
            # the oracle on the first and last ciphertext of each job
            for (kid, cnt, first), orc in zip(jobs, orcs):
                for c in (first, first + cnt - 1):
                    assert np.array_equal(got[c], orc.pbs(cts[c], v)[2]), f"replay {rep}, ct {c} differs from the oracle"
            ph = yg.lwe_decrypt_phase(got[:counts[0]], keysets[0].lwe_key)
            assert (np.abs(ph.astype(np.int64)) > (1 << 27)).all()
        ctx.graph_destroy(gid)
        for b in (d_lwe, d_v, d_out, d_out2):
            b.free()
    for o in orcs:
        o.close()


def test_graph_survives_later_key_uploads(p32, keys32):
    """Keys uploaded after a graph was captured leave its key tables in place
    (the tables never move), and a key a live graph reads cannot be released."""
    v = yg.sign_test_polynomial(p32)
    bits, cts = sign_batch(keys32, 24, seed=77)
    extra = yg.keygen(p32, seed=401)
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys32)
        d_lwe = ctx.alloc(cts.nbytes); d_lwe.upload(cts)
        d_v = ctx.alloc(v.nbytes); d_v.upload(v)
        d_out = ctx.alloc(len(cts) * (p32.n + 1) * 4)
        gid = ctx.graph_create([(kid, len(cts), 0)], len(cts), d_lwe, d_v, False, d_out)
        # more keys than any earlier table size held
        later = [ctx.upload_key(extra) for _ in range(20)]
        ctx.graph_launch(gid)
        ctx.sync()
        got = d_out.download(np.zeros((len(cts), p32.n + 1), np.int32))
        with pytest.raises(yg.YgpuError) as e:
            ctx.release_key(kid)
        assert e.value.code == yg.YGPU_ERR_BUSY
        ctx.graph_destroy(gid)
        ctx.release_key(kid)
        for k in later:
            ctx.release_key(k)
        for b in (d_lwe, d_v, d_out):
            b.free()
    orc = Oracle(p32, keys32)
    want = orc.pbs_batch(cts, v)
    orc.close()
    assert np.array_equal(got, want)
    assert np.array_equal((yg.lwe_decrypt_phase(got, keys32.lwe_key) > 0).astype(int), bits)


def test_every_tile_is_bit_identical(p32, keys32):
    _, cts = sign_batch(keys32, 37, seed=4)
    v = yg.sign_test_polynomial(p32)
    outs = []
    with yg.Context(p32, 0) as ctx:
        kid = ctx.upload_key(keys32)
        for cpw in (4, 2, 1):
            # twl 2: lane-major twiddle tables (with the digit tables only)
            for twl, tab in [(1, 0), (1, 1), (1, 2), (0, 0), (0, 1), (0, 2), (2, 2)]:
                ctx.set_tile_ex(cpw, twl, tab)
                out, acc, _ = ctx.pbs_batch(kid, cts, v, want_acc=True)
                outs.append((out, acc))
        cpw, twl, ms = ctx.autotune(kid, 512)
        assert cpw in (4, 2, 1) and twl in (0, 1, 2) and ms > 0
        assert ctx.tile() == (cpw, twl)
        assert len(ctx.autotune_results()) == 21
        out, acc, _ = ctx.pbs_batch(kid, cts, v, want_acc=True)
        outs.append((out, acc))
    for o, a in outs[1:]:
        assert np.array_equal(o, outs[0][0]) and np.array_equal(a, outs[0][1])
    orc = Oracle(p32, keys32)
    assert np.array_equal(outs[0][0][:4], orc.pbs_batch(cts[:4], v))
    orc.close()


def test_t56_tiles_bit_identical(p56, keys56):
    rng = np.random.default_rng(8)
    lut = rng.integers(0, 16, 16)
    msgs = rng.integers(0, 16, 6)
    cts = yg.lwe_encrypt(keys56.lwe_key, (msgs * (1 << 32) // 32).astype(np.int64).astype(np.int32), seed=9)
    v = yg.lut_test_polynomial(p56, lut, 16)
    with yg.Context(p56, 0) as ctx:
        kid = ctx.upload_key(keys56)
        assert ctx.tile() == (1, 1)
        assert ctx.tile_ex() == (1, 1, 1)
        # forward twiddles in LDS (1) or read through L1 (0), stage 0 from the
        # digit table (1) or computed (0)
        ctx.set_tile_ex(1, 1, 0)
        a, _, _ = ctx.pbs_batch(kid, cts, v)
        others = []
        for twl, tab in [(1, 1), (0, 1), (0, 0)]:
            ctx.set_tile_ex(1, twl, tab)
            others.append(ctx.pbs_batch(kid, cts, v)[0])
            # a plain call per ciphertext runs the same kernel with a one-ciphertext grid
            others.append(np.stack([ctx.pbs(kid, cts[i], v)[0] for i in range(len(cts))]))
        cpw, twl, ms = ctx.autotune(kid, 64)
        assert (cpw, twl) in [(1, 1), (1, 0)] and ms > 0
        assert len(ctx.autotune_results()) == 4
        others.append(ctx.pbs_batch(kid, cts, v)[0])
        others.append(np.stack([ctx.pbs(kid, cts[i], v)[0] for i in range(len(cts))]))
    for other in others:
        assert np.array_equal(a, other)
    orc = Oracle(p56, keys56)
    assert np.array_equal(a[:2], orc.pbs_batch(cts[:2], v))
    orc.close()
    ph = yg.lwe_decrypt_phase(a, keys56.lwe_key).astype(np.int64)
    assert np.array_equal(np.round(ph / (1 << 32) * 16).astype(np.int64) % 16, lut[msgs] % 16)
