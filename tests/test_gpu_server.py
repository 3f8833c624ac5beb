"""The serving loop against the oracle: every served result equals the CPU
oracle's bootstrap of the same ciphertext under the tenant's key, alone and in
the mix.

Mirrors the reference's serving benchmark setting (tenants with their own
keys, a key cache smaller than the key pool, both torus widths in one process)
on the GPU server: requests served one at a time, the same requests submitted
together from several threads (batched with other tenants' requests, across
evictions and reloads), and open-loop load whose every result the server
compares with the oracle's."""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import Oracle, derive_succinct

pytestmark = pytest.mark.gpu

P16 = 16


def t32_inputs(keys, count, seed):
    bits = np.random.default_rng(seed).integers(0, 2, count)
    return yg.lwe_encrypt(keys.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=seed)


def t56_inputs(keys, count, seed):
    msgs = np.random.default_rng(seed).integers(0, P16, count)
    return msgs, yg.lwe_encrypt(keys.lwe_key, (msgs * (1 << 32) // (2 * P16)).astype(np.int64).astype(np.int32), seed=seed)


@pytest.fixture(scope="module")
def tenants(p32, p56):
    k32 = [yg.keygen(p32, seed=500 + t) for t in range(5)]
    k56 = [yg.keygen(p56, seed=600 + t) for t in range(2)]
    return k32, k56


def oracle_results(params, keys, cts, v):
    """The oracle's keyswitched outputs (the reference's CPU path restated)."""
    orc = Oracle(params, keys)
    out = orc.pbs_batch(np.atleast_2d(cts), v)
    orc.close()
    return out


class _Derived:
    """A succinct key's full coefficient-form key, as the oracle derives it."""

    def __init__(self, params, keys):
        self.bsk = derive_succinct(params, keys)
        self.ksk = keys.ksk


def test_alone_equals_mix_across_evictions(p32, p56, tenants):
    k32, k56 = tenants
    v32 = yg.sign_test_polynomial(p32)
    lut = np.random.default_rng(7).integers(0, P16, P16)
    v56 = yg.lut_test_polynomial(p56, lut, P16)
    # budget: one TORUS56 key and a few TORUS32 keys, well under the 7-key pool
    budget = 560 << 20
    with yg.Server(0, key_budget_bytes=budget, lanes=3, max_batch=256) as srv:
        ps32 = srv.add_params(p32)
        ps56 = srv.add_params(p56)
        tids = [srv.add_tenant(ps32, k) for k in k32] + [srv.add_tenant(ps56, k) for k in k56]
        work = []  # (tenant, params, keys, cts, v)
        for i, k in enumerate(k32):
            work.append((tids[i], p32, k, t32_inputs(k, 6, 40 + i), v32))
        msgs56 = []
        for j, k in enumerate(k56):
            msgs, cts = t56_inputs(k, 4, 60 + j)
            msgs56.append(msgs)
            work.append((tids[len(k32) + j], p56, k, cts, v56))
        print("server: alone", flush=True)
        alone = {}
        for t, prm, k, cts, v in work:
            for c in range(len(cts)):
                r = srv.request(t, cts[c], v)
                srv.submit(r)
                alone[(t, c)] = srv.wait(r, timeout=60).copy()
                assert r.batch_size == 1
        print("server: mix", flush=True)
        reqs = [(t, c, srv.request(t, cts[c], v)) for t, prm, k, cts, v in work for c in range(len(cts))]
        order = np.random.default_rng(3).permutation(len(reqs))
        chunks = [order[i::4] for i in range(4)]

        def submit(idx):
            for i in idx:
                srv.submit(reqs[i][2])

        ths = [threading.Thread(target=submit, args=(ch,)) for ch in chunks]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        sizes = []
        for t, c, r in reqs:
            got = srv.wait(r, timeout=60)
            sizes.append(r.batch_size)
            assert np.array_equal(got, alone[(t, c)]), f"tenant {t} ciphertext {c}: mix differs from alone"
        st = srv.stats()
        print("server stats", st, "batch sizes", sorted(set(sizes)), flush=True)
        assert st["key_evictions"] > 0 and st["key_loads"] > len(tids)
        assert st["peak_resident_bytes"] <= budget
        assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
        assert max(sizes) > 1, "no request was batched with another"
    # the oracle gives the same bits
    print("server: oracle", flush=True)
    with ThreadPoolExecutor(4) as pool:
        refs = list(pool.map(lambda w: oracle_results(w[1], w[2], w[3], w[4]), work))
    for (t, prm, k, cts, v), ref in zip(work, refs):
        for c in range(len(cts)):
            assert np.array_equal(ref[c], alone[(t, c)]), f"tenant {t} ciphertext {c}: server differs from the oracle"
    # and they decrypt to the intended values
    for i, (t, prm, k, cts, v) in enumerate(work[:len(k32)]):
        ph = yg.lwe_decrypt_phase(np.stack([alone[(t, c)] for c in range(len(cts))]), k.lwe_key)
        assert (np.abs(ph.astype(np.int64)) > (1 << 27)).all()
    for j, (t, prm, k, cts, v) in enumerate(work[len(k32):]):
        ph = yg.lwe_decrypt_phase(np.stack([alone[(t, c)] for c in range(len(cts))]), k.lwe_key).astype(np.int64)
        assert np.array_equal(np.round(ph / (1 << 32) * P16).astype(np.int64) % P16, lut[msgs56[j]] % P16)


def test_loadgen_results_match_and_account(p32, tenants):
    k32, _ = tenants
    v32 = yg.sign_test_polynomial(p32)
    base_cts = [t32_inputs(k, 1, 90 + i)[0] for i, k in enumerate(k32[:2])]
    expected_base = [oracle_results(p32, k, ct, v32)[0] for k, ct in zip(k32[:2], base_cts)]
    with yg.Server(0, key_budget_bytes=6 * 83 << 20, lanes=2, max_batch=512) as srv:
        ps = srv.add_params(p32)
        base = [srv.add_tenant(ps, k) for k in k32[:2]]
        n_t = 24
        for t in range(2, n_t):
            srv.add_tenant_shared(base[t % 2])
        inputs = [base_cts[t % 2] for t in range(n_t)]
        expected = [expected_base[t % 2] for t in range(n_t)]
        print("server: loadgen", flush=True)
        res = srv.loadgen(rate=3000, requests=600, inputs=inputs, luts=[v32] * n_t, expected=expected, warm=32, seed=5)
        print("loadgen", res, flush=True)
        assert res["errors"] == 0 and res["mismatches"] == 0
        assert res["completed"] + res["shed"] == 600
        assert 0 < res["p50_ms"] <= res["p99_ms"] <= res["max_ms"]
        assert res["key_evictions"] > 0
        # loads happened during the run: the reference record's load concurrency
        assert res["load_concurrency_peak"] >= 1 and res["load_concurrency_mean"] > 0
        assert srv.report(res)["load_concurrency_peak"] == res["load_concurrency_peak"]


@pytest.mark.parametrize("sched", ["shared", "affinity", "soft"])
def test_scheduling_modes(p32, tenants, sched):
    """The reference Scheduler's modes with lanes as its workers
    (blindrotate_serving.h:132-273): every result is the same, affinity serves
    each tenant on lane t % lanes only, and the reference's record fields
    account for every request."""
    k32, _ = tenants
    v32 = yg.sign_test_polynomial(p32)
    base_cts = [t32_inputs(k, 1, 95 + i)[0] for i, k in enumerate(k32[:2])]
    expected_base = [oracle_results(p32, k, ct, v32)[0] for k, ct in zip(k32[:2], base_cts)]
    lanes = 3
    with yg.Server(0, lanes=lanes, max_batch=256, sched=sched, spill_threshold=2, steal_threshold=2) as srv:
        ps = srv.add_params(p32)
        base = [srv.add_tenant(ps, k) for k in k32[:2]]
        n_t = 12
        for t in range(2, n_t):
            srv.add_tenant_shared(base[t % 2])
        inputs = [base_cts[t % 2] for t in range(n_t)]
        expected = [expected_base[t % 2] for t in range(n_t)]
        print(f"server: loadgen sched={sched}", flush=True)
        res = srv.loadgen(rate=4000, requests=800, inputs=inputs, luts=[v32] * n_t, expected=expected, warm=24, seed=9)
        rep = srv.report(res)
        print("report", {k: rep[k] for k in ("throughput_rps", "locality", "spilled", "stolen",
                                             "per_worker_requests", "peak_queue_depth", "coalesced_loads")}, flush=True)
    assert res["errors"] == 0 and res["mismatches"] == 0
    assert res["completed"] + res["shed"] == 800
    assert sum(res["lane_requests"]) == res["completed"]
    assert rep["sched"] == sched and rep["shed_plus_completed"] == 800
    assert rep["sojourn"]["p50_ms"] <= rep["sojourn"]["p99_ms"] <= rep["sojourn"]["max_ms"]
    assert 0 < rep["tenant_fairness_jain"] <= 1.0
    if sched == "shared":
        assert res["locality"] == 0 and res["spilled"] == 0 and res["stolen"] == 0
    elif sched == "affinity":
        # tenant t only on lane t % lanes: with 12 tenants every lane has owners
        assert res["locality"] == 1.0 and res["spilled"] == 0 and res["stolen"] == 0
        assert all(c > 0 for c in res["lane_requests"])
    else:
        assert 0 < res["locality"] <= 1.0


def test_succinct_tenants_alone_equal_mix(p32):
    """Tenants holding WWL+24 succinct keys: each load moves the succinct form and
    derives the full rows on the device; results equal the oracle's bootstrap
    under the oracle-derived full key (blindRotateWWL24Ntt's rows), alone and in
    a mix with full-key tenants, across evictions."""
    v32 = yg.sign_test_polynomial(p32)
    full_keys = [yg.keygen(p32, seed=700 + t) for t in range(2)]
    succ_keys = [yg.keygen_succinct(p32, seed=710 + t, s2_level=4) for t in range(3)]
    budget = 3 * 83 << 20  # three of the five keys
    with yg.Server(0, key_budget_bytes=budget, lanes=2, max_batch=128) as srv:
        ps = srv.add_params(p32)
        tids = [srv.add_tenant(ps, k) for k in full_keys] + [srv.add_tenant_succinct(ps, k) for k in succ_keys]
        keys = full_keys + succ_keys
        work = [(t, k, t32_inputs(k, 5, 80 + i)) for i, (t, k) in enumerate(zip(tids, keys))]
        print("server: succinct alone", flush=True)
        alone = {}
        for t, k, cts in work:
            for c in range(len(cts)):
                r = srv.request(t, cts[c], v32)
                srv.submit(r)
                alone[(t, c)] = srv.wait(r, timeout=60).copy()
        loads_alone = srv.stats()
        print("server: succinct mix", flush=True)
        reqs = [(t, c, srv.request(t, cts[c], v32)) for t, k, cts in work for c in range(len(cts))]
        for i in np.random.default_rng(4).permutation(len(reqs)):
            srv.submit(reqs[i][2])
        for t, c, r in reqs:
            assert np.array_equal(srv.wait(r, timeout=60), alone[(t, c)]), f"tenant {t} ciphertext {c}"
        st = srv.stats()
        print("server stats", st, flush=True)
        assert st["key_evictions"] > 0 and st["failed"] == 0
        assert st["peak_resident_bytes"] <= budget
        assert loads_alone["key_loads"] >= len(tids)
        # transfer forms: 49-bit packed residues (+1 word) and the KSK byte planes
        poly_words = p32.N * 49 // 64
        planes = p32.k * p32.N * p32.ks_level * (p32.n + 1) * 4
        full_b = (p32.n * p32.level * 4 * poly_words + 1) * 8 + planes
        succ_b = ((p32.n * p32.level * 2 + 4 * 2) * poly_words + 1) * 8 + planes
        assert st["host_key_bytes"] == 2 * full_b + 3 * succ_b
        assert succ_b < 0.66 * full_b
    print("server: succinct oracle", flush=True)
    for t, k, cts in work:
        okeys = _Derived(p32, k) if isinstance(k, yg.SuccinctKeySet) else k
        ref = oracle_results(p32, okeys, cts, v32)
        for c in range(len(cts)):
            assert np.array_equal(ref[c], alone[(t, c)]), f"tenant {t} ciphertext {c}: server differs from the oracle"


def test_key_file_tier_vs_oracle(tmp_path, p32):
    """Tenants whose keys stay in key files, the tier below pinned host memory
    (the reference's DiskReader behind its serving key cache): one pinned slot
    and device room for two keys, so every round reads the files again.  Every
    served result equals the oracle's bootstrap under the tenant's key (full
    keys, and a succinct key under its oracle-derived full key); a corrupt
    file fails only its own tenant's requests."""
    v32 = yg.sign_test_polynomial(p32)
    keys = [yg.keygen(p32, seed=800 + t) for t in range(2)] + [yg.keygen_succinct(p32, seed=810, s2_level=4)]
    paths = []
    for i, k in enumerate(keys):
        paths.append(str(tmp_path / f"tenant{i}.ygkey"))
        yg.save_keys(paths[-1], k)
    bad = str(tmp_path / "corrupt.ygkey")
    yg.save_keys(bad, keys[0])
    with open(bad, "r+b") as f:  # one bit of a BSK residue, past the 128-byte header
        f.seek(4096)
        b = f.read(1)
        f.seek(4096)
        f.write(bytes([b[0] ^ 1]))
    work = [(k, t32_inputs(k, 4, 90 + i)) for i, k in enumerate(keys)]
    got = {}
    with yg.Server(0, key_budget_bytes=2 * 83 << 20, lanes=1, max_batch=64) as srv:
        ps = srv.add_params(p32)
        srv.set_file_slots(ps, 1)
        tids = [srv.add_tenant_file(ps, path) for path in paths]
        tbad = srv.add_tenant_file(ps, bad)
        print("server: file tier prefetch", flush=True)
        assert srv.prefetch(tids[0])
        for rnd in range(2):
            print(f"server: file tier round {rnd}", flush=True)
            for t, (k, cts) in zip(tids, work):
                for c in range(len(cts)):
                    r = srv.request(t, cts[c], v32)
                    srv.submit(r)
                    got[(rnd, t, c)] = srv.wait(r, timeout=60).copy()
        print("server: corrupt key file", flush=True)
        r = srv.request(tbad, work[0][1][0], v32)
        srv.submit(r)
        with pytest.raises(yg.YgpuError) as err:
            srv.wait(r, timeout=60)
        assert err.value.code == yg.YGPU_ERR_CORRUPT
        r = srv.request(tids[1], work[1][1][0], v32)  # the others are still served
        srv.submit(r)
        assert np.array_equal(srv.wait(r, timeout=60), got[(0, tids[1], 0)])
        st = srv.stats()
        print("server stats", st, flush=True)
        # keys were evicted and loaded again, and with one slot reloads read the files again
        assert st["failed"] == 1 and st["key_evictions"] > 0 and st["peak_resident_bytes"] <= 2 * 83 << 20
        assert st["key_loads"] > len(tids) and st["file_reads"] > len(tids)
        assert st["file_read_bytes"] <= st["bytes_loaded"]
    print("server: file tier oracle", flush=True)
    for t, (k, cts) in zip(tids, work):
        okeys = _Derived(p32, k) if isinstance(k, yg.SuccinctKeySet) else k
        ref = oracle_results(p32, okeys, cts, v32)
        for rnd in range(2):
            for c in range(len(cts)):
                assert np.array_equal(ref[c], got[(rnd, t, c)]), f"tenant {t} round {rnd} ciphertext {c} differs from the oracle"


def test_serving_256_tenants_mixed_widths_vs_oracle(p32, p56):
    """The serving workload's shape: 256 tenants with Zipf(1.1) popularity, 30 %
    of them TORUS56 N=2048 (16-entry LUT) and 70 % TORUS32 N=1024 (sign gate),
    key materials shared round-robin within each width but every tenant cached
    on the device on its own, under a key budget below the key pool (loads from
    pinned host memory, evictions); every result of open-loop Poisson load is
    compared by the server with the oracle's bootstrap of that tenant's input."""
    n_t, mats32, mats56 = 256, 8, 4
    perm = np.random.default_rng(5).permutation(n_t)
    is56 = perm < int(round(0.3 * n_t))
    lut = np.random.default_rng(6).integers(0, P16, P16)
    v32 = yg.sign_test_polynomial(p32)
    v56 = yg.lut_test_polynomial(p56, lut, P16)
    keys32 = [yg.keygen(p32, seed=3000 + m) for m in range(mats32)]
    keys56 = [yg.keygen(p56, seed=4000 + m) for m in range(mats56)]
    in32 = [t32_inputs(k, 1, 5000 + m)[0] for m, k in enumerate(keys32)]
    msg56, in56 = [], []
    for m, k in enumerate(keys56):
        msgs, cts = t56_inputs(k, 1, 6000 + m)
        msg56.append(int(msgs[0]))
        in56.append(cts[0])
    print("serving 256: oracle expectations", flush=True)
    with ThreadPoolExecutor(8) as pool:
        exp32 = list(pool.map(lambda a: oracle_results(p32, a[0], a[1], v32)[0], zip(keys32, in32)))
        exp56 = list(pool.map(lambda a: oracle_results(p56, a[0], a[1], v56)[0], zip(keys56, in56)))
    for m, k in enumerate(keys56):
        ph = int(yg.lwe_decrypt_phase(exp56[m], k.lwe_key)[0])
        assert round(ph / (1 << 32) * P16) % P16 == lut[msg56[m]] % P16
    budget = 12 << 30  # the pool is ~41 GB in device form: loads and evictions
    with yg.Server(0, key_budget_bytes=budget, lanes=3, max_batch=4096, queue_max=1 << 16) as srv:
        ps32 = srv.add_params(p32)
        ps56 = srv.add_params(p56)
        first, inputs, luts, expected = {}, [], [], []
        c32 = c56 = 0
        for t in range(n_t):
            if is56[t]:
                w, m = 56, c56 % mats56
                c56 += 1
            else:
                w, m = 32, c32 % mats32
                c32 += 1
            if (w, m) not in first:
                tid = srv.add_tenant(ps56 if w == 56 else ps32, keys56[m] if w == 56 else keys32[m])
                first[(w, m)] = tid
            else:
                tid = srv.add_tenant_shared(first[(w, m)])
            assert tid == t
            inputs.append(in56[m] if w == 56 else in32[m])
            luts.append(v56 if w == 56 else v32)
            expected.append(exp56[m] if w == 56 else exp32[m])
        for rate in (2000, 10000):
            print(f"serving 256: loadgen {rate}/s", flush=True)
            res = srv.loadgen(rate=rate, requests=rate, inputs=inputs, luts=luts, expected=expected, warm=0,
                              zipf_s=1.1, seed=5, slo_ms=100.0)
            print({k: res[k] for k in ("throughput_rps", "p50_ms", "p99_ms", "hit_rate", "key_loads", "key_evictions",
                                      "mismatches", "errors", "shed")}, flush=True)
            assert res["errors"] == 0 and res["mismatches"] == 0 and res["shed"] == 0
            assert res["completed"] == rate
            assert 0 < res["p50_ms"] <= res["p99_ms"]
        st = srv.stats()
        assert st["key_evictions"] > 0 and st["peak_resident_bytes"] <= budget and st["failed"] == 0
