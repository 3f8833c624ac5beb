"""The serving loop: every result is the same alone and in the mix.

Mirrors the reference's serving benchmark setting (tenants with their own
keys, a key cache smaller than the key pool, both torus widths in one process)
on the GPU server: results of requests served one at a time, of the same
requests submitted together from several threads (batched with other tenants'
requests, across evictions and reloads), and of the single-context batched
entry point (itself pinned to the oracle by the parity tests) must be
bit-identical."""
import threading

import numpy as np
import pytest

import yatfhe_amd as yg

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


def context_results(params, keys, cts, v):
    with yg.Context(params, 0) as ctx:
        kid = ctx.upload_key(keys)
        out, _, _ = ctx.pbs_batch(kid, cts, v)
    return out


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
    # the single-context batched path gives the same bits
    print("server: context", flush=True)
    for t, prm, k, cts, v in work:
        ref = context_results(prm, k, cts, v)
        for c in range(len(cts)):
            assert np.array_equal(ref[c], alone[(t, c)]), f"tenant {t} ciphertext {c}: server differs from context"
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
    expected_base = [context_results(p32, k, ct[None, :], v32)[0] for k, ct in zip(k32[:2], base_cts)]
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
    expected_base = [context_results(p32, k, ct[None, :], v32)[0] for k, ct in zip(k32[:2], base_cts)]
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
    derives the full rows on the device; results equal the single-context
    succinct upload's, alone and in a mix with full-key tenants, across evictions."""
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
    print("server: succinct context", flush=True)
    for t, k, cts in work:
        ref = context_results(p32, k, cts, v32)
        for c in range(len(cts)):
            assert np.array_equal(ref[c], alone[(t, c)]), f"tenant {t} ciphertext {c}: server differs from context"
