"""The lazy-pipe key format (BootstrappingKeyMPLazyPipe, bootstrapping.h:250-282;
genBootstrappingKeyMPLazyPipe, bootstrapping.cpp:269-320; blindRotateLazyPipeNtt,
blind_rotate.cpp:822-902) on the CPU: the keys ygpu_keygen_lazy writes and the
oracle's restatement of the bootstrap.

Pins, each independent of the oracle's own formulas:
  * the keygen's digits recompose to the torus values of the succinct key's
    half rows for the same seed (component i holding s_{i+1}), within the
    dropped bits, and bskFirst decrypts to v s_0;
  * the per-ciphertext rows by the oracle's NTT route equal the schoolbook
    (coefficient-domain) route, and decrypt to the RGSW rows of
    1 + (X^a - 1) s_{i+1} (row k: that message times the gadget; row 0: minus
    the GLWE key times it);
  * rotated digits narrow to the reference's TORUS32 Decomp (int8): a planted
    digit pair whose rotated difference is +128 yields -128;
  * whole blind rotations: schoolbook route equals NTT route, and the phase
    is X^{-b + sum a_i s_i} v (the reference's BLIND_ROT gate)."""
import ctypes
import dataclasses

import numpy as np
import pytest

import yatfhe_amd as yg
from tests.oracle_py import LazyOracle, glwe_phase, glwe_phase_modq, lib, P


def wrap(x, w=32):
    x = int(x) % (1 << w)
    return x - (1 << w) if x >= (1 << (w - 1)) else x


def centred(r, q):
    r = np.asarray(r).astype(object)
    return np.array([int(x) - q if int(x) > (q - 1) // 2 else int(x) for x in r.reshape(-1)], dtype=object).reshape(r.shape)


SMALL = dataclasses.replace(yg.PARAMS_T32_N1024, n=8, N=64)
# the reference's TORUS32 lazy-pipe shape at N = 1024: base 2^4, 8 digit levels
# (all 32 bits, yatfhe_parameters.h:42-47), 4 row levels (lApprox), GLWE noise
# of its rlweNoiseB = 3 (sigma 2^2.21)
EXACT = dataclasses.replace(yg.PARAMS_T32_N1024, n=64, radix_bits=4, level=4)
EXACT_DIGITS = 8
SIGMA_GLWE_EXACT = 2.21 - 32


@pytest.fixture(scope="module")
def small_keys():
    v = yg.sign_test_polynomial(SMALL)
    return yg.keygen_lazy(SMALL, seed=31, v=v)


@pytest.fixture(scope="module")
def real_keys():
    """The headline gadget 2^7 x 3 (21 of 32 bits: the rows carry the dropped bits)."""
    p = dataclasses.replace(yg.PARAMS_T32_N1024, n=64)
    v = yg.sign_test_polynomial(p)
    return p, yg.keygen_lazy(p, seed=32, v=v)


@pytest.fixture(scope="module")
def exact_keys():
    v = yg.sign_test_polynomial(EXACT)
    return EXACT, yg.keygen_lazy(EXACT, seed=33, v=v, digit_levels=EXACT_DIGITS, sigma_log2_glwe=SIGMA_GLWE_EXACT)


def test_keygen_digits_recompose_succinct_rows(small_keys):
    p = SMALL
    sk = yg.keygen_succinct(p, seed=31, s2_level=p.level)
    lk = small_keys
    assert np.array_equal(lk.lwe_key, sk.lwe_key) and np.array_equal(lk.glwe_key, sk.glwe_key)
    assert np.array_equal(lk.s2, sk.s2) and np.array_equal(lk.ksk, sk.ksk)
    q, w, b, L = p.q_ntt, 32, p.radix_bits, p.level
    g = [1 << (w - (l + 1) * b) for l in range(L)]
    for i in range(p.n - 1):
        for lvl in range(L):
            for c in range(p.k + 1):
                torus = np.array([wrap(x) for x in centred(sk.bsk_half[i + 1, lvl, c], q)], np.int64)
                d = lk.digits_a[i, lvl, :, c] if c < p.k else lk.digits_b[i, lvl]
                rec = sum(d[l].astype(np.int64) * g[l] for l in range(L))
                err = np.array([wrap(x) for x in rec - torus])
                assert np.abs(err).max() <= 1 << (w - L * b - 1), (i, lvl, c)
                assert np.abs(d).max() <= 1 << (b - 1)
    # bskFirst: phase v s_0 up to the GLWE noise
    ph = glwe_phase(p, lk.first.astype(np.int64), lk.glwe_key)
    want = lk.v.astype(np.int64) * int(lk.lwe_key[0])
    assert max(abs(wrap(a - b)) for a, b in zip(ph, want)) < 1 << 12


@pytest.mark.parametrize("a", [0, 1, 5, 63, 64, 65, 100, 127])
def test_rows_schoolbook_equals_ntt(small_keys, a):
    orc = LazyOracle(SMALL, small_keys)
    for comp in (0, SMALL.n - 2):
        assert np.array_equal(orc.rows(comp, a), orc.rows(comp, a, schoolbook=True)), (comp, a)
    orc.close()


def negacyclic_rotate(d, a):
    """X^a d in Z[X]/(X^N + 1), integers (last axis)."""
    d = np.asarray(d, np.int64)
    N = d.shape[-1]
    t = np.arange(N) + (a % (2 * N))
    out = np.zeros_like(d)
    out[..., t % N] = d * np.where((t // N) % 2 == 1, -1, 1)
    return out


def negacyclic_mul(x, z):
    N = len(x)
    out = np.zeros(N, object)
    for j in np.nonzero(x)[0]:
        for t in np.nonzero(z)[0]:
            d = j + t
            if d < N:
                out[d] += int(x[j])
            else:
                out[d - N] -= int(x[j])
    return out


def check_rows_decrypt(p, lk, a, comps, checks):
    """Row k of level lvl decrypts to E = m g_lvl, row 0 to -z E (see the tests)."""
    orc = LazyOracle(p, lk)
    q, N, b = p.q_ntt, p.N, p.radix_bits
    D = lk.digit_levels
    z = lk.glwe_key.astype(np.int64)
    wraps = 0
    for comp in comps:
        s = int(lk.lwe_key[comp + 1])
        m = negacyclic_rotate(np.eye(1, N, 0, np.int64)[0], a) if s else np.eye(1, N, 0, np.int64)[0]
        rows = orc.rows(comp, a)
        for lvl, bound, mask_bound in checks:
            g = [1 << (32 - (l + 1) * b) for l in range(D)]
            corr_b = np.zeros(N, object)
            corr_a = np.zeros(N, object)
            for l in range(D):
                rb_ = negacyclic_rotate(lk.digits_b[comp, lvl, l], a) - lk.digits_b[comp, lvl, l].astype(np.int64)
                if l == lvl:
                    rb_[0] += 1
                ra_ = negacyclic_rotate(lk.digits_a[comp, lvl, l, 0], a) - lk.digits_a[comp, lvl, l, 0].astype(np.int64)
                db_ = ((rb_ + 128) % 256 - 128) - rb_
                da_ = ((ra_ + 128) % 256 - 128) - ra_
                wraps += int(np.count_nonzero(db_) + np.count_nonzero(da_))
                corr_b = corr_b + db_.astype(object) * g[l]
                corr_a = corr_a + da_.astype(object) * g[l]
            E = m.astype(object) * g[lvl] + corr_b - negacyclic_mul(corr_a, z)
            ph = glwe_phase_modq(q, rows[lvl, 1], lk.glwe_key)
            err = [wrap(int(x) - int(y)) for x, y in zip(ph, E)]
            assert max(abs(e) for e in err) < bound, (comp, lvl, "body row")
            if mask_bound:
                ph0 = glwe_phase_modq(q, rows[lvl, 0], lk.glwe_key)
                zE = negacyclic_mul(np.array([wrap(int(x)) for x in E], np.int64), z)
                err0 = [wrap(int(x) + int(y)) for x, y in zip(ph0, zE)]
                assert max(abs(e) for e in err0) < mask_bound, (comp, lvl, "mask row")
    orc.close()
    return wraps


@pytest.mark.parametrize("a", [0, 3, 64, 100, 1500])
def test_rows_decrypt_exact_digits(exact_keys, a):
    """Exact digits (2^4 x 8 = 32 bits): every row level decrypts to the RGSW of
    1 + (X^a - 1) s_{i+1} within g_lvl / 4 (the TRLEV noise through the digits
    is ~2^12)."""
    p, lk = exact_keys
    checks = [(lvl, 1 << (32 - 4 * (lvl + 1) - 2), 1 << (32 - 4 * (lvl + 1) - 2)) for lvl in range(p.level)]
    check_rows_decrypt(p, lk, a, (0, 11, p.n - 2), checks)


@pytest.mark.parametrize("a", [0, 3, 64, 100, 1500])
def test_rows_decrypt_to_rgsw_of_cmux(real_keys, a):
    """Gadget 2^7 x 3: row k of level lvl decrypts to E = m g_lvl, row 0 to
    -z E, for m = 1 + (X^a - 1) s_{i+1} (X^a when the bit is 1, 1 when it is 0),
    plus the reference's int8 narrowing: where a rotated digit X^a d - d (+1 on
    the diagonal) leaves [-128, 127] it wraps by 256, which adds 256 g_l there
    (to the body digits, and through the key to the mask digits), predicted
    here from the stored digits with numpy.  The digits drop 11 bits and the
    s^2 products carry the TRLEV noise, so level 0 is checked on both rows
    (mask row within g_0 / 2: sum_l D_l e_l over 3 N products of digits up to
    128 and noise 2^7, ~2^23.5) and level 1 on the body row (within g_1 / 2:
    the dropped bits times the key, ~2^16).  A wrong rotation, sign or gadget
    slot moves some coefficient by g or more."""
    p, lk = real_keys
    check_rows_decrypt(p, lk, a, (0, 7, 21, p.n - 2), [(0, 1 << 23, 1 << 24), (1, 1 << 17, 0)])


def test_rotation_wraps_occur(real_keys):
    """Over all components the narrowing does happen: digits of +-B/2 = +-64
    occur (ties), so some rotated differences reach +128 and wrap."""
    p, lk = real_keys
    count = 0
    for a in (1, 100, 513, 1500):
        for d in (lk.digits_b.astype(np.int64), lk.digits_a[:, :, :, 0].astype(np.int64)):
            r = negacyclic_rotate(d, a) - d
            count += int(np.count_nonzero((r > 127) | (r < -128)))
    assert count > 0


def test_rotated_digits_narrow_to_int8():
    """A stored digit pair (64 at position 0, -64 at position 1) rotated by
    X^1 - 1 gives 64 - (-64) = +128 at position 1, which the reference's int8
    Decomp holds as -128 (polynomial.cpp:273-282, torus.h:14).  The narrowed
    row differs from the unnarrowed one by 256 X g_l in the body row."""
    p = SMALL
    L, N = p.level, p.N
    zeros = dict(first=np.zeros((2, N), np.int64), da=np.zeros((p.n - 1, L, L, 1, N), np.int32),
                 db=np.zeros((p.n - 1, L, L, N), np.int32), s2=np.zeros((L, 2, N), np.uint64))

    class K:
        pass

    def rows_with(db00):
        k = K()
        k.first, k.digits_a, k.ksk = zeros["first"], zeros["da"], np.zeros(p.ksk_shape, np.int32)
        db = zeros["db"].copy()
        db[0, 0, 2, :2] = db00   # component 0, level 0, digit level 2 (not the diagonal)
        k.digits_b, k.s2 = db, zeros["s2"]
        orc = LazyOracle(p, k)
        r = orc.rows(0, 1, schoolbook=True)
        orc.close()
        return r

    # level 0's body row: the diagonal digit (digit level 0) is 0 + 1 at
    # coefficient 0; digit level 2 rotates [64, -64, 0, ..] by X - 1 into
    # [-64, +128 -> -128, -64, 0, ..]
    r = rows_with([64, -64])
    q = p.q_ntt
    g0, g2 = 1 << (32 - p.radix_bits), 1 << (32 - 3 * p.radix_bits)
    want = np.zeros(N, object)
    want[0], want[1], want[2] = g0 - 64 * g2, -128 * g2, -64 * g2
    body = r[0, 1, 1]                  # level 0, row k (cPrime), column b
    assert [int(x) for x in body] == [int(x) % q for x in want]


def test_blind_rotation_schoolbook_equals_ntt(small_keys):
    orc = LazyOracle(SMALL, small_keys)
    rng = np.random.default_rng(3)
    v = small_keys.v
    for t in range(3):
        abar = rng.integers(0, 2 * SMALL.N, SMALL.n + 1).astype(np.int32)
        if t == 1:
            abar[1::2] = 0                 # skipped steps
        a = orc.blind_rotate_scaled(abar, v)
        b = orc.blind_rotate_scaled(abar, v, schoolbook=True)
        assert np.array_equal(a, b)
    orc.close()


def test_lazy_pbs_phase_and_decrypt(exact_keys):
    """BLIND_ROT gate on exact digits: the accumulator's phase is
    X^{-b + sum a_i s_i} v within the noise; the keyswitched LWE decrypts to the
    sign of the message."""
    p, lk = exact_keys
    orc = LazyOracle(p, lk)
    N = p.N
    rng = np.random.default_rng(4)
    bits = rng.integers(0, 2, 8)
    cts = yg.lwe_encrypt(lk.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=7)
    acc, ext, out = orc.pbs_batch_all(cts, lk.v)
    for c in range(len(cts)):
        sc = np.zeros(p.n + 1, np.int32)
        from tests.oracle_py import or_params
        cp = or_params(p)
        lib().or_rescale(ctypes.byref(cp), P(np.ascontiguousarray(cts[c])), P(sc))
        rot = (-int(sc[-1]) + int((sc[:-1].astype(np.int64) * lk.lwe_key).sum())) % (2 * N)
        want = np.zeros(N, np.int64)
        lib().or_rotate(32, N, rot, P(lk.v.astype(np.int64)), P(want))
        ph = glwe_phase(p, acc[c].astype(np.int64), lk.glwe_key)
        assert max(abs(wrap(a - b)) for a, b in zip(ph, want)) < 1 << 26, c
    ph = yg.lwe_decrypt_phase(out, lk.lwe_key)
    assert np.array_equal((ph > 0).astype(int), bits)
    orc.close()


# ------------------------------------------------------------------ Alt form
# BootstrappingKeyMPLazyPipeAlt / blindRotateLazyPipeAltNtt (blind_rotate.cpp:996-1136),
# the format the reference's serving benchmark runs as OURS.
from tests.oracle_py import AltOracle  # noqa: E402


@pytest.fixture(scope="module")
def alt_small():
    return yg.keygen_lazy_alt(SMALL, seed=41)


@pytest.fixture(scope="module")
def alt_exact():
    return EXACT, yg.keygen_lazy_alt(EXACT, seed=42, digit_levels=EXACT_DIGITS, sigma_log2_glwe=SIGMA_GLWE_EXACT)


def test_alt_keygen_rows_are_succinct_rows_in_torus_form(alt_small):
    sk = yg.keygen_succinct(SMALL, seed=41, s2_level=SMALL.level)
    ak = alt_small
    assert np.array_equal(ak.s2, sk.s2) and np.array_equal(ak.ksk, sk.ksk) and np.array_equal(ak.lwe_key, sk.lwe_key)
    q = SMALL.q_ntt
    torus = np.array([wrap(x) for x in centred(sk.bsk_half, q).reshape(-1)], np.int64).reshape(ak.rows.shape)
    assert np.array_equal(torus, ak.rows.astype(np.int64))


@pytest.mark.parametrize("a", [0, 1, 63, 64, 100])
def test_alt_rows_and_first_schoolbook_equals_ntt(alt_small, a):
    orc = AltOracle(SMALL, alt_small)
    for comp in (1, SMALL.n - 1):
        assert np.array_equal(orc.rows(comp, a), orc.rows(comp, a, schoolbook=True)), (comp, a)
    v = yg.sign_test_polynomial(SMALL)
    assert np.array_equal(orc.first(v), orc.first(v, schoolbook=True))
    orc.close()


@pytest.mark.parametrize("a", [0, 5, 1000])
def test_alt_rows_decrypt_exact_digits(alt_exact, a):
    """Row k of level lvl decrypts to m g_lvl and row 0 to -z m g_lvl, for
    m = 1 + (X^a - 1) s_i; the body is exact (transformed whole) and the mask's
    digits keep all 32 bits, so every level is checked within g_lvl / 4."""
    p, ak = alt_exact
    orc = AltOracle(p, ak)
    q, N = p.q_ntt, p.N
    z = ak.glwe_key.astype(np.int64)
    for comp in (1, 30, p.n - 1):
        s = int(ak.lwe_key[comp])
        m = negacyclic_rotate(np.eye(1, N, 0, np.int64)[0], a) if s else np.eye(1, N, 0, np.int64)[0]
        rows = orc.rows(comp, a)
        for lvl in range(p.level):
            g = 1 << (32 - (lvl + 1) * p.radix_bits)
            ph = glwe_phase_modq(q, rows[lvl, 1], ak.glwe_key)
            assert max(abs(wrap(int(x) - int(y) * g)) for x, y in zip(ph, m)) < g // 4, (comp, lvl, "body row")
            ph0 = glwe_phase_modq(q, rows[lvl, 0], ak.glwe_key)
            zm = negacyclic_mul(m, z)
            assert max(abs(wrap(int(x) + int(y) * g)) for x, y in zip(ph0, zm)) < g // 4, (comp, lvl, "mask row")
    orc.close()


def test_alt_first_component_and_pbs_decrypt(alt_exact):
    """deriveFirstComponentNtt gives a TRLWE of v s_0; whole Alt bootstraps
    decrypt (phase X^{-b + sum a_i s_i} v, and the keyswitched sign)."""
    p, ak = alt_exact
    orc = AltOracle(p, ak)
    N = p.N
    v = yg.sign_test_polynomial(p)
    first = orc.first(v)
    ph = glwe_phase(p, first, ak.glwe_key)
    want = v.astype(np.int64) * int(ak.lwe_key[0])
    assert max(abs(wrap(a - b)) for a, b in zip(ph, want)) < 1 << 20
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2, 8)
    cts = yg.lwe_encrypt(ak.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=9)
    acc, _, out = orc.pbs_batch_all(cts, v)
    from tests.oracle_py import or_params
    for c in range(len(cts)):
        sc = np.zeros(p.n + 1, np.int32)
        lib().or_rescale(ctypes.byref(or_params(p)), P(np.ascontiguousarray(cts[c])), P(sc))
        rot = (-int(sc[-1]) + int((sc[:-1].astype(np.int64) * ak.lwe_key).sum())) % (2 * N)
        want = np.zeros(N, np.int64)
        lib().or_rotate(32, N, rot, P(v.astype(np.int64)), P(want))
        ph = glwe_phase(p, acc[c].astype(np.int64), ak.glwe_key)
        assert max(abs(wrap(a - b)) for a, b in zip(ph, want)) < 1 << 26, c
    assert np.array_equal((yg.lwe_decrypt_phase(out, ak.lwe_key) > 0).astype(int), bits)
    orc.close()


def test_alt_blind_rotation_schoolbook_equals_ntt(alt_small):
    orc = AltOracle(SMALL, alt_small)
    rng = np.random.default_rng(6)
    v = yg.sign_test_polynomial(SMALL)
    for t in range(2):
        abar = rng.integers(0, 2 * SMALL.N, SMALL.n + 1).astype(np.int32)
        if t == 1:
            abar[::2] = 0
        assert np.array_equal(orc.blind_rotate_scaled(abar, v), orc.blind_rotate_scaled(abar, v, schoolbook=True))
    orc.close()
