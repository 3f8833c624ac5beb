"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle restates the reference's bootstrapping
arithmetic on the CPU and is the checker for the HIP engine.  Used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_DIGITS = {32: os.path.join(ORACLE_DIR, "_ref", "libref_digits32.so"),
              56: os.path.join(ORACLE_DIR, "_ref", "libref_digits56.so")}


class _OrParams(ctypes.Structure):
    _fields_ = [
        ("torus_bits", ctypes.c_int32), ("n", ctypes.c_int32), ("k", ctypes.c_int32), ("N", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32), ("level", ctypes.c_int32), ("ks_radix_bits", ctypes.c_int32),
        ("ks_level", ctypes.c_int32), ("ks_width_bits", ctypes.c_int32), ("lwe_q_bits", ctypes.c_int32),
        ("q_ntt", ctypes.c_uint64), ("extract_index", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)
        L = ctypes.CDLL(ORACLE_LIB)
        i32, i64, u64, vp, P = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p, ctypes.POINTER
        sigs = {
            "or_sext": (i64, [i64, i32]),
            "or_mod_switch_from_torus": (i64, [i64, i64, i64]),
            "or_modswitch_closed_form_check": (i64, [i32, i64, i64]),
            "or_rotate": (None, [i32, i32, i32, vp, vp]),
            "or_rotate_minus_one": (None, [i32, i32, i32, vp, vp]),
            "or_digits_scalar": (None, [i64, i32, i32, i32, vp]),
            "or_digits_row": (None, [vp, i32, i32, i32, i32, vp]),
            "or_digits_cross_check": (i64, [i32, i32, i32, u64, u64, u64, i32]),
            "or_digits_hash": (u64, [i32, i32, i32, u64, u64, u64, i32]),
            "or_negacyclic_mul_modq": (None, [u64, i32, vp, vp, vp, i32]),
            "or_ntt_roundtrip": (i32, [u64, i32, vp]),
            "or_intt_exit": (i64, [u64, u64, i32]),
            "or_key_create": (vp, [P(_OrParams), vp]),
            "or_key_free": (None, [vp]),
            "or_external_product": (None, [vp, i32, vp, vp, i32]),
            "or_rescale": (None, [P(_OrParams), vp, vp]),
            "or_init_acc": (None, [P(_OrParams), vp, i32, vp]),
            "or_blind_rotate": (None, [vp, vp, vp, i32]),
            "or_extract": (None, [P(_OrParams), vp, vp]),
            "or_keyswitch": (None, [P(_OrParams), vp, vp, vp]),
            "or_keyswitch_u32": (None, [P(_OrParams), vp, vp, vp]),
            "or_pbs": (None, [vp, vp, vp, vp, vp, vp, vp, i32]),
            "or_pbs_batch": (None, [vp, vp, i32, vp, vp, vp, i32]),
            "or_pbs_batch_all": (None, [vp, vp, i32, vp, vp, vp, vp, vp, i32]),
            "or_lwe_phase": (i32, [vp, vp, i32]),
            "or_glwe_phase": (None, [i32, i32, i32, vp, vp, vp]),
            "or_glwe_phase_modq": (None, [u64, i32, i32, vp, vp, vp]),
            "or_derive_succinct": (None, [P(_OrParams), i32, vp, vp, vp, i32]),
            "or_pbs_level_parallel": (None, [vp, vp, vp, vp, vp, i32, vp]),
            "or_derive_succinct_row": (None, [P(_OrParams), i32, vp, vp, vp, i32]),
            "or_lazy_key_new": (vp, [P(_OrParams), i32, vp, vp, vp, vp]),
            "or_lazy_key_free": (None, [vp]),
            "or_lazy_rows": (None, [vp, i32, i32, vp, i32]),
            "or_blind_rotate_lazy": (None, [vp, vp, vp, vp, i32]),
            "or_pbs_lazy_batch": (None, [vp, vp, i32, vp, vp, vp, vp, vp, i32]),
            "or_alt_key_new": (vp, [P(_OrParams), i32, vp, vp]),
            "or_alt_key_free": (None, [vp]),
            "or_alt_rows": (None, [vp, i32, i32, vp, i32]),
            "or_alt_first": (None, [vp, vp, vp, i32]),
            "or_blind_rotate_alt": (None, [vp, vp, vp, vp, i32]),
            "or_pbs_alt_batch": (None, [vp, vp, i32, vp, vp, vp, vp, vp, i32]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def P(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def or_params(p, extract_index: int = 0) -> _OrParams:
    return _OrParams(p.torus_bits, p.n, p.k, p.N, p.radix_bits, p.level, p.ks_radix_bits, p.ks_level,
                     p.ks_width_bits, p.lwe_q_bits, p.q_ntt, extract_index, 0)


def ref_digits_lib(width: int):
    path = REF_DIGITS[width]
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    i32, i64, u64, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p
    L.ref_torus_bits.restype = i32
    L.ref_decompose_row.restype = i32
    L.ref_decompose_row.argtypes = [vp, i32, i32, i32, vp]
    L.ref_decompose_ks.restype = None
    L.ref_decompose_ks.argtypes = [i64, i32, i32, i32, vp]
    L.ref_digits_hash.restype = u64
    L.ref_digits_hash.argtypes = [i32, i32, u64, u64, u64, i32]
    return L


class Oracle:
    """Oracle bound to one parameter set and one key set."""

    def __init__(self, params, keys, extract_index: int = 0):
        """extract_index: the coefficient the bootstrap extracts (param.driftPhase)."""
        self.L = lib()
        self.p = params
        self.cp = or_params(params, extract_index)
        self.bsk = np.ascontiguousarray(keys.bsk, dtype=np.uint64)
        self.ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        self.key = self.L.or_key_create(ctypes.byref(self.cp), P(self.bsk))
        if not self.key:
            raise RuntimeError("oracle key creation failed")

    def close(self):
        if self.key:
            self.L.or_key_free(self.key)
            self.key = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _v64(self, v):
        return np.ascontiguousarray(np.asarray(v).astype(np.int64))

    def pbs(self, lwe, v, schoolbook=False):
        p = self.p
        lwe = np.ascontiguousarray(lwe, dtype=np.int32)
        acc = np.zeros((p.k + 1) * p.N, np.int64)
        ext = np.zeros(p.k * p.N + 1, np.int32)
        out = np.zeros(p.n + 1, np.int32)
        self.L.or_pbs(self.key, P(self.ksk), P(lwe), P(self._v64(v)), P(acc), P(ext), P(out), 1 if schoolbook else 0)
        return acc.reshape(p.k + 1, p.N).astype(p.torus_dtype), ext, out

    def pbs_batch(self, lwes, v, threads=None):
        p = self.p
        lwes = np.ascontiguousarray(lwes, dtype=np.int32)
        out = np.zeros((lwes.shape[0], p.n + 1), np.int32)
        t = threads or min(16, os.cpu_count() or 1)
        self.L.or_pbs_batch(self.key, P(self.ksk), lwes.shape[0], P(lwes), P(self._v64(v)), P(out), t)
        return out

    def pbs_level_parallel(self, lwe, v, reps=1):
        """(keyswitched LWE, seconds per run) with the external product's levels
        on a pool of `level` workers, as the reference threads it."""
        out = np.zeros(self.p.n + 1, np.int32)
        secs = np.zeros(reps, np.float64)
        self.L.or_pbs_level_parallel(self.key, P(self.ksk), P(np.ascontiguousarray(lwe, dtype=np.int32)), P(self._v64(v)),
                                     P(out), reps, P(secs))
        return out, secs

    def pbs_batch_all(self, lwes, v, threads=None):
        """(accumulators, extracted LWEs, keyswitched LWEs) of many ciphertexts."""
        p = self.p
        lwes = np.ascontiguousarray(lwes, dtype=np.int32)
        c = lwes.shape[0]
        out = np.zeros((c, p.n + 1), np.int32)
        acc = np.zeros((c, p.k + 1, p.N), np.int64)
        ext = np.zeros((c, p.k * p.N + 1), np.int32)
        t = threads or min(16, os.cpu_count() or 1)
        self.L.or_pbs_batch_all(self.key, P(self.ksk), c, P(lwes), P(self._v64(v)), P(out), P(acc), P(ext), t)
        return acc.astype(p.torus_dtype), ext, out

    def rescale(self, lwe):
        out = np.zeros(self.p.n + 1, np.int32)
        self.L.or_rescale(ctypes.byref(self.cp), P(np.ascontiguousarray(lwe, dtype=np.int32)), P(out))
        return out

    def init_acc(self, v, bbar):
        acc = np.zeros((self.p.k + 1) * self.p.N, np.int64)
        self.L.or_init_acc(ctypes.byref(self.cp), P(self._v64(v)), int(bbar), P(acc))
        return acc

    def blind_rotate(self, lwe, v, schoolbook=False):
        sc = self.rescale(lwe)
        acc = self.init_acc(v, sc[-1])
        self.L.or_blind_rotate(self.key, P(sc), P(acc), 1 if schoolbook else 0)
        return acc.reshape(self.p.k + 1, self.p.N).astype(self.p.torus_dtype)

    def external_product(self, i, acc, schoolbook=False):
        p = self.p
        a = np.ascontiguousarray(np.asarray(acc).astype(np.int64).reshape(-1))
        out = np.zeros_like(a)
        self.L.or_external_product(self.key, int(i), P(a), P(out), 1 if schoolbook else 0)
        return out.reshape(p.k + 1, p.N)

    def extract(self, acc):
        a = np.ascontiguousarray(np.asarray(acc).astype(np.int64).reshape(-1))
        out = np.zeros(self.p.k * self.p.N + 1, np.int32)
        self.L.or_extract(ctypes.byref(self.cp), P(a), P(out))
        return out

    def keyswitch(self, ext, u32=False):
        out = np.zeros(self.p.n + 1, np.int32)
        f = self.L.or_keyswitch_u32 if u32 else self.L.or_keyswitch
        f(ctypes.byref(self.cp), P(self.ksk), P(np.ascontiguousarray(ext, dtype=np.int32)), P(out))
        return out


class LazyOracle:
    """The oracle's blindRotateLazyPipeNtt (+ extract and keyswitch) for one lazy-pipe key set."""

    def __init__(self, params, keys, extract_index: int = 0):
        self.L = lib()
        self.p = params
        self.cp = or_params(params, extract_index)
        self.first = np.ascontiguousarray(np.asarray(keys.first).astype(np.int64))
        self.da = np.ascontiguousarray(np.asarray(keys.digits_a).astype(np.int32))
        self.db = np.ascontiguousarray(np.asarray(keys.digits_b).astype(np.int32))
        self.s2 = np.ascontiguousarray(keys.s2, dtype=np.uint64)
        self.ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        self.digit_levels = int(np.asarray(keys.digits_b).shape[2])
        self.key = self.L.or_lazy_key_new(ctypes.byref(self.cp), self.digit_levels, P(self.first), P(self.da), P(self.db),
                                          P(self.s2))
        if not self.key:
            raise RuntimeError("oracle lazy-pipe key creation failed")

    def close(self):
        if self.key:
            self.L.or_lazy_key_free(self.key)
            self.key = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rows(self, comp, a, schoolbook=False):
        """Coefficient-form rows [l][k+1][k+1][N] of component comp (s_{comp+1}) rotated by a."""
        p = self.p
        out = np.zeros((p.level, p.k + 1, p.k + 1, p.N), np.uint64)
        self.L.or_lazy_rows(self.key, int(comp), int(a), P(out), 1 if schoolbook else 0)
        return out

    def blind_rotate_scaled(self, abar, v, schoolbook=False):
        p = self.p
        acc = np.zeros((p.k + 1) * p.N, np.int64)
        self.L.or_blind_rotate_lazy(self.key, P(np.ascontiguousarray(abar, dtype=np.int32)),
                                    P(np.ascontiguousarray(np.asarray(v).astype(np.int64))), P(acc), 1 if schoolbook else 0)
        return acc.reshape(p.k + 1, p.N).astype(p.torus_dtype)

    def pbs_batch_all(self, lwes, v, threads=None):
        """(accumulators, extracted LWEs, keyswitched LWEs) of many ciphertexts."""
        p = self.p
        lwes = np.ascontiguousarray(lwes, dtype=np.int32)
        c = lwes.shape[0]
        out = np.zeros((c, p.n + 1), np.int32)
        acc = np.zeros((c, p.k + 1, p.N), np.int64)
        ext = np.zeros((c, p.k * p.N + 1), np.int32)
        t = threads or min(16, os.cpu_count() or 1)
        self.L.or_pbs_lazy_batch(self.key, P(self.ksk), c, P(lwes), P(np.ascontiguousarray(np.asarray(v).astype(np.int64))),
                                 P(acc), P(ext), P(out), t)
        return acc.astype(p.torus_dtype), ext, out


class AltOracle:
    """The oracle's blindRotateLazyPipeAltNtt (+ extract and keyswitch) for one Alt key set."""

    def __init__(self, params, keys, extract_index: int = 0):
        self.L = lib()
        self.p = params
        self.cp = or_params(params, extract_index)
        self.rows_t = np.ascontiguousarray(np.asarray(keys.rows).astype(np.int64))
        self.s2 = np.ascontiguousarray(keys.s2, dtype=np.uint64)
        self.ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        self.key = self.L.or_alt_key_new(ctypes.byref(self.cp), int(self.s2.shape[0]), P(self.rows_t), P(self.s2))
        if not self.key:
            raise RuntimeError("oracle Alt key creation failed")

    def close(self):
        if self.key:
            self.L.or_alt_key_free(self.key)
            self.key = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rows(self, comp, a, schoolbook=False):
        """Coefficient-form rows [R][k+1][k+1][N] of s_comp rotated by a."""
        p = self.p
        out = np.zeros((p.level, p.k + 1, p.k + 1, p.N), np.uint64)
        self.L.or_alt_rows(self.key, int(comp), int(a), P(out), 1 if schoolbook else 0)
        return out

    def first(self, v, schoolbook=False):
        p = self.p
        out = np.zeros((p.k + 1) * p.N, np.int64)
        self.L.or_alt_first(self.key, P(np.ascontiguousarray(np.asarray(v).astype(np.int64))), P(out), 1 if schoolbook else 0)
        return out.reshape(p.k + 1, p.N)

    def blind_rotate_scaled(self, abar, v, schoolbook=False):
        p = self.p
        acc = np.zeros((p.k + 1) * p.N, np.int64)
        self.L.or_blind_rotate_alt(self.key, P(np.ascontiguousarray(abar, dtype=np.int32)),
                                   P(np.ascontiguousarray(np.asarray(v).astype(np.int64))), P(acc), 1 if schoolbook else 0)
        return acc.reshape(p.k + 1, p.N).astype(p.torus_dtype)

    def pbs_batch_all(self, lwes, v, threads=None):
        p = self.p
        lwes = np.ascontiguousarray(lwes, dtype=np.int32)
        c = lwes.shape[0]
        out = np.zeros((c, p.n + 1), np.int32)
        acc = np.zeros((c, p.k + 1, p.N), np.int64)
        ext = np.zeros((c, p.k * p.N + 1), np.int32)
        t = threads or min(16, os.cpu_count() or 1)
        self.L.or_pbs_alt_batch(self.key, P(self.ksk), c, P(lwes), P(np.ascontiguousarray(np.asarray(v).astype(np.int64))),
                                P(acc), P(ext), P(out), t)
        return acc.astype(p.torus_dtype), ext, out


def glwe_phase(params, acc, glwe_key):
    out = np.zeros(params.N, np.int64)
    a = np.ascontiguousarray(np.asarray(acc).astype(np.int64).reshape(-1))
    lib().or_glwe_phase(params.torus_bits, params.k, params.N, P(a), P(np.ascontiguousarray(glwe_key, dtype=np.int32)), P(out))
    return out


def derive_succinct(params, keys, threads=8):
    """Full coefficient-form BSK [n][l][k+1][k+1][N] of a succinct key (oracle NTT route)."""
    k1 = params.k + 1
    full = np.zeros((params.n, params.level, k1, k1, params.N), np.uint64)
    half = np.ascontiguousarray(keys.bsk_half, dtype=np.uint64)
    s2 = np.ascontiguousarray(keys.s2, dtype=np.uint64)
    cp = or_params(params)
    lib().or_derive_succinct(ctypes.byref(cp), s2.shape[0], P(half), P(s2), P(full), threads)
    return full


def derive_succinct_row(params, keys, i, lvl, schoolbook=False):
    """Rows [k+1][k+1][N] of step i, level lvl, by the NTT route or schoolbook."""
    k1 = params.k + 1
    out = np.zeros((k1, k1, params.N), np.uint64)
    cp_row = np.ascontiguousarray(keys.bsk_half[i, lvl], dtype=np.uint64)
    s2 = np.ascontiguousarray(keys.s2, dtype=np.uint64)
    cp = or_params(params)
    lib().or_derive_succinct_row(ctypes.byref(cp), s2.shape[0], P(cp_row), P(s2), P(out), 1 if schoolbook else 0)
    return out


def glwe_phase_modq(q, row, glwe_key):
    """(b - sum a_i z_i) mod q, centred, of a coefficient-form row [k+1][N]."""
    row = np.ascontiguousarray(row, dtype=np.uint64)
    k = row.shape[0] - 1
    N = row.shape[1]
    out = np.zeros(N, np.int64)
    lib().or_glwe_phase_modq(q, k, N, P(row), P(np.ascontiguousarray(glwe_key, dtype=np.int32)), P(out))
    return out
