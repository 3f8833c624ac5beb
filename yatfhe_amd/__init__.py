"""yatfhe_amd -- MI355X TFHE programmable-bootstrapping engine, Python binding.

The engine is the C-ABI library ``libyatfhe_gpu.so`` (declared in
``include/yatfhe_gpu.h``) built from ``yatfhe_amd/csrc``.  This module loads it
with ctypes and offers thin wrappers used by the tests and the benchmark.  There
is no CPU fallback: if the library is missing, importing the binding raises.

Parameter presets follow the reference's ``YatfheParameters``
(yatfhe_parameters.h:9-110); see ``PARAMS_T32_N1024`` and ``PARAMS_T56_N2048``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, replace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyatfhe_gpu.so")

Q_49P_T32 = 562941363486721  # torus.h:151
Q_61P_T56 = 1945555039024054273  # torus.h:157

YGPU_OK = 0
YGPU_ERR_BUSY = -7
YGPU_ERR_IO = -8
YGPU_ERR_CORRUPT = -9
YGPU_ERR_DEVICE = -10


class YgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ygpu error {code}: {msg}")
        self.code = code


class _Params(ctypes.Structure):
    _fields_ = [
        ("torus_bits", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("k", ctypes.c_int32),
        ("N", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32),
        ("level", ctypes.c_int32),
        ("ks_radix_bits", ctypes.c_int32),
        ("ks_level", ctypes.c_int32),
        ("ks_width_bits", ctypes.c_int32),
        ("lwe_q_bits", ctypes.c_int32),
        ("q_ntt", ctypes.c_uint64),
    ]


class _Job(ctypes.Structure):
    _fields_ = [("key_id", ctypes.c_int32), ("count", ctypes.c_int32), ("first", ctypes.c_int64)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("lanes", ctypes.c_int32), ("max_batch", ctypes.c_int32),
                ("queue_max", ctypes.c_int32), ("max_tenants", ctypes.c_int32), ("max_loads", ctypes.c_int32),
                ("eviction", ctypes.c_int32), ("reserved", ctypes.c_int32), ("key_budget_bytes", ctypes.c_int64),
                ("sched", ctypes.c_int32), ("spill_threshold", ctypes.c_int32), ("steal_threshold", ctypes.c_int32),
                ("reserved2", ctypes.c_int32)]

# Scheduling modes of the server (the reference serving Scheduler's names).
SCHED_MODES = {"shared": 0, "affinity": 1, "soft": 2}


class _Request(ctypes.Structure):
    _fields_ = [("tenant", ctypes.c_int32), ("batch_size", ctypes.c_int32), ("lwe_in", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("lwe_out", ctypes.c_void_p), ("arrival_ns", ctypes.c_int64),
                ("launch_ns", ctypes.c_int64), ("done_ns", ctypes.c_int64), ("seq", ctypes.c_uint64),
                ("status", ctypes.c_int32), ("key_hit", ctypes.c_int32), ("lane", ctypes.c_int32),
                ("coalesced", ctypes.c_int32)]


_STAT_FIELDS = ["submitted", "completed", "failed", "shed", "batches", "batched_cts", "request_hits", "key_loads",
                "key_evictions", "bytes_loaded", "memory_waits", "resident_keys", "resident_bytes",
                "peak_resident_bytes", "key_budget_bytes", "host_key_bytes", "coalesced", "spilled", "stolen",
                "peak_queue_depth", "loads_in_flight", "peak_loads_in_flight", "load_inflight_ns",
                "file_reads", "file_read_bytes", "file_read_ns"]


class _ServerStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in _STAT_FIELDS]


class _LoadgenConfig(ctypes.Structure):
    _fields_ = [("rate", ctypes.c_double), ("requests", ctypes.c_int64), ("warm", ctypes.c_int64),
                ("zipf_s", ctypes.c_double), ("seed", ctypes.c_uint64), ("n_tenants", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("lwe_in", ctypes.POINTER(ctypes.c_void_p)),
                ("v", ctypes.POINTER(ctypes.c_void_p)), ("expected", ctypes.POINTER(ctypes.c_void_p)),
                ("slo_ms", ctypes.c_double)]


class _LoadgenResult(ctypes.Structure):
    _fields_ = ([(f, ctypes.c_double) for f in ("offered_rps", "throughput_rps", "wall_sec")]
                + [(f, ctypes.c_int64) for f in ("completed", "shed", "errors", "mismatches")]
                + [(f, ctypes.c_double) for f in ("mean_ms", "p50_ms", "p90_ms", "p99_ms", "p99_9_ms", "max_ms",
                                                   "queue_p50_ms", "queue_p99_ms", "slo_violation_rate")]
                + [("batches", ctypes.c_int64), ("mean_batch", ctypes.c_double), ("hit_rate", ctypes.c_double)]
                + [(f, ctypes.c_int64) for f in ("key_loads", "key_evictions", "bytes_loaded", "memory_waits")]
                + [(f, ctypes.c_double) for f in ("hit_p50_ms", "hit_p99_ms", "miss_p50_ms", "miss_p99_ms")]
                + [("pset_p50_ms", ctypes.c_double * 4), ("pset_p99_ms", ctypes.c_double * 4)]
                + [(f, ctypes.c_double) for f in ("sd_ms", "queue_mean_ms", "queue_p90_ms", "queue_max_ms",
                                                   "service_mean_ms", "service_p50_ms", "service_p90_ms",
                                                   "service_p99_ms", "service_max_ms", "slo_violation_rate_100",
                                                   "slo_violation_rate_300", "slo_violation_rate_500")]
                + [(f, ctypes.c_int64) for f in ("coalesced_loads", "peak_queue_depth", "spilled", "stolen")]
                + [("load_concurrency_mean", ctypes.c_double), ("load_concurrency_peak", ctypes.c_int64)]
                + [("tenant_fairness_jain", ctypes.c_double), ("locality", ctypes.c_double),
                   ("lane_requests", ctypes.c_int64 * 16)])


@dataclass(frozen=True)
class Params:
    """Bootstrapping parameter set (the fields of YatfheParameters the path reads)."""

    torus_bits: int
    n: int
    k: int
    N: int
    radix_bits: int
    level: int
    ks_radix_bits: int
    ks_level: int
    ks_width_bits: int = 32
    lwe_q_bits: int = 32
    q_ntt: int = Q_49P_T32

    def c(self) -> _Params:
        return _Params(self.torus_bits, self.n, self.k, self.N, self.radix_bits, self.level,
                       self.ks_radix_bits, self.ks_level, self.ks_width_bits, self.lwe_q_bits, self.q_ntt)

    @property
    def torus_dtype(self):
        return np.int32 if self.torus_bits == 32 else np.int64

    @property
    def bsk_shape(self):
        k1 = self.k + 1
        return (self.n, self.level, k1, k1, self.N)

    @property
    def ksk_shape(self):
        return (self.k * self.N, self.ks_level, self.n + 1)


# TORUS32, n=630, k=1, N=1024, gadget 2^7 x 3, keyswitch 2^2 x 8.
PARAMS_T32_N1024 = Params(torus_bits=32, n=630, k=1, N=1024, radix_bits=7, level=3,
                          ks_radix_bits=2, ks_level=8, q_ntt=Q_49P_T32)
# TORUS56, n=742, k=1, N=2048, gadget 2^10 x 3, keyswitch 2^2 x 8.
PARAMS_T56_N2048 = Params(torus_bits=56, n=742, k=1, N=2048, radix_bits=10, level=3,
                          ks_radix_bits=2, ks_level=8, q_ntt=Q_61P_T56)

_lib = None


def load_library(path: str | None = None):
    """Load the engine library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build it with yatfhe_amd.build.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(p)
    c_int, c_i32, c_i64, c_u64, c_vp, c_dbl = ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_double
    P = ctypes.POINTER
    sigs = {
        "ygpu_last_error": (ctypes.c_char_p, []),
        "ygpu_device_count": (c_int, [P(c_i32)]),
        "ygpu_abi_struct_sizes": (c_int, [P(c_i64), c_i32]),
        "ygpu_init": (c_int, [P(_Params), c_i32, P(c_vp)]),
        "ygpu_destroy": (c_int, [c_vp]),
        "ygpu_get_params": (c_int, [c_vp, P(_Params)]),
        "ygpu_key_upload": (c_int, [c_vp, c_vp, c_vp, P(c_i32)]),
        "ygpu_key_upload_succinct": (c_int, [c_vp, c_vp, c_vp, c_i32, c_vp, P(c_i32)]),
        "ygpu_key_release": (c_int, [c_vp, c_i32]),
        "ygpu_key_bytes": (c_int, [c_vp, c_i32, P(c_i64)]),
        "ygpu_pbs": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
        "ygpu_blind_rotate": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp]),
        "ygpu_extract_keyswitch": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp]),
        "ygpu_blind_rotate_scaled": (c_int, [c_vp, c_i32, c_vp, c_vp]),
        "ygpu_keyswitch": (c_int, [c_vp, c_i32, c_vp, c_vp]),
        "ygpu_sample_extract": (c_int, [c_vp, c_vp, c_vp]),
        "ygpu_pbs_batch_device": (c_int, [c_vp, c_i32, c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
        "ygpu_pbs_batch_device_noalloc": (c_int, [c_vp, c_i32, c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
        "ygpu_pbs_jobs_device": (c_int, [c_vp, P(_Job), c_i32, c_i64, c_vp, c_vp, c_i32, c_vp, c_vp]),
        "ygpu_pbs_plain_device": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp]),
        "ygpu_malloc": (c_int, [c_vp, ctypes.c_size_t, P(c_vp)]),
        "ygpu_free": (c_int, [c_vp, c_vp]),
        "ygpu_memcpy_htod": (c_int, [c_vp, c_vp, c_vp, ctypes.c_size_t]),
        "ygpu_memcpy_dtoh": (c_int, [c_vp, c_vp, c_vp, ctypes.c_size_t]),
        "ygpu_sync": (c_int, [c_vp]),
        "ygpu_timing_enable": (c_int, [c_vp, c_i32]),
        "ygpu_timing_get": (c_int, [c_vp, ctypes.c_char_p, P(c_i64), P(c_dbl)]),
        "ygpu_timing_reset": (c_int, [c_vp]),
        "ygpu_modmul_peak": (c_int, [c_vp, c_i32, P(c_dbl)]),
        "ygpu_get_tile": (c_int, [c_vp, P(c_i32), P(c_i32)]),
        "ygpu_set_tile": (c_int, [c_vp, c_i32, c_i32]),
        "ygpu_autotune": (c_int, [c_vp, c_i32, c_i32, P(c_i32), P(c_i32), P(c_dbl)]),
        "ygpu_get_tile_ex": (c_int, [c_vp, P(c_i32), P(c_i32), P(c_i32)]),
        "ygpu_set_tile_ex": (c_int, [c_vp, c_i32, c_i32, c_i32]),
        "ygpu_autotune_results": (c_int, [c_vp, c_i32, P(c_i32), P(c_i32), P(c_i32), P(c_dbl), P(c_i32)]),
        "ygpu_graph_create": (c_int, [c_vp, P(_Job), c_i32, c_i64, c_vp, c_vp, c_i32, c_vp, c_vp, P(c_i32)]),
        "ygpu_graph_launch": (c_int, [c_vp, c_i32]),
        "ygpu_graph_destroy": (c_int, [c_vp, c_i32]),
        "ygpu_copy_device": (c_int, [c_vp, c_vp, c_vp, ctypes.c_size_t]),
        "ygpu_set_barrier_polls": (c_int, [c_vp, c_i32]),
        "ygpu_set_extract_index": (c_int, [c_vp, c_i32]),
        "ygpu_get_extract_index": (c_int, [c_vp, P(c_i32)]),
        "ygpu_sample_extract_index": (c_int, [c_vp, c_vp, c_i32, c_vp]),
        "ygpu_server_set_extract_index": (c_int, [c_vp, c_i32, c_i32]),
        "ygpu_server_lanes": (c_int, [c_vp, c_i32, P(c_i32)]),
        "ygpu_keygen": (c_int, [P(_Params), c_u64, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_i32]),
        "ygpu_keygen_classic": (c_int, [P(_Params), c_u64, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_i32]),
        "ygpu_keygen_succinct": (c_int, [P(_Params), c_u64, c_dbl, c_dbl, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32]),
        "ygpu_keygen_lazy": (c_int, [P(_Params), c_u64, c_dbl, c_dbl, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32]),
        "ygpu_key_upload_lazy": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, P(c_i32)]),
        "ygpu_pbs_lazy": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
        "ygpu_blind_rotate_lazy": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp]),
        "ygpu_pbs_lazy_batch_device": (c_int, [c_vp, c_i32, c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
        "ygpu_set_lazy_window": (c_int, [c_vp, c_i32]),
        "ygpu_keygen_lazy_alt": (c_int, [P(_Params), c_u64, c_dbl, c_dbl, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32]),
        "ygpu_key_upload_lazy_alt": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, P(c_i32)]),
        "ygpu_keygen_lazy_tuniform": (c_int, [P(_Params), c_u64, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                              c_vp, c_i32]),
        "ygpu_keygen_lazy_alt_tuniform": (c_int, [P(_Params), c_u64, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32]),
        "ygpu_lwe_encrypt": (c_int, [c_i32, c_vp, c_u64, c_dbl, c_i32, c_vp, c_vp]),
        "ygpu_lwe_encrypt_tuniform": (c_int, [c_i32, c_vp, c_u64, c_i32, c_i32, c_vp, c_vp]),
        "ygpu_keygen_tuniform": (c_int, [P(_Params), c_u64, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32]),
        "ygpu_server_create": (c_int, [P(_ServerConfig), P(c_vp)]),
        "ygpu_server_destroy": (c_int, [c_vp]),
        "ygpu_server_add_params": (c_int, [c_vp, P(_Params), P(c_i32)]),
        "ygpu_server_set_tile": (c_int, [c_vp, c_i32, c_i32, c_i32]),
        "ygpu_server_add_tenant": (c_int, [c_vp, c_i32, c_vp, c_vp, P(c_i32)]),
        "ygpu_server_add_tenant_shared": (c_int, [c_vp, c_i32, P(c_i32)]),
        "ygpu_server_add_tenant_file": (c_int, [c_vp, c_i32, ctypes.c_char_p, P(c_i32)]),
        "ygpu_server_set_file_slots": (c_int, [c_vp, c_i32, c_i32]),
        "ygpu_server_add_tenant_succinct": (c_int, [c_vp, c_i32, c_vp, c_vp, c_i32, c_vp, P(c_i32)]),
        "ygpu_server_tenant_params": (c_int, [c_vp, c_i32, P(c_i32)]),
        "ygpu_server_prefetch": (c_int, [c_vp, c_i32]),
        "ygpu_server_submit": (c_int, [c_vp, P(_Request)]),
        "ygpu_server_wait": (c_int, [c_vp, P(_Request)]),
        "ygpu_server_wait_for": (c_int, [c_vp, P(_Request), c_i32]),
        "ygpu_server_stats_get": (c_int, [c_vp, P(_ServerStats)]),
        "ygpu_server_loadgen": (c_int, [c_vp, P(_LoadgenConfig), P(_LoadgenResult)]),
        "ygpu_key_file_write": (c_int, [ctypes.c_char_p, P(_Params), c_vp, c_vp]),
        "ygpu_key_file_info": (c_int, [ctypes.c_char_p, P(_Params)]),
        "ygpu_key_file_read": (c_int, [ctypes.c_char_p, P(_Params), c_vp, c_vp]),
        "ygpu_key_upload_file": (c_int, [c_vp, ctypes.c_char_p, P(c_i32)]),
        "ygpu_key_file_write_succinct": (c_int, [ctypes.c_char_p, P(_Params), c_vp, c_vp, c_i32, c_vp]),
        "ygpu_key_file_form": (c_int, [ctypes.c_char_p, P(c_i32)]),
        "ygpu_key_file_read_succinct": (c_int, [ctypes.c_char_p, P(_Params), c_vp, c_vp, c_vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def exported_symbols():
    return [
        "ygpu_last_error", "ygpu_device_count", "ygpu_init", "ygpu_destroy", "ygpu_get_params",
        "ygpu_key_upload", "ygpu_key_release", "ygpu_key_bytes", "ygpu_pbs", "ygpu_blind_rotate",
        "ygpu_extract_keyswitch", "ygpu_pbs_batch_device", "ygpu_pbs_jobs_device", "ygpu_pbs_plain_device",
        "ygpu_malloc", "ygpu_free", "ygpu_memcpy_htod", "ygpu_memcpy_dtoh", "ygpu_sync",
        "ygpu_timing_enable", "ygpu_timing_get", "ygpu_timing_reset", "ygpu_keygen", "ygpu_lwe_encrypt",
        "ygpu_modmul_peak", "ygpu_copy_device", "ygpu_get_tile", "ygpu_set_tile", "ygpu_autotune",
        "ygpu_graph_create", "ygpu_graph_launch", "ygpu_graph_destroy", "ygpu_blind_rotate_scaled", "ygpu_keyswitch",
        "ygpu_sample_extract", "ygpu_server_create", "ygpu_server_destroy", "ygpu_server_add_params",
        "ygpu_server_set_tile", "ygpu_server_add_tenant", "ygpu_server_add_tenant_shared", "ygpu_server_tenant_params",
        "ygpu_server_prefetch",
        "ygpu_server_submit", "ygpu_server_wait", "ygpu_server_wait_for", "ygpu_server_stats_get", "ygpu_server_loadgen",
        "ygpu_key_file_write", "ygpu_key_file_info", "ygpu_key_file_read", "ygpu_key_upload_file",
        "ygpu_key_upload_succinct", "ygpu_keygen_succinct", "ygpu_server_add_tenant_succinct",
        "ygpu_key_file_write_succinct", "ygpu_key_file_form", "ygpu_key_file_read_succinct", "ygpu_keygen_classic",
        "ygpu_abi_struct_sizes", "ygpu_set_barrier_polls", "ygpu_pbs_batch_device_noalloc",
        "ygpu_set_extract_index", "ygpu_get_extract_index", "ygpu_sample_extract_index", "ygpu_server_set_extract_index", "ygpu_server_lanes",
        "ygpu_server_add_tenant_file", "ygpu_server_set_file_slots",
        "ygpu_keygen_lazy", "ygpu_key_upload_lazy", "ygpu_pbs_lazy", "ygpu_blind_rotate_lazy", "ygpu_pbs_lazy_batch_device",
        "ygpu_set_lazy_window", "ygpu_keygen_lazy_alt", "ygpu_key_upload_lazy_alt", "ygpu_keygen_lazy_tuniform",
        "ygpu_keygen_lazy_alt_tuniform",
        "ygpu_keygen_tuniform", "ygpu_lwe_encrypt_tuniform", "ygpu_get_tile_ex", "ygpu_set_tile_ex", "ygpu_autotune_results",
    ]


def abi_struct_mismatches():
    """(name, library size, binding size) of every ABI struct whose ctypes
    mirror disagrees with the loaded library (empty when they all agree)."""
    mirrors = [("ygpu_params", _Params), ("ygpu_job", _Job), ("ygpu_server_config", _ServerConfig),
               ("ygpu_request", _Request), ("ygpu_server_stats", _ServerStats),
               ("ygpu_loadgen_config", _LoadgenConfig), ("ygpu_loadgen_result", _LoadgenResult)]
    sizes = (ctypes.c_int64 * len(mirrors))()
    n = load_library().ygpu_abi_struct_sizes(sizes, len(mirrors))
    if n != len(mirrors):
        return [("struct count", n, len(mirrors))]
    return [(name, sizes[i], ctypes.sizeof(cls)) for i, (name, cls) in enumerate(mirrors)
            if sizes[i] != ctypes.sizeof(cls)]


def _check(rc):
    if rc != YGPU_OK:
        msg = load_library().ygpu_last_error()
        raise YgpuError(rc, msg.decode() if msg else "")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------------- keys
@dataclass
class KeySet:
    params: Params
    lwe_key: np.ndarray   # int32[n]
    glwe_key: np.ndarray  # int32[kN]
    bsk: np.ndarray       # uint64[n][l][k+1][k+1][N], coefficient residues mod q
    ksk: np.ndarray       # int32[kN][t][n+1]


SIGMA_LOG2_LWE = -15.0
SIGMA_LOG2_GLWE = -25.0
# TORUS56 / N=2048 / 2^10 x 3: GLWE noise 2^-25 would exceed a 4-bit slot after
# 742 external products; the reference's TORUS56 default (TUniform bound 2^8 on
# a 56-bit torus, yatfhe_parameters.h:51) is ~2^-48.
SIGMA_LOG2_GLWE_T56 = -48.0


def default_sigmas(params: "Params"):
    return SIGMA_LOG2_LWE, (SIGMA_LOG2_GLWE if params.torus_bits == 32 else SIGMA_LOG2_GLWE_T56)


def keygen(params: Params, seed: int, sigma_log2_lwe: float | None = None,
           sigma_log2_glwe: float | None = None, threads: int = 0) -> KeySet:
    """Seeded keys; noise defaults per torus width (default_sigmas)."""
    lib = load_library()
    dl, dg = default_sigmas(params)
    sigma_log2_lwe = dl if sigma_log2_lwe is None else sigma_log2_lwe
    sigma_log2_glwe = dg if sigma_log2_glwe is None else sigma_log2_glwe
    lk = np.zeros(params.n, np.int32)
    gk = np.zeros(params.k * params.N, np.int32)
    bsk = np.zeros(params.bsk_shape, np.uint64)
    ksk = np.zeros(params.ksk_shape, np.int32)
    cp = params.c()
    _check(lib.ygpu_keygen(ctypes.byref(cp), seed, sigma_log2_lwe, sigma_log2_glwe,
                           _ptr(lk), _ptr(gk), _ptr(bsk), _ptr(ksk), threads))
    return KeySet(params, lk, gk, bsk, ksk)


@dataclass
class SuccinctKeySet:
    """WWL+24 succinct key (BootstrappingKeyWWL24): body-digit rows and the TRLEV of s^2."""
    params: Params
    lwe_key: np.ndarray   # int32[n]
    glwe_key: np.ndarray  # int32[kN]
    bsk_half: np.ndarray  # uint64[n][l][k+1][N], coefficient residues mod q (rows cPrime[lvl])
    s2: np.ndarray        # uint64[s2_level][k+1][N], coefficient residues mod q
    ksk: np.ndarray       # int32[kN][t][n+1]

    @property
    def s2_level(self) -> int:
        return int(self.s2.shape[0])


def max_s2_level(params: Params) -> int:
    """Most s^2 digits that fit the torus (decomposeRow's limit), at most 8."""
    return min(8, params.torus_bits // params.radix_bits)


def keygen_succinct(params: Params, seed: int, s2_level: int | None = None, sigma_log2_lwe: float | None = None,
                    sigma_log2_glwe: float | None = None, threads: int = 0) -> SuccinctKeySet:
    """Seeded WWL+24 succinct keys (ygpu_keygen_succinct); secret keys and KSK equal keygen's for the seed."""
    lib = load_library()
    dl, dg = default_sigmas(params)
    sigma_log2_lwe = dl if sigma_log2_lwe is None else sigma_log2_lwe
    sigma_log2_glwe = dg if sigma_log2_glwe is None else sigma_log2_glwe
    s2_level = max_s2_level(params) if s2_level is None else s2_level
    k1 = params.k + 1
    lk = np.zeros(params.n, np.int32)
    gk = np.zeros(params.k * params.N, np.int32)
    half = np.zeros((params.n, params.level, k1, params.N), np.uint64)
    s2 = np.zeros((s2_level, k1, params.N), np.uint64)
    ksk = np.zeros(params.ksk_shape, np.int32)
    cp = params.c()
    _check(lib.ygpu_keygen_succinct(ctypes.byref(cp), seed, sigma_log2_lwe, sigma_log2_glwe, s2_level,
                                    _ptr(lk), _ptr(gk), _ptr(half), _ptr(s2), _ptr(ksk), threads))
    return SuccinctKeySet(params, lk, gk, half, s2, ksk)


@dataclass
class LazyKeySet:
    """Lazy-pipe key (BootstrappingKeyMPLazyPipe): digits of the half RGSW rows,
    bskFirst (a TRLWE of v s_0, so the key belongs to the test polynomial v) and
    the TRLEV of s^2."""
    params: Params
    lwe_key: np.ndarray   # int32[n]
    glwe_key: np.ndarray  # int32[kN]
    first: np.ndarray     # int32[k+1][N]
    digits_a: np.ndarray  # int8[n-1][R][D][k][N]  (R = params.level, D = digit levels)
    digits_b: np.ndarray  # int8[n-1][R][D][N]
    s2: np.ndarray        # uint64[D][k+1][N], coefficient residues mod q
    ksk: np.ndarray       # int32[kN][t][n+1]
    v: np.ndarray         # the test polynomial bskFirst encrypts

    @property
    def digit_levels(self) -> int:
        return int(self.digits_b.shape[2])


def keygen_lazy(params: Params, seed: int, v: np.ndarray, digit_levels: int | None = None, sigma_log2_lwe: float | None = None,
                sigma_log2_glwe: float | None = None, threads: int = 0) -> LazyKeySet:
    """Seeded lazy-pipe keys (ygpu_keygen_lazy) for test polynomial v, rows of
    params.level levels stored as digit_levels digits (default params.level);
    secret keys and KSK equal keygen's for the seed."""
    lib = load_library()
    dl, dg = default_sigmas(params)
    sigma_log2_lwe = dl if sigma_log2_lwe is None else sigma_log2_lwe
    sigma_log2_glwe = dg if sigma_log2_glwe is None else sigma_log2_glwe
    R, K, N = params.level, params.k, params.N
    L = params.level if digit_levels is None else digit_levels
    vv = np.ascontiguousarray(v, dtype=np.int32)
    lk = np.zeros(params.n, np.int32)
    gk = np.zeros(K * N, np.int32)
    first = np.zeros((K + 1, N), np.int32)
    da = np.zeros((params.n - 1, R, L, K, N), np.int8)
    db = np.zeros((params.n - 1, R, L, N), np.int8)
    s2 = np.zeros((L, K + 1, N), np.uint64)
    ksk = np.zeros(params.ksk_shape, np.int32)
    cp = params.c()
    _check(lib.ygpu_keygen_lazy(ctypes.byref(cp), seed, sigma_log2_lwe, sigma_log2_glwe, L, _ptr(vv), _ptr(lk), _ptr(gk),
                                _ptr(first), _ptr(da), _ptr(db), _ptr(s2), _ptr(ksk), threads))
    return LazyKeySet(params, lk, gk, first, da, db, s2, ksk, vv.copy())


@dataclass
class LazyAltKeySet:
    """Alt lazy-pipe key (BootstrappingKeyMPLazyPipeAlt): torus half rows of every
    s_i and the TRLEV of s^2; independent of the test polynomial."""
    params: Params
    lwe_key: np.ndarray   # int32[n]
    glwe_key: np.ndarray  # int32[kN]
    rows: np.ndarray      # int32[n][R][k+1][N]
    s2: np.ndarray        # uint64[D][k+1][N]
    ksk: np.ndarray       # int32[kN][t][n+1]

    @property
    def digit_levels(self) -> int:
        return int(self.s2.shape[0])


def keygen_lazy_alt(params: Params, seed: int, digit_levels: int | None = None, sigma_log2_lwe: float | None = None,
                    sigma_log2_glwe: float | None = None, threads: int = 0) -> LazyAltKeySet:
    """Seeded Alt lazy-pipe keys (ygpu_keygen_lazy_alt)."""
    lib = load_library()
    dl, dg = default_sigmas(params)
    sigma_log2_lwe = dl if sigma_log2_lwe is None else sigma_log2_lwe
    sigma_log2_glwe = dg if sigma_log2_glwe is None else sigma_log2_glwe
    D = params.level if digit_levels is None else digit_levels
    lk = np.zeros(params.n, np.int32)
    gk = np.zeros(params.k * params.N, np.int32)
    rows = np.zeros((params.n, params.level, params.k + 1, params.N), np.int32)
    s2 = np.zeros((D, params.k + 1, params.N), np.uint64)
    ksk = np.zeros(params.ksk_shape, np.int32)
    cp = params.c()
    _check(lib.ygpu_keygen_lazy_alt(ctypes.byref(cp), seed, sigma_log2_lwe, sigma_log2_glwe, D, _ptr(lk), _ptr(gk), _ptr(rows),
                                    _ptr(s2), _ptr(ksk), threads))
    return LazyAltKeySet(params, lk, gk, rows, s2, ksk)


def keygen_classic(params: Params, seed: int, sigma_log2_lwe: float | None = None,
                   sigma_log2_glwe: float | None = None, threads: int = 0) -> KeySet:
    """Seeded classic bootstrapping key (ygpu_keygen_classic): same layout as keygen."""
    lib = load_library()
    dl, dg = default_sigmas(params)
    sigma_log2_lwe = dl if sigma_log2_lwe is None else sigma_log2_lwe
    sigma_log2_glwe = dg if sigma_log2_glwe is None else sigma_log2_glwe
    lk = np.zeros(params.n, np.int32)
    gk = np.zeros(params.k * params.N, np.int32)
    bsk = np.zeros(params.bsk_shape, np.uint64)
    ksk = np.zeros(params.ksk_shape, np.int32)
    cp = params.c()
    _check(lib.ygpu_keygen_classic(ctypes.byref(cp), seed, sigma_log2_lwe, sigma_log2_glwe,
                                   _ptr(lk), _ptr(gk), _ptr(bsk), _ptr(ksk), threads))
    return KeySet(params, lk, gk, bsk, ksk)


KEY_FORMS = {"mp": 0, "classic": 1, "succinct": 2}


def keygen_tuniform(params: Params, seed: int, lwe_bound_bits: int, glwe_bound_bits: int, form: str = "mp",
                    s2_level: int | None = None, threads: int = 0):
    """Seeded keys with the reference's TUniform noise (ygpu_keygen_tuniform):
    noise in [-2^b, 2^b] torus units, b = lweNoiseB for LWE/KSK rows and
    rlweNoiseB for the bootstrapping key's GLWE rows."""
    lib = load_library()
    k1 = params.k + 1
    lk = np.zeros(params.n, np.int32)
    gk = np.zeros(params.k * params.N, np.int32)
    ksk = np.zeros(params.ksk_shape, np.int32)
    cp = params.c()
    if form == "succinct":
        s2_level = max_s2_level(params) if s2_level is None else s2_level
        half = np.zeros((params.n, params.level, k1, params.N), np.uint64)
        s2 = np.zeros((s2_level, k1, params.N), np.uint64)
        _check(lib.ygpu_keygen_tuniform(ctypes.byref(cp), seed, lwe_bound_bits, glwe_bound_bits, KEY_FORMS[form], s2_level,
                                        _ptr(lk), _ptr(gk), _ptr(half), _ptr(s2), _ptr(ksk), threads))
        return SuccinctKeySet(params, lk, gk, half, s2, ksk)
    bsk = np.zeros(params.bsk_shape, np.uint64)
    _check(lib.ygpu_keygen_tuniform(ctypes.byref(cp), seed, lwe_bound_bits, glwe_bound_bits, KEY_FORMS[form], 0,
                                    _ptr(lk), _ptr(gk), _ptr(bsk), None, _ptr(ksk), threads))
    return KeySet(params, lk, gk, bsk, ksk)


def save_keys(path: str, keys):
    """Write the keys' coefficient form to a key interchange file (ygpu_key_file_write,
    or ygpu_key_file_write_succinct for a SuccinctKeySet)."""
    cp = keys.params.c()
    if isinstance(keys, SuccinctKeySet):
        half = np.ascontiguousarray(keys.bsk_half, dtype=np.uint64)
        s2 = np.ascontiguousarray(keys.s2, dtype=np.uint64)
        ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        _check(load_library().ygpu_key_file_write_succinct(os.fsencode(path), ctypes.byref(cp), _ptr(half), _ptr(s2),
                                                           keys.s2_level, _ptr(ksk)))
        return
    bsk = np.ascontiguousarray(keys.bsk, dtype=np.uint64)
    ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
    _check(load_library().ygpu_key_file_write(os.fsencode(path), ctypes.byref(cp), _ptr(bsk), _ptr(ksk)))


def key_file_params(path: str) -> Params:
    cp = _Params()
    _check(load_library().ygpu_key_file_info(os.fsencode(path), ctypes.byref(cp)))
    return Params(cp.torus_bits, cp.n, cp.k, cp.N, cp.radix_bits, cp.level, cp.ks_radix_bits, cp.ks_level,
                  cp.ks_width_bits, cp.lwe_q_bits, cp.q_ntt)


def load_key_arrays(path: str, expect: Params | None = None):
    """(params, bsk, ksk) from a key interchange file (ygpu_key_file_read)."""
    p = key_file_params(path)
    bsk = np.zeros(p.bsk_shape, np.uint64)
    ksk = np.zeros(p.ksk_shape, np.int32)
    ce = expect.c() if expect is not None else None
    _check(load_library().ygpu_key_file_read(os.fsencode(path), ctypes.byref(ce) if ce is not None else None,
                                             _ptr(bsk), _ptr(ksk)))
    return p, bsk, ksk


def key_file_s2_level(path: str) -> int:
    """0 for a full-key file, the s^2 TRLEV levels for a succinct-key file."""
    lv = ctypes.c_int32()
    _check(load_library().ygpu_key_file_form(os.fsencode(path), ctypes.byref(lv)))
    return lv.value


def load_succinct_arrays(path: str, expect: Params | None = None):
    """(params, bsk_half, s2, ksk) from a succinct-key file (ygpu_key_file_read_succinct)."""
    p = key_file_params(path)
    lv = key_file_s2_level(path)
    k1 = p.k + 1
    half = np.zeros((p.n, p.level, k1, p.N), np.uint64)
    s2 = np.zeros((max(lv, 1), k1, p.N), np.uint64)
    ksk = np.zeros(p.ksk_shape, np.int32)
    ce = expect.c() if expect is not None else None
    _check(load_library().ygpu_key_file_read_succinct(os.fsencode(path), ctypes.byref(ce) if ce is not None else None,
                                                      _ptr(half), _ptr(s2), _ptr(ksk)))
    return p, half, s2, ksk


def lwe_encrypt(lwe_key: np.ndarray, mu: np.ndarray, seed: int, sigma_log2: float = SIGMA_LOG2_LWE) -> np.ndarray:
    lib = load_library()
    mu = np.ascontiguousarray(mu, dtype=np.int32)
    n = lwe_key.shape[0]
    out = np.zeros((mu.shape[0], n + 1), np.int32)
    key = np.ascontiguousarray(lwe_key, dtype=np.int32)
    _check(lib.ygpu_lwe_encrypt(n, _ptr(key), seed, sigma_log2, mu.shape[0], _ptr(mu), _ptr(out)))
    return out


def lwe_encrypt_tuniform(lwe_key: np.ndarray, mu: np.ndarray, seed: int, bound_bits: int) -> np.ndarray:
    """symEncTlwe with TUniform noise in [-2^bound_bits, 2^bound_bits] (the reference's lweNoiseB)."""
    lib = load_library()
    mu = np.ascontiguousarray(mu, dtype=np.int32)
    n = lwe_key.shape[0]
    out = np.zeros((mu.shape[0], n + 1), np.int32)
    key = np.ascontiguousarray(lwe_key, dtype=np.int32)
    _check(lib.ygpu_lwe_encrypt_tuniform(n, _ptr(key), seed, bound_bits, mu.shape[0], _ptr(mu), _ptr(out)))
    return out


def lwe_decrypt_phase(ct: np.ndarray, lwe_key: np.ndarray) -> np.ndarray:
    """b - <a, s> mod 2^32 as int32 (before rounding), rows of ct."""
    ct = np.atleast_2d(ct).astype(np.int64)
    s = lwe_key.astype(np.int64)
    ph = (ct[:, -1] - (ct[:, :-1] * s).sum(axis=1)) & 0xFFFFFFFF
    return ph.astype(np.uint32).astype(np.int32)


# ------------------------------------------------------------- test polynomials
def sign_test_polynomial(params: Params, value: int | None = None) -> np.ndarray:
    """v_i = 1/8 of the torus for every i (the sign gate)."""
    val = value if value is not None else (1 << (params.torus_bits - 3))
    return np.full(params.N, val, dtype=params.torus_dtype)


def lut_test_polynomial(params: Params, lut, p: int) -> np.ndarray:
    """Programmable-bootstrap test polynomial for a p-entry LUT with one padding
    bit (messages m in [0, p) encoded at m * 2^32 / (2p)), laid out like
    generateTestPolynomialFR (polynomial.cpp:92-107): slot s covers 2N/(2p)
    consecutive coefficients, each slot centred on its sample point, the left
    half of slot 0 wrapped (negated) to the end.  Output values are
    lut[s] mod p scaled to the torus (modSwitchToTorusGeneral)."""
    N = params.N
    box = 2 * N // p
    half = box // 2
    halfhalf = half // 2
    w = params.torus_bits
    scale = 1 << (w - int(np.log2(p)))
    v = np.zeros(N, dtype=np.int64)
    for i in range(N):
        s = i // half
        m = int(lut[s]) % p
        m = m - p if m > (p - 1) // 2 else m  # intModP centring
        v[i] = m * scale
    v[:halfhalf] = -v[:halfhalf]
    v = np.roll(v, -halfhalf)
    mask = (1 << w) - 1
    v = ((v & mask) ^ (1 << (w - 1))) - (1 << (w - 1))
    return v.astype(params.torus_dtype)


# ----------------------------------------------------------------- device side
class DeviceBuffer:
    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        _check(ctx.lib.ygpu_malloc(ctx.h, nbytes, ctypes.byref(self.ptr)))

    @property
    def value(self):
        return self.ptr.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(self.ctx.lib.ygpu_memcpy_htod(self.ctx.h, self.ptr, _ptr(a), a.nbytes))

    def download(self, a: np.ndarray):
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        _check(self.ctx.lib.ygpu_memcpy_dtoh(self.ctx.h, _ptr(a), self.ptr, a.nbytes))
        return a

    def free(self):
        if self.ptr.value:
            _check(self.ctx.lib.ygpu_free(self.ctx.h, self.ptr))
            self.ptr = ctypes.c_void_p()


def device_count() -> int:
    lib = load_library()
    c = ctypes.c_int32(0)
    rc = lib.ygpu_device_count(ctypes.byref(c))
    return c.value if rc == YGPU_OK else 0


class Context:
    """One engine context on one GPU (replaces initYatfhe, initializer.cpp:50-85)."""

    def __init__(self, params: Params, device: int = 0):
        self.lib = load_library()
        self.params = params
        self.h = ctypes.c_void_p()
        cp = params.c()
        _check(self.lib.ygpu_init(ctypes.byref(cp), device, ctypes.byref(self.h)))

    def close(self):
        if self.h.value:
            self.lib.ygpu_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_key(self, keys) -> int:
        if isinstance(keys, SuccinctKeySet):
            return self.upload_key_succinct(keys)
        kid = ctypes.c_int32(-1)
        bsk = np.ascontiguousarray(keys.bsk, dtype=np.uint64)
        ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        _check(self.lib.ygpu_key_upload(self.h, _ptr(bsk), _ptr(ksk), ctypes.byref(kid)))
        return kid.value

    def upload_key_succinct(self, keys: SuccinctKeySet) -> int:
        kid = ctypes.c_int32(-1)
        half = np.ascontiguousarray(keys.bsk_half, dtype=np.uint64)
        s2 = np.ascontiguousarray(keys.s2, dtype=np.uint64)
        ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        _check(self.lib.ygpu_key_upload_succinct(self.h, _ptr(half), _ptr(s2), keys.s2_level, _ptr(ksk), ctypes.byref(kid)))
        return kid.value

    def upload_key_lazy(self, keys) -> int:
        """A lazy-pipe key set (LazyKeySet or LazyAltKeySet)."""
        kid = ctypes.c_int32(-1)
        if isinstance(keys, LazyAltKeySet):
            arrs = [np.ascontiguousarray(keys.rows, dtype=np.int32), np.ascontiguousarray(keys.s2, dtype=np.uint64),
                    np.ascontiguousarray(keys.ksk, dtype=np.int32)]
            _check(self.lib.ygpu_key_upload_lazy_alt(self.h, keys.digit_levels, *[_ptr(a) for a in arrs], ctypes.byref(kid)))
            return kid.value
        arrs = [np.ascontiguousarray(keys.first, dtype=np.int32), np.ascontiguousarray(keys.digits_a, dtype=np.int8),
                np.ascontiguousarray(keys.digits_b, dtype=np.int8), np.ascontiguousarray(keys.s2, dtype=np.uint64),
                np.ascontiguousarray(keys.ksk, dtype=np.int32)]
        _check(self.lib.ygpu_key_upload_lazy(self.h, keys.digit_levels, *[_ptr(a) for a in arrs], ctypes.byref(kid)))
        return kid.value

    def pbs_lazy(self, key_id: int, lwe_in: np.ndarray, v: np.ndarray, want_acc=False, want_extract=False):
        p = self.params
        lwe = np.ascontiguousarray(lwe_in, dtype=np.int32)
        vv = np.ascontiguousarray(v, dtype=p.torus_dtype)
        out = np.zeros(p.n + 1, np.int32)
        acc = np.zeros((p.k + 1, p.N), p.torus_dtype) if want_acc else None
        ext = np.zeros(p.k * p.N + 1, np.int32) if want_extract else None
        _check(self.lib.ygpu_pbs_lazy(self.h, key_id, _ptr(lwe), _ptr(vv), _ptr(out),
                                      _ptr(acc) if acc is not None else None, _ptr(ext) if ext is not None else None))
        return out, acc, ext

    def blind_rotate_lazy(self, key_id: int, abar: np.ndarray, v: np.ndarray):
        """blindRotateLazyPipeNtt on a rescaled input (a_0 .. a_{n-1}, b mod 2N)."""
        p = self.params
        ab = np.ascontiguousarray(abar, dtype=np.int32)
        vv = np.ascontiguousarray(v, dtype=p.torus_dtype)
        acc = np.zeros((p.k + 1, p.N), p.torus_dtype)
        _check(self.lib.ygpu_blind_rotate_lazy(self.h, key_id, _ptr(ab), _ptr(vv), _ptr(acc)))
        return acc

    def pbs_lazy_batch_device(self, key_id: int, count: int, d_lwe, d_v, v_per_ct: bool, d_out, d_acc=None, d_ext=None):
        _check(self.lib.ygpu_pbs_lazy_batch_device(self.h, key_id, count, d_lwe.value, d_v.value, 1 if v_per_ct else 0,
                                                   d_out.value, d_acc.value if d_acc else None,
                                                   d_ext.value if d_ext else None))

    def set_lazy_window(self, steps: int):
        _check(self.lib.ygpu_set_lazy_window(self.h, steps))

    def pbs_lazy_batch(self, key_id: int, lwe_in: np.ndarray, v: np.ndarray, want_acc=False, want_extract=False):
        """Host-buffer convenience over ygpu_pbs_lazy_batch_device."""
        p = self.params
        lwe = np.ascontiguousarray(lwe_in, dtype=np.int32)
        count = lwe.shape[0]
        vv = np.ascontiguousarray(v, dtype=p.torus_dtype)
        tb = np.dtype(p.torus_dtype).itemsize
        bufs = []
        try:
            d_lwe = self.alloc(lwe.nbytes); bufs.append(d_lwe)
            d_v = self.alloc(vv.nbytes); bufs.append(d_v)
            d_out = self.alloc(count * (p.n + 1) * 4); bufs.append(d_out)
            d_acc = self.alloc(count * (p.k + 1) * p.N * tb) if want_acc else None
            d_ext = self.alloc(count * (p.k * p.N + 1) * 4) if want_extract else None
            bufs += [b for b in (d_acc, d_ext) if b]
            d_lwe.upload(lwe)
            d_v.upload(vv)
            self.pbs_lazy_batch_device(key_id, count, d_lwe, d_v, vv.ndim == 2, d_out, d_acc, d_ext)
            self.sync()
            out = d_out.download(np.zeros((count, p.n + 1), np.int32))
            acc = d_acc.download(np.zeros((count, p.k + 1, p.N), p.torus_dtype)) if d_acc else None
            ext = d_ext.download(np.zeros((count, p.k * p.N + 1), np.int32)) if d_ext else None
            return out, acc, ext
        finally:
            for b in bufs:
                b.free()

    def upload_key_file(self, path: str) -> int:
        kid = ctypes.c_int32()
        _check(self.lib.ygpu_key_upload_file(self.h, os.fsencode(path), ctypes.byref(kid)))
        return kid.value

    def release_key(self, key_id: int):
        _check(self.lib.ygpu_key_release(self.h, key_id))

    def key_bytes(self, key_id: int) -> int:
        b = ctypes.c_int64(0)
        _check(self.lib.ygpu_key_bytes(self.h, key_id, ctypes.byref(b)))
        return b.value

    # -- plain drop-in form (host buffers)
    def pbs(self, key_id: int, lwe_in: np.ndarray, v: np.ndarray, want_acc=False, want_extract=False):
        p = self.params
        lwe = np.ascontiguousarray(lwe_in, dtype=np.int32)
        vv = np.ascontiguousarray(v, dtype=p.torus_dtype)
        out = np.zeros(p.n + 1, np.int32)
        acc = np.zeros((p.k + 1, p.N), p.torus_dtype) if want_acc else None
        ext = np.zeros(p.k * p.N + 1, np.int32) if want_extract else None
        _check(self.lib.ygpu_pbs(self.h, key_id, _ptr(lwe), _ptr(vv), _ptr(out),
                                 _ptr(acc) if acc is not None else None, _ptr(ext) if ext is not None else None))
        return out, acc, ext

    def blind_rotate(self, key_id: int, lwe_in: np.ndarray, v: np.ndarray):
        p = self.params
        lwe = np.ascontiguousarray(lwe_in, dtype=np.int32)
        vv = np.ascontiguousarray(v, dtype=p.torus_dtype)
        acc = np.zeros((p.k + 1, p.N), p.torus_dtype)
        _check(self.lib.ygpu_blind_rotate(self.h, key_id, _ptr(lwe), _ptr(vv), _ptr(acc)))
        return acc

    def blind_rotate_scaled(self, key_id: int, abar: np.ndarray, acc: np.ndarray):
        p = self.params
        ab = np.ascontiguousarray(abar, dtype=np.int32)
        a = np.ascontiguousarray(acc, dtype=p.torus_dtype).copy()
        _check(self.lib.ygpu_blind_rotate_scaled(self.h, key_id, _ptr(ab), _ptr(a)))
        return a

    def sample_extract(self, acc: np.ndarray, index: int | None = None):
        """extractTlweFromTrlwe at `index` (default: the context's extraction index)."""
        p = self.params
        out = np.zeros(p.k * p.N + 1, np.int32)
        a = np.ascontiguousarray(acc, dtype=p.torus_dtype)
        if index is None:
            _check(self.lib.ygpu_sample_extract(self.h, _ptr(a), _ptr(out)))
        else:
            _check(self.lib.ygpu_sample_extract_index(self.h, _ptr(a), index, _ptr(out)))
        return out

    def set_extract_index(self, index: int):
        """The bootstraps' extraction coefficient (the reference's param.driftPhase)."""
        _check(self.lib.ygpu_set_extract_index(self.h, index))

    def extract_index(self) -> int:
        i = ctypes.c_int32(0)
        _check(self.lib.ygpu_get_extract_index(self.h, ctypes.byref(i)))
        return i.value

    def keyswitch(self, key_id: int, lwe_in: np.ndarray):
        out = np.zeros(self.params.n + 1, np.int32)
        _check(self.lib.ygpu_keyswitch(self.h, key_id, _ptr(np.ascontiguousarray(lwe_in, dtype=np.int32)), _ptr(out)))
        return out

    def extract_keyswitch(self, key_id: int, acc: np.ndarray):
        p = self.params
        a = np.ascontiguousarray(acc, dtype=p.torus_dtype)
        ext = np.zeros(p.k * p.N + 1, np.int32)
        out = np.zeros(p.n + 1, np.int32)
        _check(self.lib.ygpu_extract_keyswitch(self.h, key_id, _ptr(a), _ptr(ext), _ptr(out)))
        return ext, out

    # -- batched form (device buffers)
    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def pbs_batch_device(self, key_id: int, count: int, d_lwe, d_v, v_per_ct: bool, d_out, d_acc=None, d_ext=None):
        _check(self.lib.ygpu_pbs_batch_device(self.h, key_id, count, d_lwe.value, d_v.value, 1 if v_per_ct else 0,
                                              d_out.value, d_acc.value if d_acc else None,
                                              d_ext.value if d_ext else None))

    def pbs_batch_device_noalloc(self, key_id: int, count: int, d_lwe, d_v, v_per_ct: bool, d_out, d_acc=None, d_ext=None):
        """ygpu_pbs_batch_device with its scratch in the library's static device array."""
        _check(self.lib.ygpu_pbs_batch_device_noalloc(self.h, key_id, count, d_lwe.value, d_v.value, 1 if v_per_ct else 0,
                                                      d_out.value, d_acc.value if d_acc else None,
                                                      d_ext.value if d_ext else None))

    def pbs_jobs_device(self, jobs, total: int, d_lwe, d_v, v_per_ct: bool, d_out, d_acc=None):
        arr = (_Job * max(1, len(jobs)))()
        for i, (kid, cnt, first) in enumerate(jobs):
            arr[i] = _Job(kid, cnt, first)
        _check(self.lib.ygpu_pbs_jobs_device(self.h, arr, len(jobs), total, d_lwe.value, d_v.value,
                                             1 if v_per_ct else 0, d_out.value, d_acc.value if d_acc else None))

    def pbs_plain_device(self, key_id: int, d_lwe, d_v, d_out, d_acc=None):
        _check(self.lib.ygpu_pbs_plain_device(self.h, key_id, d_lwe.value, d_v.value, d_out.value,
                                              d_acc.value if d_acc else None))

    def pbs_batch(self, key_id: int, lwe_in: np.ndarray, v: np.ndarray, want_acc=False, want_extract=False):
        """Host-buffer convenience over the batched device form."""
        p = self.params
        lwe = np.ascontiguousarray(lwe_in, dtype=np.int32)
        count = lwe.shape[0]
        vv = np.ascontiguousarray(v, dtype=p.torus_dtype)
        v_per_ct = vv.ndim == 2
        tb = np.dtype(p.torus_dtype).itemsize
        bufs = []
        try:
            d_lwe = self.alloc(lwe.nbytes); bufs.append(d_lwe)
            d_v = self.alloc(vv.nbytes); bufs.append(d_v)
            d_out = self.alloc(count * (p.n + 1) * 4); bufs.append(d_out)
            d_acc = self.alloc(count * (p.k + 1) * p.N * tb) if want_acc else None
            d_ext = self.alloc(count * (p.k * p.N + 1) * 4) if want_extract else None
            bufs += [b for b in (d_acc, d_ext) if b]
            d_lwe.upload(lwe)
            d_v.upload(vv)
            self.pbs_batch_device(key_id, count, d_lwe, d_v, v_per_ct, d_out, d_acc, d_ext)
            self.sync()
            out = d_out.download(np.zeros((count, p.n + 1), np.int32))
            acc = d_acc.download(np.zeros((count, p.k + 1, p.N), p.torus_dtype)) if d_acc else None
            ext = d_ext.download(np.zeros((count, p.k * p.N + 1), np.int32)) if d_ext else None
            return out, acc, ext
        finally:
            for b in bufs:
                b.free()

    def sync(self):
        _check(self.lib.ygpu_sync(self.h))

    def modmul_peak(self, kind: int = 0) -> float:
        """Measured modmul/s of the microbenchmark (0: binary64 exact, 1: 64-bit Shoup)."""
        r = ctypes.c_double(0)
        _check(self.lib.ygpu_modmul_peak(self.h, kind, ctypes.byref(r)))
        return r.value

    def tile(self):
        c, t = ctypes.c_int32(0), ctypes.c_int32(0)
        _check(self.lib.ygpu_get_tile(self.h, ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value

    def set_tile(self, cpw: int, twl: int):
        _check(self.lib.ygpu_set_tile(self.h, cpw, twl))

    def autotune(self, key_id: int, count: int):
        c, t, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_double(0)
        _check(self.lib.ygpu_autotune(self.h, key_id, count, ctypes.byref(c), ctypes.byref(t), ctypes.byref(ms)))
        return c.value, t.value, ms.value

    def tile_ex(self):
        c, t, b = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        _check(self.lib.ygpu_get_tile_ex(self.h, ctypes.byref(c), ctypes.byref(t), ctypes.byref(b)))
        return c.value, t.value, b.value

    def set_tile_ex(self, cpw: int, twl: int, tab: int):
        _check(self.lib.ygpu_set_tile_ex(self.h, cpw, twl, tab))

    def autotune_results(self):
        """[(cpw, twl, tab, best ms)] of the last autotune's candidates."""
        m = 32
        c, t, b = (ctypes.c_int32 * m)(), (ctypes.c_int32 * m)(), (ctypes.c_int32 * m)()
        ms, n = (ctypes.c_double * m)(), ctypes.c_int32(0)
        _check(self.lib.ygpu_autotune_results(self.h, m, c, t, b, ms, ctypes.byref(n)))
        return [(c[i], t[i], b[i], ms[i]) for i in range(min(m, n.value))]

    def graph_create(self, jobs, total: int, d_lwe, d_v, v_per_ct: bool, d_out, d_acc=None) -> int:
        arr = (_Job * max(1, len(jobs)))()
        for i, (kid, cnt, first) in enumerate(jobs):
            arr[i] = _Job(kid, cnt, first)
        gid = ctypes.c_int32(-1)
        _check(self.lib.ygpu_graph_create(self.h, arr, len(jobs), total, d_lwe.value, d_v.value, 1 if v_per_ct else 0,
                                          d_out.value, d_acc.value if d_acc else None, ctypes.byref(gid)))
        return gid.value

    def graph_launch(self, graph_id: int):
        _check(self.lib.ygpu_graph_launch(self.h, graph_id))

    def graph_destroy(self, graph_id: int):
        _check(self.lib.ygpu_graph_destroy(self.h, graph_id))

    def set_barrier_polls(self, polls: int):
        """Bound of the blind-rotation kernels' team-barrier waits (0: every
        team barrier fails at once, to exercise the YGPU_ERR_DEVICE report)."""
        _check(self.lib.ygpu_set_barrier_polls(self.h, polls))

    def copy_device(self, d_dst, d_src, nbytes: int):
        _check(self.lib.ygpu_copy_device(self.h, d_dst.value, d_src.value, nbytes))

    def timing(self, on: bool):
        _check(self.lib.ygpu_timing_enable(self.h, 1 if on else 0))

    def timing_get(self, kernel: str):
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0)
        _check(self.lib.ygpu_timing_get(self.h, kernel.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def timing_reset(self):
        _check(self.lib.ygpu_timing_reset(self.h))


class Request:
    """One bootstrap request: owns copies of its input, test polynomial and output."""

    def __init__(self, params: Params, tenant: int, lwe_in: np.ndarray, v: np.ndarray, arrival_ns: int = 0):
        self.lwe_in = np.ascontiguousarray(lwe_in, dtype=np.int32)
        self.v = np.ascontiguousarray(v, dtype=params.torus_dtype)
        assert self.lwe_in.shape == (params.n + 1,) and self.v.shape == (params.N,)
        self.out = np.zeros(params.n + 1, np.int32)
        self.c = _Request(tenant, 0, self.lwe_in.ctypes.data, self.v.ctypes.data, self.out.ctypes.data, arrival_ns)

    @property
    def status(self):
        return self.c.status

    @property
    def latency_ms(self):
        return (self.c.done_ns - self.c.arrival_ns) * 1e-6

    @property
    def batch_size(self):
        return self.c.batch_size


class Server:
    """Continuous-batching bootstrapping server (ygpu_server_*)."""

    def __init__(self, device: int = 0, key_budget_bytes: int = 0, lanes: int = 0, max_batch: int = 0,
                 queue_max: int = 0, max_tenants: int = 0, max_loads: int = 0, eviction: str = "freq_bytes",
                 sched: str = "shared", spill_threshold: int = 0, steal_threshold: int = 0):
        self.lib = load_library()
        self.h = ctypes.c_void_p()
        policy = {"freq_bytes": 0, "lru": 1}[eviction]
        cfg = _ServerConfig(device, lanes, max_batch, queue_max, max_tenants, max_loads, policy, 0, key_budget_bytes,
                            SCHED_MODES[sched], spill_threshold, steal_threshold, 0)
        _check(self.lib.ygpu_server_create(ctypes.byref(cfg), ctypes.byref(self.h)))
        self.params = []
        self.tenant_params = []
        # lanes 0: the server fits its streams to the hardware queues (lanes())
        self.config = {"lanes": lanes, "max_batch": max_batch or 4096, "queue_max": queue_max or (1 << 20),
                       "eviction": eviction, "sched": sched, "spill_threshold": spill_threshold or 2,
                       "steal_threshold": steal_threshold or 2, "key_budget_bytes": key_budget_bytes}

    def close(self):
        if self.h:
            self.lib.ygpu_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_params(self, params: Params) -> int:
        cp = params.c()
        ps = ctypes.c_int32()
        _check(self.lib.ygpu_server_add_params(self.h, ctypes.byref(cp), ctypes.byref(ps)))
        self.params.append(params)
        return ps.value

    def set_tile(self, pset: int, cpw: int, twl: int):
        _check(self.lib.ygpu_server_set_tile(self.h, pset, cpw, twl))

    def lanes(self, pset: int) -> int:
        n = ctypes.c_int32(0)
        _check(self.lib.ygpu_server_lanes(self.h, pset, ctypes.byref(n)))
        return n.value

    def set_extract_index(self, pset: int, index: int):
        _check(self.lib.ygpu_server_set_extract_index(self.h, pset, index))

    def add_tenant(self, pset: int, keys) -> int:
        if isinstance(keys, SuccinctKeySet):
            return self.add_tenant_succinct(pset, keys)
        t = ctypes.c_int32()
        bsk = np.ascontiguousarray(keys.bsk, dtype=np.uint64)
        ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        _check(self.lib.ygpu_server_add_tenant(self.h, pset, _ptr(bsk), _ptr(ksk), ctypes.byref(t)))
        self.tenant_params.append(pset)
        return t.value

    def add_tenant_succinct(self, pset: int, keys: SuccinctKeySet) -> int:
        t = ctypes.c_int32()
        half = np.ascontiguousarray(keys.bsk_half, dtype=np.uint64)
        s2 = np.ascontiguousarray(keys.s2, dtype=np.uint64)
        ksk = np.ascontiguousarray(keys.ksk, dtype=np.int32)
        _check(self.lib.ygpu_server_add_tenant_succinct(self.h, pset, _ptr(half), _ptr(s2), keys.s2_level, _ptr(ksk),
                                                        ctypes.byref(t)))
        self.tenant_params.append(pset)
        return t.value

    def add_tenant_shared(self, src: int) -> int:
        t = ctypes.c_int32()
        _check(self.lib.ygpu_server_add_tenant_shared(self.h, src, ctypes.byref(t)))
        self.tenant_params.append(self.tenant_params[src])
        return t.value

    def add_tenant_file(self, pset: int, path: str) -> int:
        """A tenant whose key stays in a key file until a load needs it (read on a miss)."""
        t = ctypes.c_int32()
        _check(self.lib.ygpu_server_add_tenant_file(self.h, pset, os.fsencode(path), ctypes.byref(t)))
        self.tenant_params.append(pset)
        return t.value

    def set_file_slots(self, pset: int, slots: int):
        _check(self.lib.ygpu_server_set_file_slots(self.h, pset, slots))

    def prefetch(self, tenant: int) -> bool:
        """Make a tenant's key resident now; False when no memory can be freed."""
        rc = self.lib.ygpu_server_prefetch(self.h, tenant)
        if rc == -7:
            return False
        _check(rc)
        return True

    def request(self, tenant: int, lwe_in: np.ndarray, v: np.ndarray) -> Request:
        return Request(self.params[self.tenant_params[tenant]], tenant, lwe_in, v)

    def submit(self, req: Request):
        _check(self.lib.ygpu_server_submit(self.h, ctypes.byref(req.c)))

    def wait(self, req: Request, timeout: float | None = None) -> np.ndarray:
        """Result of a request; with `timeout` (seconds), raises TimeoutError
        carrying the server's counters if it has not completed by then."""
        if timeout is None:
            _check(self.lib.ygpu_server_wait(self.h, ctypes.byref(req.c)))
            return req.out
        rc = self.lib.ygpu_server_wait_for(self.h, ctypes.byref(req.c), int(timeout * 1000))
        if rc == 1:
            raise TimeoutError(f"request of tenant {req.c.tenant} still pending after {timeout} s; server {self.stats()}")
        _check(rc)
        return req.out

    def stats(self) -> dict:
        st = _ServerStats()
        _check(self.lib.ygpu_server_stats_get(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f in _STAT_FIELDS}

    def loadgen(self, rate: float, requests: int, inputs, luts, expected=None, warm: int = 0, zipf_s: float = 1.1,
                seed: int = 5, slo_ms: float = 100.0) -> dict:
        """Open-loop Poisson/Zipf load over tenants 0..len(inputs)-1 (see ygpu_server_loadgen)."""
        n = len(inputs)
        keep = [np.ascontiguousarray(a, dtype=np.int32) for a in inputs]
        keepv = [np.ascontiguousarray(v) for v in luts]
        arr_in = (ctypes.c_void_p * n)(*[a.ctypes.data for a in keep])
        arr_v = (ctypes.c_void_p * n)(*[v.ctypes.data for v in keepv])
        arr_e = None
        if expected is not None:
            keepe = [np.ascontiguousarray(e, dtype=np.int32) for e in expected]
            arr_e = (ctypes.c_void_p * n)(*[e.ctypes.data for e in keepe])
        cfg = _LoadgenConfig(rate, requests, warm, zipf_s, seed, n, 0,
                             ctypes.cast(arr_in, ctypes.POINTER(ctypes.c_void_p)),
                             ctypes.cast(arr_v, ctypes.POINTER(ctypes.c_void_p)),
                             ctypes.cast(arr_e, ctypes.POINTER(ctypes.c_void_p)) if arr_e is not None else None,
                             slo_ms)
        res = _LoadgenResult()
        _check(self.lib.ygpu_server_loadgen(self.h, ctypes.byref(cfg), ctypes.byref(res)))
        out = {}
        for f, _ in _LoadgenResult._fields_:
            v = getattr(res, f)
            if f.startswith("pset_"):
                v = list(v)[:len(self.params)]
            elif f == "lane_requests":
                v = list(v)[:min(16, len(self.params) * self.config["lanes"])]
            out[f] = v
        out["load"] = {"rate": rate, "requests": requests, "warm": warm, "zipf_s": zipf_s, "seed": seed,
                       "slo_ms": slo_ms, "users": n}
        return out

    def report(self, res: dict) -> dict:
        """One loadgen result as the reference serving benchmark's JSON record
        (ya_benchmark/blindrotate_serving.cpp:289-353): the same keys, the
        server's lanes standing for its workers.  Keys with no counterpart on
        a GPU (CPU affinity, device read concurrency) are None."""
        ld, cfg = res["load"], self.config

        def summary(prefix):
            if prefix == "sojourn":
                keys = {"mean_ms": "mean_ms", "sd_ms": "sd_ms", "p50_ms": "p50_ms", "p90_ms": "p90_ms",
                        "p99_ms": "p99_ms", "p99_9_ms": "p99_9_ms", "max_ms": "max_ms"}
            elif prefix == "queue_delay":
                keys = {"mean_ms": "queue_mean_ms", "p50_ms": "queue_p50_ms", "p90_ms": "queue_p90_ms",
                        "p99_ms": "queue_p99_ms", "max_ms": "queue_max_ms"}
            else:
                keys = {"mean_ms": "service_mean_ms", "p50_ms": "service_p50_ms", "p90_ms": "service_p90_ms",
                        "p99_ms": "service_p99_ms", "max_ms": "service_max_ms"}
            return {"count": res["completed"], **{k: res[v] for k, v in keys.items()}}

        return {
            "method": "GPU-MP21", "schema_version": 1,
            "config": {"arrival": "poisson", "offered_rate_rps": ld["rate"], "workers": cfg["lanes"],
                       "requests": ld["requests"], "warm": ld["warm"], "users": ld["users"],
                       "cache_capacity_keys": None, "key_budget_bytes": cfg["key_budget_bytes"],
                       "zipf_s": ld["zipf_s"], "seed": ld["seed"], "slo_ms": ld["slo_ms"],
                       "queue_max": cfg["queue_max"], "batch": cfg["max_batch"], "sched": cfg["sched"],
                       "spill_threshold": cfg["spill_threshold"], "steal_threshold": cfg["steal_threshold"]},
            "environment": {"affinity_cpus": None, "cpus_available": os.cpu_count()},
            "throughput_rps": res["throughput_rps"], "offered_rps": res["offered_rps"], "wall_sec": res["wall_sec"],
            "shed": res["shed"], "shed_rate": res["shed"] / ld["requests"] if ld["requests"] else 0.0,
            "hit_rate": res["hit_rate"], "coalesced_loads": res["coalesced_loads"], "slo_ms": ld["slo_ms"],
            "slo_violation_rate": res["slo_violation_rate"],
            "slo_violation_rate_100": res["slo_violation_rate_100"],
            "slo_violation_rate_300": res["slo_violation_rate_300"],
            "slo_violation_rate_500": res["slo_violation_rate_500"],
            "shed_plus_completed": res["shed"] + res["completed"], "peak_queue_depth": res["peak_queue_depth"],
            "load_concurrency_peak": res["load_concurrency_peak"], "load_concurrency_mean": res["load_concurrency_mean"],
            "device": "hbm",
            "device_inflight_reads_peak": None, "device_inflight_reads_mean": None, "device_inflight_samples": 0,
            "batch": cfg["max_batch"], "sched": cfg["sched"], "per_worker_requests": res["lane_requests"],
            "tenant_fairness_jain": res["tenant_fairness_jain"], "locality": res["locality"],
            "spilled": res["spilled"], "stolen": res["stolen"], "spill_threshold": cfg["spill_threshold"],
            "sojourn": summary("sojourn"), "queue_delay": summary("queue_delay"), "service": summary("service"),
            "mean_batch": res["mean_batch"], "key_loads": res["key_loads"], "bytes_loaded": res["bytes_loaded"],
            "mismatches": res["mismatches"],
        }


__all__ = [
    "Params", "PARAMS_T32_N1024", "PARAMS_T56_N2048", "Q_49P_T32", "Q_61P_T56", "KeySet", "keygen",
    "lwe_encrypt", "lwe_decrypt_phase", "sign_test_polynomial", "lut_test_polynomial", "Context",
    "DeviceBuffer", "device_count", "load_library", "exported_symbols", "YgpuError", "replace", "Server", "Request", "LazyKeySet", "keygen_lazy", "LazyAltKeySet", "keygen_lazy_alt",
    "save_keys", "key_file_params", "load_key_arrays", "SuccinctKeySet", "keygen_succinct", "max_s2_level",
    "key_file_s2_level", "load_succinct_arrays", "keygen_classic", "SCHED_MODES", "abi_struct_mismatches",
    "keygen_tuniform", "lwe_encrypt_tuniform", "KEY_FORMS",
]
