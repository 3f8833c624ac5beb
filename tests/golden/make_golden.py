"""Writes the golden fixtures of the bootstrapping path (run from the repo root):

    python tests/golden/make_golden.py

Keys come from the engine's seeded host key generator (seed 0); expected outputs
from the CPU oracle (oracle/, the restatement of the reference's path).  Stored:
  golden.json     parameters, seeds, SHA-256 of the key material and of outputs
  golden.npz      config-0 chain (input LWE, accumulator, extracted and keyswitched
                  LWE), 64 TORUS32 batch outputs, 16 TORUS56 LUT outputs
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import yatfhe_amd as yg  # noqa: E402
from tests.oracle_py import Oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def t32_inputs(keys, count, seed):
    bits = np.random.default_rng(seed).integers(0, 2, count)
    return bits, yg.lwe_encrypt(keys.lwe_key, ((2 * bits - 1) * (1 << 29)).astype(np.int32), seed=seed)


def t56_inputs(keys, count, seed, p=16):
    rng = np.random.default_rng(seed)
    lut = rng.integers(0, p, p)
    msgs = rng.integers(0, p, count)
    cts = yg.lwe_encrypt(keys.lwe_key, (msgs * (1 << 32) // (2 * p)).astype(np.int64).astype(np.int32), seed=seed + 1)
    return lut, msgs, cts


def main():
    p32, p56 = yg.PARAMS_T32_N1024, yg.PARAMS_T56_N2048
    k32 = yg.keygen(p32, seed=0)
    k56 = yg.keygen(p56, seed=0)
    v32 = yg.sign_test_polynomial(p32)
    o32 = Oracle(p32, k32)
    bits0, ct0 = t32_inputs(k32, 1, seed=0)
    acc0, ext0, out0 = o32.pbs(ct0[0], v32)
    bits1, ct1 = t32_inputs(k32, 64, seed=1)
    out1 = o32.pbs_batch(ct1, v32)
    o32.close()
    lut, msgs, ct2 = t56_inputs(k56, 16, seed=2)
    v56 = yg.lut_test_polynomial(p56, lut, 16)
    o56 = Oracle(p56, k56)
    out2 = o56.pbs_batch(ct2, v56)
    o56.close()
    np.savez_compressed(os.path.join(OUT, "golden.npz"),
                        t32_ct0=ct0[0], t32_acc0=acc0, t32_ext0=ext0, t32_out0=out0, t32_bits0=bits0,
                        t32_out1=out1, t32_bits1=bits1, t56_lut=lut, t56_msgs=msgs, t56_out2=out2)
    meta = {
        "params": {"t32": p32.__dict__, "t56": p56.__dict__},
        "seeds": {"keys": 0, "t32_chain": 0, "t32_batch": 1, "t56_batch": 2},
        "sha256": {
            "t32_keys": sha(k32.lwe_key, k32.glwe_key, k32.bsk, k32.ksk),
            "t56_keys": sha(k56.lwe_key, k56.glwe_key, k56.bsk, k56.ksk),
            "t32_batch_inputs": sha(ct1),
            "t32_batch_outputs": sha(out1),
            "t56_batch_inputs": sha(ct2),
            "t56_batch_outputs": sha(out2),
        },
    }
    with open(os.path.join(OUT, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(OUT, "golden.npz"), os.path.getsize(os.path.join(OUT, "golden.npz")), "bytes")


if __name__ == "__main__":
    main()
