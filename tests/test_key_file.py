"""Key interchange file: round trip, header-only reads, and every corruption the
format promises to catch (bad magic, truncation, flipped payload bits, wrong
parameter set, residues out of range), without a GPU."""
import os

import numpy as np
import pytest

import yatfhe_amd as yg

HEADER = 128


def write(tmp_path, keys, name="k.ygkey"):
    path = str(tmp_path / name)
    yg.save_keys(path, keys)
    return path


def test_round_trip_t32(tmp_path, p32, keys32):
    path = write(tmp_path, keys32)
    assert yg.key_file_params(path) == p32
    p, bsk, ksk = yg.load_key_arrays(path, expect=p32)
    assert p == p32
    assert np.array_equal(bsk, keys32.bsk) and np.array_equal(ksk, keys32.ksk)
    expect = HEADER + keys32.bsk.size * 8 + keys32.ksk.size * 4
    assert os.path.getsize(path) == expect
    assert not os.path.exists(path + ".tmp")


def test_header_layout(tmp_path, p32, keys32):
    path = write(tmp_path, keys32)
    with open(path, "rb") as f:
        h = f.read(HEADER)
    assert h[:8] == b"YGPUKEY1"
    assert int.from_bytes(h[8:12], "little") == 1 and int.from_bytes(h[12:16], "little") == HEADER
    fields = np.frombuffer(h[16:56], dtype="<i4")
    assert list(fields) == [p32.torus_bits, p32.n, p32.k, p32.N, p32.radix_bits, p32.level, p32.ks_radix_bits,
                            p32.ks_level, p32.ks_width_bits, p32.lwe_q_bits]
    assert int.from_bytes(h[56:64], "little") == p32.q_ntt
    assert int.from_bytes(h[64:72], "little") == keys32.bsk.size
    assert int.from_bytes(h[72:80], "little") == keys32.ksk.size
    assert h[104:128] == bytes(24)


def corrupt(path, offset, xor=0x01):
    with open(path, "r+b") as f:
        f.seek(offset)
        b = f.read(1)
        f.seek(offset)
        f.write(bytes([b[0] ^ xor]))


@pytest.mark.parametrize("where", ["magic", "params", "bsk", "ksk_last"])
def test_corruption_detected(tmp_path, p32, keys32, where):
    path = write(tmp_path, keys32)
    size = os.path.getsize(path)
    offset = {"magic": 0, "params": 20, "bsk": HEADER + 12345 * 8 + 3, "ksk_last": size - 1}[where]
    corrupt(path, offset)
    with pytest.raises(yg.YgpuError) as e:
        yg.load_key_arrays(path)
    assert e.value.code == -9  # YGPU_ERR_CORRUPT


def test_truncated_and_trailing(tmp_path, keys32):
    path = write(tmp_path, keys32)
    data = open(path, "rb").read()
    short = str(tmp_path / "short.ygkey")
    open(short, "wb").write(data[:-4])
    with pytest.raises(yg.YgpuError) as e:
        yg.load_key_arrays(short)
    assert e.value.code == -9
    longer = str(tmp_path / "long.ygkey")
    open(longer, "wb").write(data + b"\0")
    with pytest.raises(yg.YgpuError) as e:
        yg.load_key_arrays(longer)
    assert e.value.code == -9
    tiny = str(tmp_path / "tiny.ygkey")
    open(tiny, "wb").write(data[:40])
    with pytest.raises(yg.YgpuError) as e:
        yg.key_file_params(tiny)
    assert e.value.code == -9


def test_parameter_mismatch_and_missing(tmp_path, p32, keys32):
    path = write(tmp_path, keys32)
    other = yg.replace(p32, ks_level=7)
    with pytest.raises(yg.YgpuError) as e:
        yg.load_key_arrays(path, expect=other)
    assert e.value.code == -1  # YGPU_ERR_INVALID_ARG
    with pytest.raises(yg.YgpuError) as e:
        yg.key_file_params(str(tmp_path / "absent.ygkey"))
    assert e.value.code == -8  # YGPU_ERR_IO


def test_residue_out_of_range_rejected_at_write(tmp_path, p32, keys32):
    bad = yg.KeySet(p32, keys32.lwe_key, keys32.glwe_key, keys32.bsk.copy(), keys32.ksk)
    bad.bsk.reshape(-1)[777] = p32.q_ntt
    with pytest.raises(yg.YgpuError) as e:
        yg.save_keys(str(tmp_path / "bad.ygkey"), bad)
    assert e.value.code == -1
    assert not os.path.exists(str(tmp_path / "bad.ygkey"))


def test_digest_is_order_sensitive(tmp_path, p32, keys32):
    """Swapping two BSK words keeps every word in range but must fail the digest."""
    path = write(tmp_path, keys32)
    with open(path, "r+b") as f:
        f.seek(HEADER + 8 * 100)
        a = f.read(8)
        f.seek(HEADER + 8 * 5000)
        b = f.read(8)
        assert a != b
        f.seek(HEADER + 8 * 100)
        f.write(b)
        f.seek(HEADER + 8 * 5000)
        f.write(a)
    with pytest.raises(yg.YgpuError) as e:
        yg.load_key_arrays(path)
    assert e.value.code == -9


@pytest.fixture(scope="module")
def sk_small(p32):
    return yg.keygen_succinct(yg.replace(p32, n=24), seed=31, s2_level=4)


def test_succinct_round_trip_and_layout(tmp_path, sk_small):
    p = sk_small.params
    path = write(tmp_path, sk_small, "s.ygkey")
    assert yg.key_file_params(path) == p
    assert yg.key_file_s2_level(path) == 4
    q, half, s2, ksk = yg.load_succinct_arrays(path, expect=p)
    assert q == p
    assert np.array_equal(half, sk_small.bsk_half) and np.array_equal(s2, sk_small.s2) and np.array_equal(ksk, sk_small.ksk)
    bsk_words = sk_small.bsk_half.size + sk_small.s2.size
    assert os.path.getsize(path) == HEADER + bsk_words * 8 + sk_small.ksk.size * 4
    with open(path, "rb") as f:
        h = f.read(HEADER)
    assert int.from_bytes(h[8:12], "little") == 2
    assert int.from_bytes(h[64:72], "little") == bsk_words
    assert int.from_bytes(h[104:108], "little") == 4 and h[108:112] == bytes(4)
    assert int.from_bytes(h[112:120], "little") != 0 and h[120:128] == bytes(8)
    # s2 starts right after the body-digit rows
    with open(path, "rb") as f:
        f.seek(HEADER + sk_small.bsk_half.size * 8)
        assert np.array_equal(np.frombuffer(f.read(8 * p.N), "<u8"), sk_small.s2[0, 0])


def test_succinct_form_checks(tmp_path, p32, keys32, sk_small):
    spath = write(tmp_path, sk_small, "s.ygkey")
    fpath = write(tmp_path, keys32, "f.ygkey")
    assert yg.key_file_s2_level(fpath) == 0
    with pytest.raises(yg.YgpuError):
        yg.load_key_arrays(spath)          # a succinct file is not a full key
    with pytest.raises(yg.YgpuError):
        yg.load_succinct_arrays(fpath)     # nor the other way round
    for offset in (104, 108, 112, HEADER + 5 * 8 + 1, HEADER + sk_small.bsk_half.size * 8 + 2):
        bad = str(tmp_path / f"bad{offset}.ygkey")
        with open(spath, "rb") as f, open(bad, "wb") as g:
            g.write(f.read())
        corrupt(bad, offset)
        with pytest.raises(yg.YgpuError) as e:
            yg.load_succinct_arrays(bad)
        assert e.value.code == -9, offset  # YGPU_ERR_CORRUPT
    with pytest.raises(yg.YgpuError):
        yg.save_keys(str(tmp_path / "x.ygkey"), yg.replace(sk_small, s2=sk_small.s2[:0]))  # s2_level 0
