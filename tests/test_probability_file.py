"""CPU checks of the importance-probability files (the reference's UAP v2
format, src/io/importance_probability_io.cpp) and of the SHA-256 digest that
binds them to a representative matrix file (src/io/sha256.cpp).

Pins: the reference's own tests (tests/io/importance_probability_io_test.cpp),
the byte layout written out field by field with struct, FIPS 180-4 SHA-256
known answers and Python's hashlib, and -- where oracle/_ref is built -- the
reference's own reader, writer and hash on the same files."""

import hashlib
import os
import struct

import numpy as np
import pytest

from conftest import rng as np_rng

import ultrahigh_ann_amd as uh

DIGEST = bytes(range(32))


def test_round_trip_preserves_metadata_and_values(tmp_path):
    # importance_probability_io_test.cpp:67-118
    path = tmp_path / "nested" / "probabilities.uap"
    expected = np.array([0.0, 0.125, 0.5, 1.0])
    uh.save_importance_probabilities(path, uh.METRIC_L2, 7028, DIGEST, expected)
    f = uh.load_importance_probabilities(path)
    assert f.metric == uh.METRIC_L2
    assert f.representative_count == 7028
    assert f.sha256 == DIGEST
    np.testing.assert_array_equal(f.probabilities, expected)
    assert os.path.getsize(path) == 64 + expected.size * 8
    assert uh.probability_source_matches(f, uh.METRIC_L2, 7028, 4, DIGEST)
    wrong = bytes([DIGEST[0] ^ 0xFF]) + DIGEST[1:]
    assert not uh.probability_source_matches(f, uh.METRIC_L2, 7028, 4, wrong)
    assert not uh.probability_source_matches(f, uh.METRIC_L1, 7028, 4, DIGEST)
    assert not uh.probability_source_matches(f, uh.METRIC_L2, 33, 4, DIGEST)
    assert not uh.probability_source_matches(f, uh.METRIC_L2, 7028, 5, DIGEST)


def test_byte_layout(tmp_path):
    """magic, kind (1 = L1, 2 = L2), 7 zero bytes, u64 counts, digest, then
    little-endian binary64 values and nothing else."""
    p = np.array([0.25, 1.0, 0.0, 2.0**-30])
    for metric, kind in ((uh.METRIC_L1, 1), (uh.METRIC_L2, 2)):
        path = tmp_path / f"k{kind}.uap"
        uh.save_importance_probabilities(path, metric, 33, DIGEST, p)
        want = b"UHANNPR2" + bytes([kind]) + bytes(7) + struct.pack("<QQ", 33, p.size) + DIGEST
        want += struct.pack(f"<{p.size}d", *p)
        assert path.read_bytes() == want


@pytest.mark.parametrize("payload", [b"abc", b"", b"a" * 55, b"a" * 56, b"a" * 64, b"a" * 1_000_001,
                                     b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq"])
def test_sha256_known_answers(tmp_path, payload):
    # importance_probability_io_test.cpp:120-133 ("abc") and FIPS 180-4 examples
    path = tmp_path / "x.bin"
    path.write_bytes(payload)
    assert uh.sha256_file(path) == hashlib.sha256(payload).digest()
    if payload == b"abc":
        assert uh.sha256_file(path).hex() == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"


def test_rejects_invalid_probability_values(tmp_path):
    # importance_probability_io_test.cpp:135-157
    path = tmp_path / "invalid.uap"
    for bad in ([1.1], [np.nan], [-0.5], [np.inf]):
        with pytest.raises(uh.UhannError, match=r"finite and in \[0, 1\]"):
            uh.save_importance_probabilities(path, uh.METRIC_L1, 2, bytes(32), np.array(bad))
    assert not path.exists()  # checked before the file is opened


def test_rejects_malformed_files(tmp_path):
    # importance_probability_io_test.cpp:159-190, plus the loader's other checks
    wrong = tmp_path / "wrong-magic.uap"
    wrong.write_bytes(b"not a probability file")
    with pytest.raises(uh.UhannError, match="wrong magic"):
        uh.load_importance_probabilities(wrong)
    good = tmp_path / "good.uap"
    uh.save_importance_probabilities(good, uh.METRIC_L2, 3, bytes(32), np.array([0.25, 0.75]))
    raw = good.read_bytes()
    cases = {
        "truncated": raw[:-1],
        "unknown probability kind": raw[:8] + b"\x03" + raw[9:],
        "reserved header bytes are not zero": raw[:12] + b"\x01" + raw[13:],
        "unexpected trailing data": raw + b"\x00",
        "finite and in": raw[:64] + struct.pack("<d", 1.5) + raw[72:],
    }
    for message, data in cases.items():
        path = tmp_path / "case.uap"
        path.write_bytes(data)
        with pytest.raises(uh.UhannError, match=message):
            uh.load_importance_probabilities(path)
    with pytest.raises(uh.UhannError, match="cannot be opened"):
        uh.load_importance_probabilities(tmp_path / "missing.uap")


def test_files_interchange_with_reference(tmp_path, ref_mod):
    """Same bytes as the reference's writer; each side reads the other's file;
    same digest as the reference's hash on the representative matrix bytes."""
    g = np_rng(5)
    refs = (g.random((30, 500), dtype=np.float32) * 2).astype(np.float32)
    matrix = tmp_path / "reps.bin"
    matrix.write_bytes(refs.tobytes())
    digest = uh.sha256_file(matrix)
    theirs_digest = bytearray(32)
    buf = (np.ctypeslib.as_ctypes_type(np.uint8) * 32).from_buffer(theirs_digest)
    assert ref_mod.ref().ref_sha256_file(os.fsencode(matrix), buf) == 0
    assert digest == bytes(theirs_digest)
    for metric_l1 in (False, True):
        probs = ref_mod.importance_probabilities(metric_l1, refs)
        metric = uh.METRIC_L1 if metric_l1 else uh.METRIC_L2
        mine, theirs = tmp_path / f"mine{metric}.uap", tmp_path / f"theirs{metric}.uap"
        uh.save_importance_probabilities(mine, metric, 30, digest, probs)
        assert ref_mod.ref().ref_save_probabilities(os.fsencode(theirs), int(metric_l1), 30, bytes(digest),
                                                    ref_mod._p(probs, ref_mod._f64p), probs.size) == 0
        assert mine.read_bytes() == theirs.read_bytes()
        f = uh.load_importance_probabilities(theirs)
        assert (f.metric, f.representative_count, f.sha256) == (metric, 30, digest)
        np.testing.assert_array_equal(f.probabilities, probs)
        out = np.zeros(probs.size, np.float64)
        kind, rows, d = np.zeros(1, np.int32), np.zeros(1, np.uint64), np.zeros(1, np.uint64)
        dig = bytearray(32)
        dbuf = (np.ctypeslib.as_ctypes_type(np.uint8) * 32).from_buffer(dig)
        assert ref_mod.ref().ref_load_probabilities(
            os.fsencode(mine), kind.ctypes.data_as(ref_mod.POINTER(ref_mod.c_int)), ref_mod._p(rows, ref_mod._u64p),
            dbuf, ref_mod._p(out, ref_mod._f64p), out.size, ref_mod._p(d, ref_mod._u64p)) == 0
        assert (int(kind[0]), int(rows[0]), int(d[0]), bytes(dig)) == (int(metric_l1), 30, probs.size, digest)
        np.testing.assert_array_equal(out, probs)
