"""PETSc binary vectors (.bin) through the C ABI (CPU only), checked against
files assembled here from PETSc's published layout: big-endian class id
1211214 and length (8-byte words for a 64-bit-index PETSc such as the
reference's, 4-byte words otherwise), then big-endian float32 values."""
import numpy as np
import pytest

import claire_amd

CLASSID = 1211214


def _file(values, word=8):
    it = ">i8" if word == 8 else ">i4"
    return np.array([CLASSID, values.size], dtype=it).tobytes() + values.astype(">f4").tobytes()


def test_write_matches_published_layout(tmp_path):
    a = np.random.default_rng(0).standard_normal((6, 5, 4)).astype(np.float32)
    p = tmp_path / "m.bin"
    claire_amd.bin_write(p, a)
    assert p.read_bytes() == _file(a.reshape(-1))


@pytest.mark.parametrize("word", [8, 4])
def test_read_both_header_widths(tmp_path, word):
    a = np.random.default_rng(1).standard_normal(1000).astype(np.float32)
    a[:4] = [0.0, -0.0, np.inf, 1e-40]  # signed zero, infinity and a subnormal survive bit for bit
    p = tmp_path / "v.bin"
    p.write_bytes(_file(a, word))
    got = claire_amd.bin_read(p)
    assert got.dtype == np.float32 and np.array_equal(got.view(np.uint32), a.view(np.uint32))


def test_round_trip_large_and_shaped(tmp_path):
    shape = (64, 80, 72)  # spans several internal 1 MiB chunks
    a = np.random.default_rng(2).standard_normal(shape).astype(np.float32)
    p = tmp_path / "big.bin"
    claire_amd.bin_write(p, a)
    assert np.array_equal(claire_amd.bin_read(p, shape), a)


def test_empty_vector(tmp_path):
    p = tmp_path / "e.bin"
    claire_amd.bin_write(p, np.zeros(0, np.float32))
    assert claire_amd.bin_read(p).size == 0


def test_errors(tmp_path):
    bad = tmp_path / "bad.bin"
    bad.write_bytes(np.array([1234, 4], dtype=">i8").tobytes() + b"\0" * 16)
    with pytest.raises(claire_amd.ClaireError, match="class id"):
        claire_amd.bin_read(bad)
    short = tmp_path / "short.bin"
    short.write_bytes(_file(np.ones(10, np.float32))[:-8])
    with pytest.raises(claire_amd.ClaireError, match="shorter"):
        claire_amd.bin_read(short)
    with pytest.raises(claire_amd.ClaireError):
        claire_amd.bin_read(tmp_path / "missing.bin")
