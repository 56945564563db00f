"""NIfTI-1 reading and writing through the C ABI (CPU only), checked against
an independent pure-Python encoder/decoder of the NIfTI-1 header layout
(struct + gzip), both byte orders and the integer and float datatypes, and
against the reference's own data files when they are present."""
import gzip
import os
import struct

import numpy as np
import pytest

import claire_amd

DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8, 512: np.uint16}


def _encode(a, code, order="<", pixdim=(1.0, 1.0, 1.0), vox_offset=352):
    """NIfTI-1 single file: header fields at their published byte offsets."""
    nz, ny, nx = a.shape
    h = bytearray(348)
    struct.pack_into(order + "i", h, 0, 348)
    struct.pack_into(order + "8h", h, 40, 3, nx, ny, nz, 1, 1, 1, 1)
    struct.pack_into(order + "hh", h, 70, code, 8 * np.dtype(DTYPES[code]).itemsize)
    struct.pack_into(order + "8f", h, 76, 1.0, pixdim[2], pixdim[1], pixdim[0], 0, 0, 0, 0)
    struct.pack_into(order + "f", h, 108, float(vox_offset))
    h[344:348] = b"n+1\0"
    data = np.ascontiguousarray(a, dtype=np.dtype(DTYPES[code]).newbyteorder(order)).tobytes()
    return bytes(h) + b"\0" * (vox_offset - 348) + data


def _decode(raw):
    order = "<" if struct.unpack_from("<i", raw, 0)[0] == 348 else ">"
    dim = struct.unpack_from(order + "8h", raw, 40)
    code = struct.unpack_from(order + "h", raw, 70)[0]
    off = int(struct.unpack_from(order + "f", raw, 108)[0])
    n = dim[1] * dim[2] * dim[3]
    dt = np.dtype(DTYPES[code]).newbyteorder(order)
    return np.frombuffer(raw, dtype=dt, count=n, offset=off).reshape(dim[3], dim[2], dim[1]).astype(np.float32)


@pytest.mark.parametrize("code", sorted(DTYPES))
@pytest.mark.parametrize("order", ["<", ">"])
@pytest.mark.parametrize("gz", [False, True])
def test_read_datatypes_and_byte_orders(tmp_path, code, order, gz):
    rng = np.random.default_rng(code)
    shape = (5, 7, 9)
    info = np.iinfo(DTYPES[code]) if np.issubdtype(DTYPES[code], np.integer) else None
    a = (rng.integers(info.min, info.max, size=shape, endpoint=True) if info else rng.standard_normal(shape))
    raw = _encode(a.astype(DTYPES[code]), code, order)
    path = tmp_path / ("img.nii.gz" if gz else "img.nii")
    path.write_bytes(gzip.compress(raw) if gz else raw)
    out, hdr = claire_amd.nifti_read(path)
    assert out.shape == shape
    np.testing.assert_array_equal(out, _decode(raw))
    assert hdr == raw[:348]


@pytest.mark.parametrize("name", ["img.nii", "img.nii.gz"])
def test_write_read_round_trip(tmp_path, name):
    rng = np.random.default_rng(7)
    a = rng.standard_normal((6, 4, 10)).astype(np.float32)
    path = tmp_path / name
    claire_amd.nifti_write(path, a, spacing=(0.5, 2.0, 1.5))
    raw = path.read_bytes()
    raw = gzip.decompress(raw) if name.endswith(".gz") else raw
    np.testing.assert_array_equal(_decode(raw), a)
    assert struct.unpack_from("<3f", raw, 80) == (1.5, 2.0, 0.5)  # pixdim of (x, y, z)
    b, hdr = claire_amd.nifti_read(path)
    np.testing.assert_array_equal(a, b)
    # header reuse (nifti_copy_nim_info): pixdim kept, datatype FLOAT32
    claire_amd.nifti_write(tmp_path / "copy.nii", b * 2, header=hdr)
    c, hdr2 = claire_amd.nifti_read(tmp_path / "copy.nii")
    np.testing.assert_array_equal(c, 2 * a)
    assert hdr2[76:108] == hdr[76:108]


def test_errors(tmp_path):
    bad = tmp_path / "bad.nii"
    bad.write_bytes(b"\0" * 400)
    with pytest.raises(claire_amd.ClaireError):
        claire_amd.nifti_read(bad)
    with pytest.raises(claire_amd.ClaireError):
        claire_amd.nifti_read(tmp_path / "missing.nii")


REF_DATA = "/root/reference/doc/data"


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's data files are not present")
@pytest.mark.parametrize("name", ["brain01.nii.gz", "brain02.nii.gz"])
def test_reference_brain_images(name):
    """The reference's example images (128×150×128 per doc/README-RUNME.md:68-72)."""
    path = os.path.join(REF_DATA, name)
    a, hdr = claire_amd.nifti_read(path)
    raw = gzip.decompress(open(path, "rb").read())
    np.testing.assert_array_equal(a, _decode(raw))
    assert sorted(a.shape) == [128, 128, 150]
