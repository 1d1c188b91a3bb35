"""TIFF reader/writer for the reference's grey-scale images (my_image_io.py:259-267):
round trips over the supported types, byte order, and the reference's own data files
when the reference tree is present."""
import os

import numpy as np
import pytest

from iterative_tomographic_reconstruction_algorithms_amd.tiff import read_tiff, write_tiff

REF_DATA = "/root/reference/data"


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8, np.uint16, np.int32])
def test_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(1)
    img = (rng.random((37, 53)) * 200).astype(dtype)
    p = str(tmp_path / "x.tif")
    write_tiff(p, img)
    out = read_tiff(p)
    assert out.dtype == np.dtype(dtype) and out.shape == img.shape
    np.testing.assert_array_equal(out, img)


def test_big_endian_file(tmp_path):
    # hand-built big-endian float32 file, 2 x 3, one strip
    import struct
    img = np.arange(6, dtype=">f4").reshape(2, 3)
    entries = [(256, 4, 3), (257, 4, 2), (258, 3, 32), (259, 3, 1), (273, 4, 8 + 2 + 12 * 7 + 4),
               (279, 4, 24), (339, 3, 3)]
    out = bytearray(b"MM" + struct.pack(">HI", 42, 8) + struct.pack(">H", len(entries)))
    for tag, typ, val in entries:
        out += (struct.pack(">HHIHH", tag, typ, 1, val, 0) if typ == 3
                else struct.pack(">HHII", tag, typ, 1, val))
    out += struct.pack(">I", 0) + img.tobytes()
    p = tmp_path / "be.tif"
    p.write_bytes(bytes(out))
    np.testing.assert_array_equal(read_tiff(str(p)), img.astype(np.float32))


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference data not present")
def test_reads_reference_data_files():
    sino = read_tiff(os.path.join(REF_DATA, "shepp_logan_pix0256_ang0304_sino.tif"))
    assert sino.shape == (304, 256) and sino.dtype == np.float32
    ph = read_tiff(os.path.join(REF_DATA, "shepp_logan_pix0256.tif"))
    assert ph.shape == (256, 256) and ph.dtype == np.float64
    assert np.isfinite(ph).all() and ph.max() > 0
