"""DMP format (common/myimage/my_image_io.py:222-250): header layout, round trip,
non-square images, dtype conversion and malformed files."""
import numpy as np
import pytest

from iterative_tomographic_reconstruction_algorithms_amd.dmp import read_dmp, write_dmp


def test_header_and_payload_layout(tmp_path):
    img = np.arange(12, dtype=np.float32).reshape(3, 4)      # height 3, width 4
    p = tmp_path / "a.DMP"
    write_dmp(str(p), img)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:6], np.uint16).tolist() == [4, 3, 0]
    np.testing.assert_array_equal(np.frombuffer(raw[6:], np.float32), img.ravel())
    assert len(raw) == 6 + 4 * img.size


@pytest.mark.parametrize("shape", [(256, 256), (304, 256), (1, 7)])
def test_round_trip_is_exact(tmp_path, shape):
    img = np.random.default_rng(0).standard_normal(shape).astype(np.float32)
    p = str(tmp_path / "x.DMP")
    write_dmp(p, img)
    out = read_dmp(p)
    assert out.dtype == np.float32 and out.shape == shape
    np.testing.assert_array_equal(out, img)


def test_float64_input_is_written_as_float32(tmp_path):
    img = np.linspace(0, 1, 20).reshape(4, 5)
    p = str(tmp_path / "d.DMP")
    write_dmp(p, img)
    np.testing.assert_array_equal(read_dmp(p), img.astype(np.float32))


def test_truncated_file_raises(tmp_path):
    p = tmp_path / "bad.DMP"
    p.write_bytes(np.array([4, 3, 0], np.uint16).tobytes() + b"\0" * 8)
    with pytest.raises(ValueError):
        read_dmp(str(p))
