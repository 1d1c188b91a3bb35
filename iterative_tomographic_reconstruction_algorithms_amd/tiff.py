"""Single-image TIFF files as the reference reads and writes them through
``skimage.io`` (common/myimage/my_image_io.py:259-267): one grey-scale image, no
compression.  skimage is not a dependency here; this is a small reader/writer for that
case: baseline uncompressed strips, one sample per pixel, unsigned/signed integers of
8-64 bits or IEEE floats of 32/64 bits, either byte order.  ``write_tiff`` writes
little-endian float32 (or the array's own integer/float type) in one strip.
"""
import struct

import numpy as np

_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 16: ("Q", 8)}


def _tags(buf, e, off):
    n = struct.unpack_from(e + "H", buf, off)[0]
    tags = {}
    for k in range(n):
        tag, typ, cnt = struct.unpack_from(e + "HHI", buf, off + 2 + 12 * k)
        if typ not in _TYPES:
            continue
        fmt, size = _TYPES[typ]
        pos = off + 10 + 12 * k
        if size * cnt > 4:
            pos = struct.unpack_from(e + "I", buf, pos)[0]
        if typ == 2:
            tags[tag] = buf[pos:pos + cnt]
        else:
            tags[tag] = list(struct.unpack_from(e + fmt * cnt, buf, pos))
    return tags


def read_tiff(path):
    """First image of an uncompressed grey-scale TIFF as a (rows, cols) array."""
    with open(path, "rb") as fd:
        buf = fd.read()
    if buf[:2] not in (b"II", b"MM"):
        raise ValueError("%s: not a TIFF file" % path)
    e = "<" if buf[:2] == b"II" else ">"
    if struct.unpack_from(e + "H", buf, 2)[0] != 42:
        raise ValueError("%s: not a classic TIFF file" % path)
    t = _tags(buf, e, struct.unpack_from(e + "I", buf, 4)[0])
    width, height = t[256][0], t[257][0]
    bits = t.get(258, [1])[0]
    if t.get(259, [1])[0] != 1:
        raise ValueError("%s: compressed TIFF is not supported" % path)
    if t.get(277, [1])[0] != 1:
        raise ValueError("%s: only one sample per pixel is supported" % path)
    kind = {1: "u", 2: "i", 3: "f"}.get(t.get(339, [1])[0])
    if kind is None or bits not in (8, 16, 32, 64) or (kind == "f" and bits < 32):
        raise ValueError("%s: unsupported sample format" % path)
    dtype = np.dtype(e + kind + str(bits // 8))
    data = b"".join(buf[o:o + c] for o, c in zip(t[273], t[279]))
    if len(data) < width * height * dtype.itemsize:
        raise ValueError("%s: strips hold fewer pixels than the header says" % path)
    arr = np.frombuffer(data, dtype, width * height).reshape(height, width)
    return arr.astype(dtype.newbyteorder("="))


def write_tiff(path, image):
    """One uncompressed little-endian strip; float64 input is written as float64, other
    floats as float32, integers as they are."""
    if hasattr(image, "detach"):
        image = image.detach().cpu().numpy()
    arr = np.asarray(image)
    if arr.ndim != 2:
        raise ValueError("one 2-D image expected, got shape %s" % (arr.shape,))
    if arr.dtype.kind == "f" and arr.dtype != np.float64:
        arr = arr.astype(np.float32)
    if arr.dtype.kind not in "uif" or arr.dtype.itemsize not in (1, 2, 4, 8):
        raise ValueError("unsupported dtype %s" % arr.dtype)
    arr = np.ascontiguousarray(arr, dtype=arr.dtype.newbyteorder("<"))
    height, width = arr.shape
    fmt = {"u": 1, "i": 2, "f": 3}[arr.dtype.kind]
    entries = [(256, 4, width), (257, 4, height), (258, 3, 8 * arr.dtype.itemsize),
               (259, 3, 1), (262, 3, 1), (273, 4, 0), (277, 3, 1), (278, 4, height),
               (279, 4, arr.nbytes), (339, 3, fmt)]
    ifd = 8
    data_off = ifd + 2 + 12 * len(entries) + 4
    out = bytearray(b"II" + struct.pack("<HI", 42, ifd))
    out += struct.pack("<H", len(entries))
    for tag, typ, val in entries:
        if tag == 273:
            val = data_off
        if typ == 3:
            out += struct.pack("<HHIHH", tag, typ, 1, val, 0)
        else:
            out += struct.pack("<HHII", tag, typ, 1, val)
    out += struct.pack("<I", 0)
    with open(path, "wb") as fd:
        fd.write(bytes(out))
        fd.write(arr.tobytes())
