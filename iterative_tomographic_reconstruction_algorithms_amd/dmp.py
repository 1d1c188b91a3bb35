"""DMP image files, the reference's own raw format (common/myimage/my_image_io.py:222-250).

Layout: three 16-bit header values ``[width, height, 0]`` (written as uint16, read back
as int16 by the reference), then ``height x width`` float32 values in row-major order,
machine byte order.  ``read_dmp`` returns a float32 ``(height, width)`` array;
``write_dmp`` converts anything else to float32 first, and accepts GPU tensors (copied
to the host).
"""
import numpy as np


def read_dmp(path):
    """``readImageDmp`` (my_image_io.py:222-232)."""
    with open(path, "rb") as fd:
        header = np.fromfile(fd, np.int16, 3)
        if header.size != 3:
            raise ValueError("%s: truncated DMP header" % path)
        width, height = int(header[0]), int(header[1])
        data = np.fromfile(fd, np.float32, -1)
    if width < 0 or height < 0 or data.size != width * height:
        raise ValueError("%s: header says %d x %d but %d float32 values follow"
                         % (path, height, width, data.size))
    return data.reshape(height, width).astype(np.float32, copy=False)


def write_dmp(path, image):
    """``writeImageDmp`` (my_image_io.py:237-246)."""
    if hasattr(image, "detach"):          # torch tensor, possibly on the GPU
        image = image.detach().cpu().numpy()
    arr = np.asarray(image)
    if arr.ndim != 2:
        raise ValueError("DMP holds one 2-D image, got shape %s" % (arr.shape,))
    height, width = arr.shape
    if width > 32767 or height > 32767:
        raise ValueError("DMP header holds sizes up to 32767 (read back as int16)")
    with open(path, "wb") as fd:
        np.array([width, height, 0], np.uint16).tofile(fd)
        np.ascontiguousarray(arr, dtype=np.float32).tofile(fd)
