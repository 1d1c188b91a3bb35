"""ctypes bindings of the penalty restatement in ``penalty_oracle.c`` (test
infrastructure only), in the shape of the reference's ``penalty`` module
(common/operators/regularization_module/penalty.pyx:16-106): ``*_num(image, delta)`` and
``*_den(image, delta)`` return float32 arrays of the image's shape."""
import ctypes

import numpy as np

from .radon import FAST, IEEE, lib

KINDS = {"huber": 0, "tikhonov": 1, "haar": 2}
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def penalty(kind, image, delta=0.1, mode=FAST):
    """(R_num, R_den) of one 2-D image."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    nr, nc = img.shape
    rn = np.zeros((nr, nc), np.float32)
    rd = np.zeros((nr, nc), np.float32)
    f = lib().oracle_penalty
    f.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                  ctypes.c_int, _f32p, _f32p]
    f.restype = ctypes.c_int
    f(img, nr, nc, KINDS[kind], float(delta), 1 if mode == FAST else 0, rn, rd)
    return rn, rd


def sobel_magnitude(x):
    """sir.py:339-343: sqrt(gx^2 + gy^2) of scipy.ndimage.sobel along both axes with
    mode='reflect', in float32 (a library routine)."""
    from scipy import ndimage as nd
    x = np.asarray(x, np.float32)
    gx = nd.sobel(x, axis=0, mode="reflect")
    gy = nd.sobel(x, axis=1, mode="reflect")
    return np.sqrt(gx ** 2 + gy ** 2)
