"""ctypes bindings of the C restatement in ``radon_oracle.c`` (test infrastructure only).

The wrappers reproduce what the reference's Cython modules do around the C call:

* ``radon_distance_driven.pyx:14-62``: degrees -> radians when ``max(angles) > 2*pi``,
  ``np.fft.fftshift`` of the angles, a zeroed float32 output, ``radon_dd(...)``.
* ``radon_ray_driven.pyx:15-61``: the same without the fftshift.
* ``radon_pixel_driven.pyx:15-73``: degrees -> radians by ``2*pi/180`` (sic) when
  ``max(angles) > 2*pi``, fftshift, zeroed output, ``radon_pd(..., method, ...)``.
* ``radon_slant_stacking.pyx:15-73``: the same without the fftshift.

Unlike the reference wrappers they never modify the caller's angle array in place.
"""
import ctypes
import os

import numpy as np

IEEE = 0   # the C as written, IEEE float32, no contraction
FAST = 1   # the C as gcc -O3 -ffast-math compiles it (the reference's own flags)
F64 = 2    # double-precision restatement

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    """Compile ``liboracle.so`` with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-C", _HERE, "_build/liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("oracle_radon_dd", "oracle_radon_rd"):
            f = getattr(L, name)
            f.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, _f32p,
                          ctypes.c_int]
            f.restype = ctypes.c_int
        for name in ("oracle_radon_dd_f64", "oracle_radon_rd_f64"):
            f = getattr(L, name)
            f.argtypes = [_f64p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, _f64p]
            f.restype = ctypes.c_int
        for kind in ("pd", "ss"):
            f = getattr(L, "oracle_radon_%s" % kind)
            f.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int, _f32p, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(L, "oracle_radon_%s_f64" % kind)
            f.argtypes = [_f64p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int, _f64p]
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _prep_angles(angles, shift, factor=np.pi / 180.0):
    a = np.array(angles, dtype=np.float32, copy=True).reshape(-1)
    if a.size and np.max(a) > 2 * np.pi:
        a *= np.float32(factor)  # in-place float32 multiply, as the .pyx does
    if shift:
        a = np.fft.fftshift(a)
    return np.ascontiguousarray(a, dtype=np.float32)


def _call(kind, x, angles, oper, mode, shift):
    L = lib()
    npix = x.shape[1] if oper == 0 else x.shape[1]
    a32 = _prep_angles(angles, shift)
    nang = a32.size
    if mode == F64:
        a = a32.astype(np.float64)
        fn = getattr(L, "oracle_radon_%s_f64" % kind)
        if oper == 0:
            img = np.ascontiguousarray(x, dtype=np.float64)
            out = np.zeros((nang, npix), dtype=np.float64)
            rc = fn(img, npix, a, nang, 0, out)
        else:
            sino = np.ascontiguousarray(x, dtype=np.float64)
            out = np.zeros((npix, npix), dtype=np.float64)
            rc = fn(out, npix, a, nang, 1, sino)
    else:
        fn = getattr(L, "oracle_radon_%s" % kind)
        if oper == 0:
            img = np.ascontiguousarray(x, dtype=np.float32)
            out = np.zeros((nang, npix), dtype=np.float32)
            rc = fn(img, npix, a32, nang, 0, out, mode)
        else:
            sino = np.ascontiguousarray(x, dtype=np.float32)
            out = np.zeros((npix, npix), dtype=np.float32)
            rc = fn(out, npix, a32, nang, 1, sino, mode)
    if rc != 0:
        raise MemoryError("oracle_radon_%s failed" % kind)
    return out


def dd_forward(image, angles, mode=FAST):
    """Distance-driven forward projection, (N,N) -> (V,N), sinogram rows in the
    reference's order (fftshift + reversal)."""
    return _call("dd", image, angles, 0, mode, shift=True)


def dd_back(sino, angles, mode=FAST):
    return _call("dd", sino, angles, 1, mode, shift=True)


def rd_forward(image, angles, mode=FAST):
    """Ray-driven forward projection, (N,N) -> (V,N)."""
    return _call("rd", image, angles, 0, mode, shift=False)


def rd_back(sino, angles, mode=FAST):
    return _call("rd", sino, angles, 1, mode, shift=False)


def _call_pd(x, angles, oper, mode, method, init=None, kind="pd"):
    L = lib()
    # radon_pixel_driven.pyx:27-30 / radon_slant_stacking.pyx:27-30: degrees times 2*pi/180,
    # fftshift for pd only (commented out in the slant-stacking wrapper)
    a32 = _prep_angles(angles, kind == "pd", 2 * np.pi / 180.0)
    nang = a32.size
    npix = x.shape[1]
    dt = np.float64 if mode == F64 else np.float32
    src = np.ascontiguousarray(x, dtype=dt)
    shape = (nang, npix) if oper == 0 else (npix, npix)
    # the C adds into its output; the wrapper hands it zeros, `init` another start
    out = (np.zeros(shape, dtype=dt) if init is None
           else np.array(init, dtype=dt, copy=True).reshape(shape))
    if oper == 0:
        args = (src, npix, None, nang, 0, int(method), out)
    else:
        args = (out, npix, None, nang, 1, int(method), src)
    if mode == F64:
        rc = getattr(L, "oracle_radon_%s_f64" % kind)(args[0], npix, a32.astype(np.float64),
                                                      nang, oper, int(method), args[6])
    else:
        rc = getattr(L, "oracle_radon_%s" % kind)(args[0], npix, a32, nang, oper,
                                                  int(method), args[6], mode)
    if rc != 0:
        raise MemoryError("oracle_radon_%s failed" % kind)
    return out


def pd_forward(image, angles, mode=FAST, method=1, init=None):
    """Pixel-driven forward projection, (N,N) -> (V,N); sinogram row v holds view v of
    the fftshift-ed angle list.  The facade always passes method 1
    (class_projectors_radon.py:62).  ``init``: the sinogram the C adds into (zeros)."""
    return _call_pd(image, angles, 0, mode, method, init)


def pd_back(sino, angles, mode=FAST, method=1, init=None):
    return _call_pd(sino, angles, 1, mode, method, init)


def ss_forward(image, angles, mode=FAST, method=1, init=None):
    """Slant-stacking forward projection, (N,N) -> (V,N); sinogram row V-1-v holds view
    v (radon_ss.c:111); no fftshift.  The facade passes method 1
    (class_projectors_radon.py:68)."""
    return _call_pd(image, angles, 0, mode, method, init, kind="ss")


def ss_back(sino, angles, mode=FAST, method=1, init=None):
    return _call_pd(sino, angles, 1, mode, method, init, kind="ss")


def ss_edge_bins(npix, c_angles, method=1, mode=FAST):
    """(flags, (lo, hi)) for the slant-stacking C receiving ``c_angles`` (already
    converted to radians): flags[row, bin] = 1 where the reference steps or samples
    outside the image square, (lo, hi) = the flat image-index extent it reaches."""
    a = np.ascontiguousarray(c_angles, dtype=np.float32)
    flags = np.zeros((a.size, npix), dtype=np.uint8)
    ext = np.zeros(2, dtype=np.int64)
    L = lib()
    f = L.oracle_ss_edge_bins
    f.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                  np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
    f.restype = ctypes.c_int
    f(npix, a, a.size, int(method), 1 if mode == FAST else 0, flags, ext)
    return flags.astype(bool), (int(ext[0]), int(ext[1]))


def forward(oper, image, angles, mode=FAST):
    return {"dd": dd_forward, "rd": rd_forward, "pd": pd_forward,
            "ss": ss_forward}[oper](image, angles, mode)


def back(oper, sino, angles, mode=FAST):
    return {"dd": dd_back, "rd": rd_back, "pd": pd_back, "ss": ss_back}[oper](sino, angles, mode)
