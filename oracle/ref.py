"""ctypes bindings of the reference's own radon C (test infrastructure only).

``oracle/Makefile`` (``make ref``) compiles the reference's
``common/operators/radon_module/*/radon_{dd,rd,pd,ss}.c`` where they lie into
``oracle/_ref/libradon_<proj>_<flags>.so`` with ``<flags>`` in ``fast`` (the reference's
own ``-O3 -ffast-math`` flags) and ``strict`` (IEEE).  Those libraries are git-ignored;
when they are absent every function here raises :class:`FileNotFoundError` and the
tests that need them skip.

The C signatures bound here are ``radon_dd.c:54`` / ``radon_rd.c:76``
(``void f(float* image, int npix, float* angles, int nang, int oper, float* sino)``)
and ``radon_pd.c:20`` / ``radon_ss.c:93`` (an extra ``int method`` before ``sino``).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
_libs = {}
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def path(proj, flags):
    return os.path.join(REF_DIR, "libradon_%s_%s.so" % (proj, flags))


def available(proj="dd", flags="fast"):
    return os.path.exists(path(proj, flags))


def _fn(proj, flags):
    key = (proj, flags)
    if key not in _libs:
        p = path(proj, flags)
        if not os.path.exists(p):
            raise FileNotFoundError(p)
        L = ctypes.CDLL(p)
        f = getattr(L, "radon_" + proj)
        if proj in ("pd", "ss"):
            f.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int, _f32p]
        else:
            f.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, _f32p]
        f.restype = None
        _libs[key] = f
    return _libs[key]


def _angles(angles, proj):
    a = np.array(angles, dtype=np.float32, copy=True).reshape(-1)
    if a.size and np.max(a) > 2 * np.pi:
        # radon_distance_driven.pyx:25-26 (pd/ss multiply by 2*pi/180 instead:
        # radon_pixel_driven.pyx:27-28)
        a *= np.float32((2 * np.pi if proj in ("pd", "ss") else np.pi) / 180.0)
    if proj in ("dd", "pd"):      # radon_slant_stacking.pyx:30 has the fftshift commented out
        a = np.fft.fftshift(a)
    return np.ascontiguousarray(a, dtype=np.float32)


# The pixel-driven C with method 0 indexes sino[v*npix + l] without a range check
# (radon_pd.c:58-63); l stays within (-npix/4, 5*npix/4), so one zero guard row on each
# side of the sinogram keeps every access inside the buffer handed to the C.  Forward
# contributions that land in the guards are dropped; in the back projection the guards
# read as zeros, i.e. they contribute nothing.
_GUARD = 1


def _ss_image_guard(npix, a, flags, method):
    """Zero rows to put before / after the image so that every flat index the
    slant-stacking C reaches for these angles stays inside the buffer (for lines that
    miss the image square it indexes rows outside the image, radon_ss.c:121-191)."""
    from . import radon
    _, (lo, hi) = radon.ss_edge_bins(npix, a, method,
                                     radon.FAST if flags == "fast" else radon.IEEE)
    before = max(0, -(lo // npix)) + 1
    after = max(0, hi // npix - (npix - 1)) + 1
    return before, after


def _ss_call(f, x, a, oper, method, flags):
    npix = x.shape[1]
    before, after = _ss_image_guard(npix, a, flags, method)
    buf = np.zeros((before + npix + after, npix), dtype=np.float32)
    img = buf[before:before + npix]
    if oper == 0:
        img[:] = x
        sino = np.zeros((a.size, npix), dtype=np.float32)
        f(img, npix, a, a.size, 0, method, sino)
        return sino
    f(img, npix, a, a.size, 1, method, np.ascontiguousarray(x, dtype=np.float32))
    return img.copy()


def forward(proj, image, angles, flags="fast", method=1):
    f = _fn(proj, flags)
    img = np.ascontiguousarray(image, dtype=np.float32)
    npix = img.shape[0]
    a = _angles(angles, proj)
    if proj == "ss":
        return _ss_call(f, img, a, 0, method, flags)
    if proj == "pd":
        g = np.zeros((a.size + 2 * _GUARD, npix), dtype=np.float32)
        f(img, npix, a, a.size, 0, method, g[_GUARD:])
        return g[_GUARD:_GUARD + a.size].copy()
    sino = np.zeros((a.size, npix), dtype=np.float32)
    f(img, npix, a, a.size, 0, sino)
    return sino


def back(proj, sino, angles, flags="fast", method=1):
    f = _fn(proj, flags)
    s = np.ascontiguousarray(sino, dtype=np.float32)
    nang, npix = s.shape
    a = _angles(angles, proj)
    if proj == "ss":
        return _ss_call(f, s, a, 1, method, flags)
    image = np.zeros((npix, npix), dtype=np.float32)
    if proj == "pd":
        g = np.zeros((nang + 2 * _GUARD, npix), dtype=np.float32)
        g[_GUARD:_GUARD + nang] = s
        f(image, npix, a, nang, 1, method, g[_GUARD:])
    else:
        f(image, npix, a, nang, 1, s)
    return image


def penalty(kind, image, delta=0.1, flags="fast"):
    """(R_num, R_den) from the reference's huber.c / tikhonov.c / haar.c, called the way
    penalty.pyx:16-106 calls them (fresh zero outputs; tikhonov den = 2, haar den = 0)."""
    p = os.path.join(REF_DIR, "libpenalty_%s.so" % flags)
    if not os.path.exists(p):
        raise FileNotFoundError(p)
    L = ctypes.CDLL(p)
    img = np.ascontiguousarray(image, dtype=np.float32)
    nr, nc = img.shape
    rn = np.zeros((nr, nc), np.float32)
    if kind == "huber":
        L.huber.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_char_p,
                            _f32p]
        rd = np.zeros((nr, nc), np.float32)
        L.huber(img, nr, nc, float(delta), b"nom", rn)
        L.huber(img, nr, nc, float(delta), b"den", rd)
        return rn, rd
    f = getattr(L, kind)
    f.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p]
    f(img, nr, nc, rn)
    return rn, np.full((nr, nc), 2.0 if kind == "tikhonov" else 0.0, np.float32)
