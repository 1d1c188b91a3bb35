"""The one-test searches of the distance-driven kernels, checked by brute force.

The forward's staged and gather paths find the first pixel of a bin on a line as
floor(fl(Dlo + 1/2) + K) (Dlo + 1/2 rounded once per bin, then the sum rounded per line),
corrected once downwards if Dlo < P_fe; the back walk finds the bin
of a pixel's left end lo as floor(lo * est + 1/2), corrected once downwards if lo < D_fd.
Both rest on a rounding bound: the estimate is the exact index or the one after it, for
every npix <= 32766 (the forward only for bins narrower than 30.75 pixels, the views it
is used for).  Here the estimate is compared with the exact index -- the largest q with
fl(q - K) <= Dlo, the largest j with fl(j * rinv) <= lo, searched around the double
quotient -- on random keys of extreme geometries, in float32 as the kernels compute them.
"""
import numpy as np
import pytest

f32 = np.float32


def view(rinv_target, sign):
    s = f32(1.0 / rinv_target)
    c = f32(np.sqrt(max(0.0, 1.0 - float(s) ** 2)) * sign)
    if float(s) == 1.0:
        c = f32(1e-3 * sign)
    rinv = f32(f32(1) / s)
    return s, rinv, f32(c * rinv)


def largest(pred, start):
    best = start - 4
    for k in range(-3, 4):
        best = np.where(pred(start + k), np.maximum(best, start + k), best)
    return best


@pytest.mark.parametrize("npix", [32766, 32765, 2048, 63])
@pytest.mark.parametrize("rinv_target", [1.0000001, 1.3, 1.74, 4.7, 14.7, 30.7, 1000.0, 50000.0])
def test_one_test_searches_hit_the_exact_index(npix, rinv_target):
    rng = np.random.default_rng(npix + int(rinv_target * 10))
    nh = npix // 2
    for sign in (1, -1):
        s, rinv, t = view(rinv_target, sign)
        n = 20000
        u = (rng.integers(0, npix, n) - nh).astype(f32) + f32(0.5)
        K = (u * t).astype(f32)
        if rinv_target < 30.75:   # forward: first pixel of bin j on the line with offset K
            j = rng.integers(-nh - 1, npix - nh + 1, n).astype(f32)
            D = (j * rinv).astype(f32)
            Dh = (D + f32(0.5)).astype(f32)
            e = np.floor((Dh + K).astype(f32))
            e = np.where(D < (e - K).astype(f32), e - 1, e)
            q = largest(lambda q: (q.astype(f32) - K).astype(f32) <= D,
                        np.floor(D.astype(np.float64) + K.astype(np.float64)))
            np.testing.assert_array_equal(e, q)
        # back: bin of the pixel left end lo = fl(x - K)
        x = (rng.integers(0, npix, n) - nh).astype(f32)
        lo = (x - K).astype(f32)
        fd = np.floor(((lo * s).astype(f32) + f32(0.5)).astype(f32))
        fd = np.where(lo < (fd * rinv).astype(f32), fd - 1, fd)
        jj = largest(lambda j: (j.astype(f32) * rinv).astype(f32) <= lo,
                     np.floor(lo.astype(np.float64) / float(rinv)))
        np.testing.assert_array_equal(fd, jj)
