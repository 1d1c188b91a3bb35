"""The back walk's single-crossing bound, checked by brute force.

A view whose bin width is not within rounding of the pixel width lets each pixel span
hold at most one bin key, so the back walk takes one crossing per pixel with selects;
only views flagged `multi` (plan.hip, dd_multi_cross: (r - 1) <= 4 * M * 2^-23 with
r = min(rinv, 1/|b|) and M = (npix + 4) * (1 + |t| + r)) run the loop for further
crossings.  Here every unflagged view of several geometries is checked in float32: the
smallest gap between consecutive bin keys (both arithmetic forms) exceeds the largest
pixel span on every line.
"""
import numpy as np
import pytest

f32 = np.float32


def multi_flag(rinv, t, b, npix):
    r = min(float(rinv), 1.0 / abs(float(b)))
    m = (npix + 4.0) * (1.0 + abs(float(t)) + r)
    return (r - 1.0) <= 4.0 * m * 2.0 ** -23


@pytest.mark.parametrize("npix", [16, 63, 256, 2048])
def test_unflagged_views_have_one_key_per_pixel_span(npix):
    nh = npix // 2
    deg = list(np.linspace(0, 180, 360, endpoint=False))
    deg += [90 + e for e in (1e-3, -1e-3, 1e-2, -1e-2, 0.5, -0.5, 3.6, 3.7)]
    deg += [e for e in (1e-3, 1e-2, 0.5, 3.6, 3.7)]
    flagged = checked = 0
    for th in (np.asarray(deg) * np.pi / 180).astype(f32):
        s, c = f32(np.sin(th)), f32(np.cos(th))
        if abs(float(s)) < 1e-7 or abs(float(c)) < 1e-7:
            continue
        if float(th) > np.pi / 2:
            rinv, b = f32(1) / s, s
            t = f32(c * rinv)
        else:
            rinv, b = f32(1) / c, c
            t = f32(s * rinv)
        if multi_flag(rinv, t, b, npix):
            flagged += 1
            continue
        checked += 1
        j = np.arange(-nh - 2, npix - nh + 3, dtype=f32)
        gap = min(np.diff((j * rinv).astype(f32)).min(), np.diff((j / b).astype(f32)).min())
        u = (np.arange(npix) - nh).astype(f32) + f32(0.5)
        K = (u * t).astype(f32)
        x = np.arange(-nh - 1, npix - nh + 2, dtype=f32)
        span = np.diff((x[None, :] - K[:, None]).astype(f32), axis=1).max()
        assert gap > span, (npix, float(th))
    assert flagged > 0 and checked > 300
