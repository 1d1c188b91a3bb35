"""The LDS window of the distance-driven forward's staged path, checked by brute force.

For a block of B bins and views with at most NC candidate pixels per bin and line (bins
narrower than NC - 1.25 pixels), the kernel (csrc/dd_device.h, fwd_staged) copies W
pixels of each line into LDS, starting at q0 = (floor(fl(D0lo + K)) + nh - 2) & ~3, and
every bin reads its NC candidates from its first pixel x (the one-test search, clamped
into the line) at offset x + nh - q0, clamped into [0, W - NC].  The clamp must never act
on a bin that meets the line, or the bin would read shifted candidates.  Here, in float32
as the kernel computes it, every (block, line, bin) of full-size geometries whose bins sit
at the top of each candidate bracket is checked against W = (NC-1) B + max(0, NC + 7 - B/4)
(rounded up to 4); the same check finds overflows with W = (NC-1) B, the earlier size,
where the last bin of a block starts left of the line's first pixel.
"""
import numpy as np
import pytest

f32 = np.float32


def extra(b, nc):
    e = nc + 7 - b // 4
    return (e + 3) // 4 * 4 if e > 0 else 0


def window_overflows(npix, w, b, nc, width):
    """(block, line, bin) triples of a view with bins w pixels wide whose candidates do
    not fit a window of `width` pixels; bins that miss the line are skipped."""
    nh = npix // 2
    rinv = f32(w)
    t = f32(np.sqrt(float(w) ** 2 - 1.0))             # tan of the view's angle
    lines = np.arange(npix)
    K = (((lines - nh).astype(f32) + f32(0.5)) * t).astype(f32)[:, None]
    P0 = (f32(-nh) - K).astype(f32)
    PN = (f32(npix - nh) - K).astype(f32)
    bad = 0
    for d0 in range(0, npix, b):
        d = np.arange(d0, min(d0 + b, npix))
        D0lo = f32(f32(d0 - nh) * rinv)
        est = np.floor((D0lo + K).astype(f32))
        est = np.clip(est, -nh - width, npix - nh + width)
        q0 = (est.astype(np.int64) + nh - 2) & ~3
        Dlo = ((d - nh).astype(f32) * rinv).astype(f32)[None, :]
        Dhi = ((d + 1 - nh).astype(f32) * rinv).astype(f32)[None, :]
        Dlh = (Dlo + f32(0.5)).astype(f32)
        fe = np.floor((Dlh + K).astype(f32))
        fe = np.where(Dlo < (fe - K).astype(f32), fe - 1, fe)
        x = np.clip(fe, -nh, npix - 1 - nh).astype(np.int64)
        off = x + nh - q0
        meets = (Dhi > P0) & (Dlo < PN)
        bad += int(np.count_nonzero(meets & ((off < 0) | (off > width - nc))))
    return bad


# bins just below the top of each staged bracket (12, 16, 24, 32: the launch of the wide
# views; 6: the main launch), with the bracket of the candidate count they get
CASES = [(30.74, 32), (30.5, 32), (22.74, 24), (14.74, 16), (14.70, 16), (10.74, 12),
         (4.74, 6), (3.74, 6), (2.74, 4), (1.74, 3)]


@pytest.mark.parametrize("b", [64, 128, 256])
@pytest.mark.parametrize("w,nc", CASES)
def test_window_holds_every_candidate(w, nc, b):
    assert int(np.floor(w + 0.25)) + 2 <= nc
    width = (nc - 1) * b + extra(b, nc)
    assert window_overflows(2048, w, b, nc, width) == 0


def test_the_earlier_window_overflows():
    # (NC-1) B alone is too small for 64-bin blocks at the top of the wide brackets (30.5:
    # the 88.125-degree view of a 1440-view list, 14.74: the 86.1-degree view of 1800)
    for w, nc in [(30.74, 32), (30.5, 32), (22.74, 24), (14.74, 16)]:
        assert window_overflows(2048, w, 64, nc, (nc - 1) * 64) > 0, (w, nc)
