"""Seeded synthetic phantoms (random ellipses / ellipsoids) for tests and benchmarks.

The field of view is [-1, 1]^2 sampled at pixel centres; row index r maps to
y = -1 + (r + 0.5) * 2 / N and column index q to x = -1 + (q + 0.5) * 2 / N.
Each ellipse has its centre uniform in the unit disk (r = sqrt(U1), phi = 2*pi*U2),
semi-axes uniform in [0.05, 0.4] * FOV width (i.e. [0.1, 0.8]), rotation uniform in
[0, pi) and intensity uniform in [0, 1]; ellipses add.  The generator is
``numpy.random.default_rng(seed)`` (PCG64), so a seed fixes the phantom.
"""
import numpy as np


def ellipse_params(n_ellipses, seed):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(0.0, 1.0, n_ellipses))
    phi = rng.uniform(0.0, 2.0 * np.pi, n_ellipses)
    cx, cy = r * np.cos(phi), r * np.sin(phi)
    a = rng.uniform(0.05, 0.4, n_ellipses) * 2.0
    b = rng.uniform(0.05, 0.4, n_ellipses) * 2.0
    rot = rng.uniform(0.0, np.pi, n_ellipses)
    val = rng.uniform(0.0, 1.0, n_ellipses)
    return np.stack([cx, cy, a, b, rot, val], axis=1)


def ellipses(npix, n_ellipses=10, seed=1, dtype=np.float32):
    """(npix, npix) additive random-ellipse phantom."""
    p = ellipse_params(n_ellipses, seed)
    c = -1.0 + (np.arange(npix) + 0.5) * (2.0 / npix)
    y, x = np.meshgrid(c, c, indexing="ij")
    img = np.zeros((npix, npix), dtype=np.float64)
    for cx, cy, a, b, rot, val in p:
        cr, sr = np.cos(rot), np.sin(rot)
        u = (x - cx) * cr + (y - cy) * sr
        w = -(x - cx) * sr + (y - cy) * cr
        img[(u / a) ** 2 + (w / b) ** 2 <= 1.0] += val
    return img.astype(dtype)


def ellipsoid_params(n_ellipsoids, seed):
    rng = np.random.default_rng(seed)
    # centre uniform in the unit ball
    d = rng.normal(size=(n_ellipsoids, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.0, 1.0, n_ellipsoids) ** (1.0 / 3.0)
    ctr = d * r[:, None]
    axes = rng.uniform(0.05, 0.4, (n_ellipsoids, 3)) * 2.0
    rot = rng.uniform(0.0, np.pi, n_ellipsoids)
    val = rng.uniform(0.0, 1.0, n_ellipsoids)
    return ctr, axes, rot, val


def ellipsoid_slices(npix, nslices, n_ellipsoids=60, seed=3000, dtype=np.float32,
                     first=0, count=None):
    """Slices [first, first+count) of an (nslices, npix, npix) random-ellipsoid volume
    (rotation about the slice axis)."""
    ctr, axes, rot, val = ellipsoid_params(n_ellipsoids, seed)
    count = nslices - first if count is None else count
    c = -1.0 + (np.arange(npix) + 0.5) * (2.0 / npix)
    y, x = np.meshgrid(c, c, indexing="ij")
    out = np.zeros((count, npix, npix), dtype=dtype)
    for k in range(count):
        z = -1.0 + (first + k + 0.5) * (2.0 / nslices)
        img = np.zeros((npix, npix), dtype=np.float64)
        for (cx, cy, cz), (a, b, h), ro, v in zip(ctr, axes, rot, val):
            dz = ((z - cz) / h) ** 2
            if dz > 1.0:
                continue
            cr, sr = np.cos(ro), np.sin(ro)
            u = (x - cx) * cr + (y - cy) * sr
            w = -(x - cx) * sr + (y - cy) * cr
            img[(u / a) ** 2 + (w / b) ** 2 + dz <= 1.0] += v
        out[k] = img
    return out


def textured(npix, n_ellipses=10, seed=1, texture=0.15, corr=0.04, dtype=np.float32):
    """Random ellipses plus a band-limited Gaussian random field texture on their
    support: white noise low-passed by a Gaussian of correlation length ``corr`` (in
    FOV units), normalised to unit standard deviation, scaled by ``texture`` and
    added where the ellipses are nonzero; the result is clipped at 0."""
    img = ellipses(npix, n_ellipses, seed, np.float64)
    rng = np.random.default_rng(seed + 7919)
    f = np.fft.fftfreq(npix) * npix / 2.0           # cycles per FOV half-width
    k2 = f[:, None] ** 2 + f[None, :] ** 2
    spec = np.exp(-2.0 * (np.pi * corr) ** 2 * k2)
    grf = np.real(np.fft.ifft2(np.fft.fft2(rng.standard_normal((npix, npix))) * spec))
    grf /= max(grf.std(), 1e-12)
    img = img + texture * grf * (img > 0)
    return np.maximum(img, 0.0).astype(dtype)


def angles_deg(nang, start=0.0, end=180.0):
    """Equally spaced projection angles in degrees, right end excluded
    (reference algorithms/utils.py:104-106)."""
    return np.linspace(start, end, nang, endpoint=False)
