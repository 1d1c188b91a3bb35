"""numpy restatement of the reference's gridding (gridrec) projector pair (test
infrastructure only: the GPU gridrec of the package is checked against it; nothing in the
product imports this file).

The reference builds these from C over FFTW3f, which this image lacks, so its own build
cannot run here; this restatement follows the C source line by line instead, in float64
with numpy's FFT (FFTW_FORWARD = numpy's fft, FFTW_BACKWARD = n * numpy's ifft, both
unnormalised):

  configure_regridding : common/operators/gridrec_module/gridrec_lut.py:268-296 (LUT of the
                         PSWF or Kaiser-Bessel kernel and the deapodisation matrix,
                         gridrec_lut.py:47-80, 89-131, 204-249)
  calc_filter          : pymodule_gridrec_v4/filters.c:151-261 (ramp magnitude from the FFT
                         of the spatial Ram-Lak taps, filters.c:17-38, times a window,
                         times the centre-of-rotation phase)
  backproj             : pymodule_gridrec_v4/gridrec_v4_backproj.c:64-446 (edge-padded
                         projection -> 1D FFT -> filter -> polar-to-Cartesian gridding
                         with the kernel LUT -> 2D inverse FFT -> deapodisation, crop)
  forwproj             : pymodule_gridrec_v4/gridrec_v4_forwproj.c:59-442 (deapodised
                         image -> 2D inverse FFT -> Cartesian-to-polar gridding -> phase
                         shift -> 1D FFT per angle, crop)

Angles are in degrees, as class_projectors_grid.py passes them (gridrec_v4_backproj.c:308,
SIN = sin(angle * PI / 180)).

Parity with the reference build is unpinned (no FFTW here).  The restatement is pinned by
what the mathematics fixes instead (tests/test_gridrec_oracle.py): its forward projection
of random ellipses against their analytic Radon transform, its filtered back projection of
that analytic sinogram against the phantom, and the LUT against the kernels' closed forms.
"""
import numpy as np

PI = 3.141592653589793
EPS = 1e-8          # gridrec_lut.py:13


# ---- kernel LUT and deapodiser (gridrec_lut.py) -------------------------------------

_PSWF_COEFFS = np.array([0.5767616E+02, 0.0, -0.8931343E+02, 0.0, 0.4167596E+02, 0.0,
                         -0.1053599E+02, 0.0, 0.1662374E+01, 0.0, -0.1780527E-00, 0.0,
                         0.1372983E-01, 0.0, -0.7963169E-03, 0.0, 0.3593372E-04, 0.0,
                         -0.1295941E-05, 0.0, 0.3817796E-07])
_KB_LAMBDA = np.array([[1.1, 0.9373402868152573], [1.2, 0.8762210440212161],
                       [1.3, 0.8296167052898601], [1.4, 0.7915122594523216],
                       [1.5, 0.7631378768943125], [1.6, 0.7391321589920385],
                       [1.7, 0.7186085894388523], [1.8, 0.7016852052327338],
                       [1.9, 0.6859972251713252], [2.0, 0.6735177854455913]], np.float32)


def grid_size(npix, oversampl):
    """gridrec_lut.py:277-279 / gridrec_v4_backproj.c:188-190."""
    n = int(2 ** (np.ceil(np.log2(npix))) * oversampl)
    while n % 4 != 0:
        n += 1
    return n


def _pswf(ns_grid, W, alpha, err, interp, ns_ker):
    """gridrec_lut.py:204-249."""
    if ns_ker != 2048:
        ns_ker = int(0.91 / (err * alpha)) if interp == "nn" else int((0.37 / err) / alpha)
    x = np.arange(0, ns_ker) / np.float32(ns_ker)
    lut = (np.polynomial.legendre.legval(x, _PSWF_COEFFS)
           / np.polynomial.legendre.legval(0.0, _PSWF_COEFFS))
    lut = np.pad(lut, (0, 5), "constant", constant_values=0)
    nh = int(ns_grid * 0.5)
    norm = np.float32(np.sqrt(1.0 / (W * 0.99998546)))
    ratio = np.float32(ns_ker) / np.float32(nh + 0.5)
    nn = lambda v: int(np.round(v))
    deapod = np.zeros(ns_grid, dtype=np.float32)
    deapod[nh] = norm / lut[nn(0.0)]
    for i in range(1, nh + 1):
        val = norm / (lut[nn(i * ratio)] + EPS)
        if i != nh:
            deapod[nh + i] = val
            deapod[nh - i] = val
        else:
            deapod[0] = val
    return lut, deapod


def _kb_lambda(alpha):
    """gridrec_lut.py:135-165 (the table is searched after reversing its rows)."""
    t = _KB_LAMBDA
    ind = np.argwhere(np.abs(t - alpha) == np.min(np.abs(t - alpha)))
    r = ind[0, 0]
    if alpha != t[r, 0]:
        if alpha > t[r, 0]:
            d = alpha - t[r, 0]
            return (1 - d) * t[r, 1] + d * t[r, 1]
        d = alpha - t[r - 1, 0]
        return (1 - d) * t[r - 1, 1] + d * t[r, 1]
    return t[r, 1]


def _kaiser_bessel(ns_grid, W, alpha, err, interp):
    """gridrec_lut.py:89-131."""
    beta = np.pi * np.sqrt(((W / alpha) * (alpha - 0.5)) ** 2 - 0.8)
    ns_ker = int(0.91 / (err * alpha)) if interp == "nn" else int(np.sqrt(0.37 / err) / alpha)
    k_max = W / np.float32(2 * ns_grid)
    k = np.linspace(0, k_max, ns_ker)
    lut = ns_grid / np.float32(W) * np.i0(beta * np.sqrt(1 - (2 * ns_grid * k / W) ** 2))
    lut /= lut[0]
    lut = np.pad(lut, (0, 4), "constant", constant_values=0)
    nh = int(ns_grid * 0.5)
    x = np.linspace(-nh, nh, ns_grid)
    f = np.sqrt(((np.pi * W * x / ns_grid) ** 2 - beta ** 2).astype(complex))
    deapod = np.abs(np.sin(f) / f)
    deapod[:] /= deapod[nh]
    deapod += EPS
    norm = np.float32(np.sqrt(1.0 / (W * _kb_lambda(alpha))))
    deapod[:] = norm / deapod
    return lut, deapod


def configure_regridding(npix, kernel="pswf", oversampl=2.0, interp="nn", W=7.0, errs=1e-3):
    """gridrec_lut.py:268-296: (W, lut, deapod 2-D) with W scaled by 2/pi."""
    W = W * 2 / np.pi
    nfreq = grid_size(npix, oversampl)
    ns_ker = 2048 if (kernel == "prolate" and oversampl == 2.0 and interp == "nn") else None
    if kernel in ("prolate", "pswf"):
        lut, deapod = _pswf(nfreq, W, oversampl, errs, interp, ns_ker)
    elif kernel in ("kaiser-bessel", "kb"):
        lut, deapod = _kaiser_bessel(nfreq, W, oversampl, errs, interp)
    else:
        raise ValueError(kernel)
    nh = int(nfreq * 0.5)                     # gridrec_lut.py:58-72: the (-1)^i shift
    sign = 1.0
    for i in range(1, nh + 1):
        sign = 1.0 if i % 2 == 0 else -1.0
        if i != nh:
            deapod[nh + i] *= sign
            deapod[nh - i] *= sign
        else:
            deapod[0] *= sign
    deapod[nh] *= sign
    return W, lut, np.outer(deapod, deapod)


# ---- the reconstruction filter (filters.c:151-261) ------------------------------------

FILTERS = ["none", "ramp", "shepp-logan", "hanning", "hamming", "lanczos", "parzen"]


def _window(kind, x, fm=0.5):
    """filters.c:45-136 for x >= 0."""
    if kind == 2:
        return 1.0 if x == 0 else (np.sin(PI * x / (2 * fm)) / (PI * x / (2 * fm)) if x < fm else 0.0)
    if kind == 3:
        return 0.5 * (1.0 + np.cos(PI * x / fm)) if abs(x) <= fm else 0.0
    if kind == 4:
        return 0.54 + 0.46 * np.cos(PI * x / fm) if abs(x) <= fm else 0.0
    if kind == 5:
        return 1.0 if x == 0 else (np.sin(PI * x / fm) / (PI * x / fm) if x < fm else 0.0)
    if kind == 6:
        f = abs(x) / fm
        if abs(x) <= fm / 2:
            return 1 - 6 * f * f * (1 - f)
        return 2 * (1 - f) ** 3 if abs(x) <= fm else 0.0
    raise ValueError(kind)


def calc_filter(nang, N, center, type_filter, radon_degree=0):
    """Complex filter[k], k < N/2 (filters.c:151-261)."""
    N2 = int(N * 0.5)
    norm = PI / N / nang
    ramp = None
    if type_filter and radon_degree == 0:
        c = (np.arange(N) - (N2 - 1)).astype(np.float64)      # filters.c:21-35
        odd = np.abs(c) % 2 == 1
        taps = np.zeros(N)
        taps[c == 0] = 0.25
        taps[odd] = -1.0 / (PI * PI * c[odd] ** 2)
        ramp = np.abs(np.fft.fft(taps))[:N2]
    out = np.zeros(N2, dtype=complex)
    for k in range(N2):
        x = k * (2 * PI * center / N)
        if radon_degree == 0:
            if type_filter and type_filter != 1:
                r = _window(type_filter, k / N) * ramp[k] * norm * N
            elif type_filter == 1:
                r = ramp[k] * norm * N
            else:
                r = 1.0
        else:
            if type_filter and type_filter != 1:
                r = _window(type_filter, k / N) * norm * N
            elif type_filter == 1:
                r = norm * N
            else:
                r = 1.0
        fr, fi = r * np.cos(x), -r * np.sin(x)
        if radon_degree:
            if k == 0:
                fr, fi = 0.0, 0.0
            elif type_filter:
                fr, fi = fi * (-1) / (2 * PI), fr / (2 * PI)
            else:
                fr, fi = fi * (-2 * PI * k) / N, fr * (2 * PI * k) / N
        out[k] = fr + 1j * fi
    return out


# ---- the projector pair -----------------------------------------------------------------

class Geometry:
    """The quantities both C functions derive from their arguments."""

    def __init__(self, npix, angles_deg, kernel="pswf", oversampl=2.0, interp="nn", W=7.0,
                 errs=1e-3, ctr=0.0, radon_degree=0):
        self.npix = npix
        self.angles = np.asarray(angles_deg, np.float32).reshape(-1)
        self.W, self.lut, self.deapod = configure_regridding(npix, kernel, oversampl, interp,
                                                             W, errs)
        self.ltbl = len(self.lut) - 5                      # class_projectors_grid.py:107
        self.pdim = grid_size(npix, oversampl)
        self.padleft = int((self.pdim - npix) * 0.5)
        self.lconv = float(np.float32(self.W))
        self.interp = 0 if interp == "nn" else 1
        self.radon_degree = radon_degree
        self.ctr = float(np.float32(ctr))
        rad = self.angles.astype(np.float64) * PI / 180
        # SIN, COS, U, V, the window and the LUT argument are float in the C
        # (backproj.c:308-325, filters.c:285-296); the LUT is float32
        # (class_projectors_grid.py:121)
        self.sin = np.sin(rad).astype(np.float32)
        self.cos = np.cos(rad).astype(np.float32)
        self.lut32 = self.lut.astype(np.float32)

    def kernel(self, x):
        """convolv_nn / convolv_lin (filters.c:285-296) at the float32 x = |offset| *
        tblspcg; returns float64 values of the float32 result."""
        x = np.asarray(x, np.float32)
        if self.interp == 0:
            return self.lut32[np.floor(x.astype(np.float64) + 0.5).astype(int)].astype(np.float64)
        xi = np.floor(x)
        d = (x - xi).astype(np.float32)
        i = xi.astype(int)
        one = np.float32(1)
        return ((one - d) * self.lut32[i] + d * self.lut32[i + 1]).astype(np.float64)

    def windows(self, w, n, lconv):
        """Cartesian neighbours of the polar samples w of angle n (backproj.c:316-325),
        in float32 as the C computes them."""
        p = self.pdim
        w = np.asarray(w, np.float32)
        U = w * self.cos[n] + np.float32(p // 2)
        V = w * self.sin[n] + np.float32(p // 2)
        h = np.float32(np.float32(lconv) * np.float32(0.5))
        iul = np.maximum(np.ceil(U - h), 0).astype(int)
        iuh = np.minimum(np.floor(U + h), p - 1).astype(int)
        ivl = np.maximum(np.ceil(V - h), 0).astype(int)
        ivh = np.minimum(np.floor(V + h), p - 1).astype(int)
        return U, V, iul, iuh, ivl, ivh

    def offsets(self, U, iu, tblspcg):
        """myAbs(U - iu) * tblspcg in float32 (backproj.c:331, 340)."""
        return (np.abs(np.float32(U) - iu.astype(np.float32)).astype(np.float32)
                * np.float32(tblspcg))


def backproj(g, sino, type_filter=0):
    """gridrec_v4_backproj.c:64-446 with flag_filter = 0 (the built-in filter of type
    `type_filter`: 0 for At, the filter index for fbp).  Returns the (npix, npix) image."""
    npix, p = g.npix, g.pdim
    ph = p // 2
    sino = np.asarray(sino, np.float64)
    nang = sino.shape[0]
    lconv = g.lconv
    tblspcg = np.float32(2 * g.ltbl) / np.float32(lconv)
    ctr = g.ctr if g.ctr != 0 else npix * 0.5
    if p != npix:
        ctr += g.padleft
    filt = calc_filter(nang, p, ctr, type_filter, g.radon_degree)
    H = np.zeros((p, p), dtype=complex)
    w = np.arange(ph, dtype=np.float64)
    for n in range(nang):
        proj = np.empty(p)
        edge_l = 0.0 if g.radon_degree else sino[n, 0]
        edge_r = 0.0 if g.radon_degree else sino[n, npix - 1]
        proj[:g.padleft] = edge_l
        proj[g.padleft:g.padleft + npix] = sino[n]
        proj[g.padleft + npix:] = edge_r
        P = np.fft.fft(proj) / p
        C1 = np.empty(ph, complex)
        C2 = np.zeros(ph, complex)
        C1[0] = filt[0].real * P[0]
        C1[1:] = filt[1:] * P[p - np.arange(1, ph)]
        C2[1:] = np.conj(filt[1:]) * P[1:ph]
        U, V, iul, iuh, ivl, ivh = g.windows(w, n, lconv)
        for j in range(ph):
            iu = np.arange(iul[j], iuh[j] + 1)
            iv = np.arange(ivl[j], ivh[j] + 1)
            cu = g.kernel(g.offsets(U[j], iu, tblspcg))
            cv = g.kernel(g.offsets(V[j], iv, tblspcg))
            conv = np.outer(cu, cv)
            H[np.ix_(iu, iv)] += conv * C1[j]
            if j != 0:
                m = (iu != 0)[:, None] & (iv != 0)[None, :]
                H2 = np.zeros_like(conv, dtype=complex)
                H2[m] = (conv * C2[j])[m]
                iu2 = (p - iu) % p
                iv2 = (p - iv) % p
                np.add.at(H, (iu2[:, None], iv2[None, :]), H2)
    H = np.fft.ifft2(H) * (p * p)                     # FFTW_BACKWARD, unnormalised
    sh = np.fft.fftshift(H)                           # j <-> iu = (j + p/2) mod p
    img = np.empty((npix, npix))
    jj = np.arange(g.padleft, g.padleft + npix)
    block = (g.deapod[np.ix_(jj, jj)] * sh.real[np.ix_(jj, jj)])   # [j][k]
    # rec[(npix - 1 - (k - padleft)) * npix + (j - padleft)]
    img[:, :] = block.T[::-1, :]
    return img


def forwproj(g, image):
    """gridrec_v4_forwproj.c:59-442.  Returns the (nang, npix) sinogram."""
    npix, p = g.npix, g.pdim
    ph = p // 2
    nang = g.angles.size
    img = np.asarray(image, np.float64)
    lconv = 2 * 7.0 * 1.0 / PI                        # forwproj.c:164 (the fixed C)
    lconv = float(np.float32(lconv))
    tblspcg = np.float32(2 * g.ltbl) / np.float32(lconv)
    ctr = npix * 0.5 + (g.padleft if p != npix else 0)
    k = np.arange(ph)
    x = k * (2 * PI * ctr / p)
    fr, fi = np.cos(x), np.sin(x)
    if g.radon_degree:
        fr, fi = fi * (2 * PI * k) / p, fr * (-2 * PI * k) / p
    jj = np.arange(g.padleft, g.padleft + npix)
    G = np.zeros((p, p))
    # H[iu][iv] = deapod[j][k] * I[(npix - 1 - (k - padleft)) * npix + (j - padleft)]
    G[np.ix_(jj, jj)] = g.deapod[np.ix_(jj, jj)] * img[::-1, :].T
    H = np.fft.ifftshift(G).astype(complex)
    H = np.fft.ifft2(H) * (p * p)                     # FFTW_BACKWARD
    sino = np.zeros((nang, npix))
    w = np.arange(ph, dtype=np.float64)
    for n in range(nang):
        U, V, iul, iuh, ivl, ivh = g.windows(w, n, lconv)
        cproj = np.zeros(p, complex)
        for j in range(ph):
            iu = np.arange(iul[j], iuh[j] + 1)
            iv = np.arange(ivl[j], ivh[j] + 1)
            cu = g.kernel(g.offsets(U[j], iu, tblspcg))
            cv = g.kernel(g.offsets(V[j], iv, tblspcg))
            conv = np.outer(cu, cv)
            d1 = np.sum(H[np.ix_(iu, iv)] * conv)
            ct = fr[j] - 1j * fi[j]
            if j != 0:
                m = (iu != 0)[:, None] & (iv != 0)[None, :]
                iu2 = (p - iu) % p
                iv2 = (p - iv) % p
                d2 = np.sum((H[np.ix_(iu2, iv2)] * conv)[m])
                cproj[p - j] = ct * d1
                cproj[j] = np.conj(ct) * d2
            else:
                cproj[0] = ct * d1
        cproj[ph] = 0.0
        out = np.fft.fft(cproj)
        sino[n] = out[g.padleft:g.padleft + npix].real / p
    return sino
