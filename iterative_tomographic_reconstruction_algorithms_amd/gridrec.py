"""Gridding (gridrec) projector pair on the GPU: the host-side set-up of the reference's
regridding projector and the Python face of the libitomo gridrec calls.

Set-up (numpy, once per geometry), as the reference computes it:
  * the interpolation kernel's look-up table and the 1-D deapodiser, for the prolate
    spheroidal (PSWF) or Kaiser-Bessel kernel (gridrec_lut.py:47-80, 89-131, 204-296);
  * per-angle sin/cos of the angle in degrees, rounded to float
    (gridrec_v4_backproj.c:308-309);
  * the back projection's filter: the magnitude of the FFT of the spatial Ram-Lak taps,
    times a window, times the centre-of-rotation phase (filters.c:151-261); and the
    forward projection's centre phase (gridrec_v4_forwproj.c:183-199).
Everything else (FFTs, gridding, deapodisation) runs in csrc/gridrec.hip.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from .core import _check_dev, device, stream_handle

__all__ = ["configure_regridding", "grid_size", "calc_filter", "GridrecPlan", "FILTERS"]

PI = 3.141592653589793
_EPS = 1e-8

# class_projectors_grid.py:43-45: index = the filter type of filters.c:188-208
FILTERS = [("none", ""), ("ramp", ""), ("shepp-logan", "shlo"), ("hanning", "hann"),
           ("hamming", "hamm"), ("lanczos", "lanc"), ("parzen", "parz")]

_PSWF = np.array([0.5767616E+02, 0.0, -0.8931343E+02, 0.0, 0.4167596E+02, 0.0, -0.1053599E+02,
                  0.0, 0.1662374E+01, 0.0, -0.1780527E-00, 0.0, 0.1372983E-01, 0.0,
                  -0.7963169E-03, 0.0, 0.3593372E-04, 0.0, -0.1295941E-05, 0.0, 0.3817796E-07])
# Kaiser-Bessel normalisation per oversampling (gridrec_lut.py:137-147)
_KB_NORM = np.array([[1.1, 0.9373402868152573], [1.2, 0.8762210440212161],
                     [1.3, 0.8296167052898601], [1.4, 0.7915122594523216],
                     [1.5, 0.7631378768943125], [1.6, 0.7391321589920385],
                     [1.7, 0.7186085894388523], [1.8, 0.7016852052327338],
                     [1.9, 0.6859972251713252], [2.0, 0.6735177854455913]], np.float32)


def grid_size(npix, oversampl):
    """Side of the Cartesian Fourier grid: 2^ceil(log2 npix) * oversampl, rounded up to a
    multiple of 4 (gridrec_lut.py:277-279)."""
    n = int(2 ** (np.ceil(np.log2(npix))) * oversampl)
    return n + (-n) % 4


def _kb_norm(alpha):
    t = _KB_NORM
    r = int(np.argwhere(np.abs(t - alpha) == np.min(np.abs(t - alpha)))[0, 0])
    if alpha == t[r, 0]:
        return t[r, 1]
    if alpha > t[r, 0]:
        return t[r, 1]                                   # (1 - d) v + d v
    d = alpha - t[r - 1, 0]
    return (1 - d) * t[r - 1, 1] + d * t[r, 1]


def configure_regridding(npix, kernel="pswf", oversampl=2.0, interp="nn", W=7.0, errs=1e-3):
    """(W scaled by 2/pi, kernel LUT, 1-D deapodiser with the (-1)^i centring signs): the
    reference's configure_regridding (gridrec_lut.py:268-296) keeps the outer product
    of the deapodiser as its 2-D matrix; the GPU forms that product on the fly."""
    W = W * 2 / np.pi
    n = grid_size(npix, oversampl)
    nh = n // 2
    if kernel in ("prolate", "pswf"):
        ns = 2048 if (kernel == "prolate" and oversampl == 2.0 and interp == "nn") else (
            int(0.91 / (errs * oversampl)) if interp == "nn" else int((0.37 / errs) / oversampl))
        x = np.arange(ns) / np.float32(ns)
        lut = (np.polynomial.legendre.legval(x, _PSWF)
               / np.polynomial.legendre.legval(0.0, _PSWF))
        lut = np.concatenate([lut, np.zeros(5)])
        norm = np.float32(np.sqrt(1.0 / (W * 0.99998546)))
        ratio = np.float32(ns) / np.float32(nh + 0.5)
        i = np.arange(1, nh + 1)
        # i * ratio in float32, as the reference's int times numpy float32 scalar
        vals = norm / (lut[np.round(i.astype(np.float32) * ratio).astype(int)] + _EPS)
        dea = np.zeros(n, np.float32)
        dea[nh] = norm / lut[0]
        dea[nh + i[:-1]] = vals[:-1]
        dea[nh - i[:-1]] = vals[:-1]
        dea[0] = vals[-1]
    elif kernel in ("kaiser-bessel", "kb"):
        beta = np.pi * np.sqrt(((W / oversampl) * (oversampl - 0.5)) ** 2 - 0.8)
        ns = int(0.91 / (errs * oversampl)) if interp == "nn" else int(np.sqrt(0.37 / errs) / oversampl)
        k = np.linspace(0, W / np.float32(2 * n), ns)
        lut = n / np.float32(W) * np.i0(beta * np.sqrt(1 - (2 * n * k / W) ** 2))
        lut = np.concatenate([lut / lut[0], np.zeros(4)])
        x = np.linspace(-nh, nh, n)
        f = np.sqrt(((np.pi * W * x / n) ** 2 - beta ** 2).astype(complex))
        dea = np.abs(np.sin(f) / f)
        dea = dea / dea[nh] + _EPS
        dea = np.float32(np.sqrt(1.0 / (W * _kb_norm(oversampl)))) / dea
    else:
        raise ValueError("gridding kernel %r (have 'pswf', 'prolate', 'kb')" % (kernel,))
    i = np.arange(1, nh + 1)
    sign = np.where(i % 2 == 0, 1.0, -1.0)
    dea = np.asarray(dea, np.float64)
    dea[nh + i[:-1]] *= sign[:-1]
    dea[nh - i[:-1]] *= sign[:-1]
    dea[0] *= sign[-1]
    dea[nh] *= sign[-1]
    return W, lut, dea


def _window(kind, x, fm=0.5):
    if kind == 2:      # shepp-logan
        return np.where(x == 0, 1.0, np.where(x < fm, np.sinc(x / (2 * fm)), 0.0))
    if kind == 3:      # hanning
        return np.where(np.abs(x) <= fm, 0.5 * (1.0 + np.cos(PI * x / fm)), 0.0)
    if kind == 4:      # hamming
        return np.where(np.abs(x) <= fm, 0.54 + 0.46 * np.cos(PI * x / fm), 0.0)
    if kind == 5:      # lanczos
        return np.where(x == 0, 1.0, np.where(x < fm, np.sinc(x / fm), 0.0))
    if kind == 6:      # parzen
        f = np.abs(x) / fm
        return np.where(np.abs(x) <= fm / 2, 1 - 6 * f * f * (1 - f),
                        np.where(np.abs(x) <= fm, 2 * (1 - f) ** 3, 0.0))
    raise ValueError(kind)


def calc_filter(nang, n, center, ftype, radon_degree=0):
    """Complex filter[k], k < n/2 (filters.c:151-261)."""
    k = np.arange(n // 2)
    norm = PI / n / nang
    if ftype and radon_degree == 0:
        c = (np.arange(n) - (n // 2 - 1)).astype(np.float64)
        taps = np.where(c == 0, 0.25, 0.0)
        odd = np.abs(c) % 2 == 1
        taps[odd] = -1.0 / (PI * PI * c[odd] ** 2)
        ramp = np.abs(np.fft.fft(taps))[: n // 2]
    if radon_degree == 0:
        if ftype == 1:
            r = ramp * norm * n
        elif ftype:
            r = _window(ftype, k / n) * ramp * norm * n
        else:
            r = np.ones(n // 2)
    else:
        r = (norm * n * np.ones(n // 2) if ftype == 1 else
             _window(ftype, k / n) * norm * n if ftype else np.ones(n // 2))
    x = k * (2 * PI * center / n)
    fr, fi = r * np.cos(x), -r * np.sin(x)
    if radon_degree:
        if ftype:
            fr, fi = fi * (-1) / (2 * PI), fr / (2 * PI)
        else:
            fr, fi = fi * (-2 * PI * k) / n, fr * (2 * PI * k) / n
        fr[0] = fi[0] = 0.0
    return fr + 1j * fi


def _c64(z):
    """Complex array -> contiguous float32 (re, im) pairs."""
    return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], -1), dtype=np.float32)


class GridrecPlan:
    """One gridding geometry on the GPU (class_projectors_grid.projectors' set-up):
    angles in degrees, npix detector bins, the reference's defaults for kernel, grid
    oversampling, kernel interpolation and size.  ``filt`` names the filter of the
    filtered back projection (default ramp)."""

    def __init__(self, npix, angles_deg, ctr=0.0, kernel="pswf", oversampl=2.0, interp="nn",
                 radon_degree=0, W=7.0, errs=1e-3, filt=None):
        device()
        self.npix = int(npix)
        self.angles = np.ascontiguousarray(np.asarray(angles_deg, np.float32).reshape(-1))
        self.nang = self.angles.size
        self.W, self.lut, self.deapod = configure_regridding(self.npix, kernel, oversampl,
                                                             interp, W, errs)
        self.ltbl = len(self.lut) - 5                 # class_projectors_grid.py:107
        self.pdim = grid_size(self.npix, oversampl)
        self.interp = 0 if interp == "nn" else 1
        self.radon_degree = int(radon_degree)
        if filt is None:
            self.filter_type = 1
        else:
            names = [n for pair in FILTERS for n in pair if n]
            if filt not in names:
                raise ValueError("filter %r (have %s)" % (filt, names))
            self.filter_type = [i for i, pair in enumerate(FILTERS) if filt in pair][0]
        padleft = int((self.pdim - self.npix) * 0.5)
        c_back = (np.float32(ctr) if ctr != 0 else np.float32(self.npix * 0.5))
        c_back = float(c_back + (padleft if self.pdim != self.npix else 0))
        c_fwd = float(np.float32(self.npix * 0.5) + (padleft if self.pdim != self.npix else 0))
        rad = self.angles.astype(np.float64) * PI / 180
        trig = np.ascontiguousarray(np.stack([np.sin(rad), np.cos(rad)], -1), np.float32)
        ph = self.pdim // 2
        k = np.arange(ph)
        xf = k * (2 * PI * c_fwd / self.pdim)
        cr, ci = np.cos(xf), np.sin(xf)
        if self.radon_degree:
            cr, ci = ci * (2 * PI * k) / self.pdim, cr * (-2 * PI * k) / self.pdim
        phase = cr - 1j * ci                          # (CtmpR, CtmpI), forwproj.c:363-364
        f_at = calc_filter(self.nang, self.pdim, c_back, 0, self.radon_degree)
        f_fbp = calc_filter(self.nang, self.pdim, c_back, self.filter_type, self.radon_degree)
        lut32 = np.ascontiguousarray(self.lut, np.float32)
        dea32 = np.ascontiguousarray(self.deapod, np.float32)
        lconv_fwd = float(np.float32(2 * 7.0 / PI))    # gridrec_v4_forwproj.c:164
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self._keep = [trig, lut32, dea32, _c64(f_at), _c64(f_fbp), _c64(phase)]
        h = ctypes.c_void_p()
        L.check(L.lib().itomo_gridrec_create(
            ctypes.byref(h), self.npix, self.nang, self.pdim, vp(trig), vp(lut32), lut32.size,
            self.ltbl, vp(dea32), float(np.float32(self.W)), lconv_fwd, self.interp,
            self.radon_degree, vp(self._keep[3]), vp(self._keep[4]), vp(self._keep[5])),
            "gridrec_create")
        self._h = h

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and h.value:
            torch.cuda.current_stream().synchronize()
            L.lib().itomo_gridrec_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, image):
        """(..., N, N) float32 GPU tensor -> (..., V, N) sinogram."""
        _check_dev(image, "image")
        nz = image.numel() // (self.npix * self.npix)
        if nz * self.npix * self.npix != image.numel():
            raise ValueError("image is not a stack of %dx%d slices" % (self.npix, self.npix))
        out = torch.empty(image.shape[:-2] + (self.nang, self.npix), dtype=torch.float32,
                          device=image.device)
        L.check(L.lib().itomo_gridrec_forward(self._h, ctypes.c_void_p(image.data_ptr()),
                                              ctypes.c_void_p(out.data_ptr()), nz,
                                              stream_handle()), "gridrec_forward")
        return out

    def back(self, sino, filtered=False):
        """(..., V, N) float32 GPU tensor -> (..., N, N) image; ``filtered`` applies the
        plan's reconstruction filter (the reference's fbp)."""
        _check_dev(sino, "sino")
        nz = sino.numel() // (self.nang * self.npix)
        if nz * self.nang * self.npix != sino.numel():
            raise ValueError("sinogram is not a stack of %dx%d slices" % (self.nang, self.npix))
        out = torch.empty(sino.shape[:-2] + (self.npix, self.npix), dtype=torch.float32,
                          device=sino.device)
        L.check(L.lib().itomo_gridrec_back(self._h, ctypes.c_void_p(sino.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), nz, int(filtered),
                                           stream_handle()), "gridrec_back")
        return out

    # -- SIRT on the gridding pair (art.py:211-237 with the grid-pswf / grid-kb projectors):
    # the projections, then the elementwise halves of the step (itomo_sirt_residual /
    # itomo_sirt_update), with the method signatures of core.Plan so that solvers.sirt
    # drives either.
    @property
    def image_numel(self):
        return self.npix * self.npix

    @property
    def sino_numel(self):
        return self.nang * self.npix

    def _work(self, nz, n):
        need = max(1, 3 * nz * ((n + 2047) // 2048))
        w = getattr(self, "_epi_work", None)
        if w is None or w.numel() < need:
            w = torch.empty(need, dtype=torch.float64, device=device())
            self._epi_work = w
        return w

    def sirt_weights(self):
        """(R, C) = (1/A(1), 1/A^T(1)) with zero sums replaced by 1e-10 (art.py:226-237)."""
        dev = device()
        ones_img = torch.ones((self.npix, self.npix), dtype=torch.float32, device=dev)
        ones_sino = torch.ones((self.nang, self.npix), dtype=torch.float32, device=dev)
        R = self.forward(ones_img)
        C = self.back(ones_sino)
        st = stream_handle()
        L.check(L.lib().itomo_reciprocal(ctypes.c_void_p(R.data_ptr()),
                                         ctypes.c_void_p(R.data_ptr()), R.numel(), st),
                "reciprocal")
        L.check(L.lib().itomo_reciprocal(ctypes.c_void_p(C.data_ptr()),
                                         ctypes.c_void_p(C.data_ptr()), C.numel(), st),
                "reciprocal")
        return R, C

    def sirt_forward(self, x, b, R, work, stats=None):
        """work = R * (b - A x); stats[3z] = sum (b - A x)^2."""
        nz = x.numel() // self.image_numel
        work.copy_(self.forward(x.contiguous()).reshape(work.shape))
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        L.check(L.lib().itomo_sirt_residual(p(work), p(b), p(R), self.sino_numel, nz, p(stats),
                                            p(self._work(nz, self.sino_numel)),
                                            stream_handle()), "sirt_residual")
        return work

    def sirt_back(self, work, C, x, pc_beta=-1.0, stats=None):
        """x = x + C * A^T work, then x[x < 0] *= pc_beta when pc_beta >= 0."""
        nz = x.numel() // self.image_numel
        atw = self.back(work)
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        L.check(L.lib().itomo_sirt_update(p(atw), p(C), p(x), self.image_numel, nz,
                                          float(pc_beta), p(stats),
                                          p(self._work(nz, self.image_numel)), stream_handle()),
                "sirt_update")
        return x

    def sirt_iterate(self, x, b, R, C, work, pc_beta=-1.0, stats=None):
        self.sirt_forward(x, b, R, work, stats)
        return self.sirt_back(work, C, x, pc_beta, stats)
