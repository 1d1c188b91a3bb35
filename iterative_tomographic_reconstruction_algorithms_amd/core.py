"""Device-side projector plans: the Python face of the libitomo C ABI.

A :class:`Plan` holds one geometry (projector, image side, angles in the order the
reference C receives them) uploaded to the GPU once, and runs the projectors and the
fused SIRT / MLEM steps on PyTorch ROCm tensors (float32, C-contiguous, on the current
device) on torch's current stream.  Nothing here computes on the CPU.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

__all__ = ["Plan", "device", "stream_handle", "dot_f64"]


def device():
    if not torch.cuda.is_available():
        raise L.ItomoError("no ROCm GPU is visible: libitomo runs on MI355X (gfx950) only")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check_dev(t, name, numel=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor on the GPU" % name)
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError("%s must be a C-contiguous float32 GPU tensor (got %s %s%s)"
                         % (name, t.dtype, t.device, "" if t.is_contiguous() else ", strided"))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s has %d elements, expected %d" % (name, t.numel(), numel))


class Plan:
    """Geometry of one projector on the GPU.

    ``angles`` are radians exactly as the reference C receives them: already
    ``fftshift``-ed for the distance- and pixel-driven projectors
    (radon_distance_driven.pyx:27, radon_pixel_driven.pyx:30), as given for the
    ray-driven and slant-stacking ones.  ``arith`` selects the float arithmetic: ``"fast"`` (the reference
    built with its own -ffast-math flags, the default) or ``"ieee"`` (the reference C as
    written).  ``method`` is the pixel-driven and slant-stacking projectors' interpolation
    (0 nearest, 1 linear, radon_pd.c:57-85, radon_ss.c:123-190; the facade uses 1) and is
    ignored by the others.
    """

    def __init__(self, oper, npix, angles, arith="fast", method=1):
        if oper not in L.PROJ:
            raise ValueError("projector %r not supported on the GPU path (have %s)"
                             % (oper, sorted(L.PROJ)))
        self.oper = oper
        self.npix = int(npix)
        self.angles = np.ascontiguousarray(np.asarray(angles, dtype=np.float32).reshape(-1))
        self.nang = int(self.angles.size)
        self.arith = arith
        self.method = int(method) if oper in ("pd", "ss") else None
        device()
        h = ctypes.c_void_p()
        aptr = self.angles.ctypes.data_as(ctypes.c_void_p)
        if oper in ("pd", "ss"):
            create = getattr(L.lib(), "itomo_plan_create_" + oper)
            rc = create(ctypes.byref(h), self.npix, aptr, self.nang, self.method,
                        L.ARITH[arith])
        else:
            rc = L.lib().itomo_plan_create(ctypes.byref(h), L.PROJ[oper], self.npix, aptr,
                                           self.nang, L.ARITH[arith])
        L.check(rc, "plan_create")
        self._h = h

    # -- lifetime -------------------------------------------------------------------
    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and h.value:
            torch.cuda.current_stream().synchronize()
            L.lib().itomo_plan_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- shapes ---------------------------------------------------------------------
    @property
    def image_numel(self):
        return self.npix * self.npix

    @property
    def sino_numel(self):
        return self.nang * self.npix

    def _nslices(self, t, per):
        if t.numel() % per:
            raise ValueError("array of %d elements is not a whole number of %d-element "
                             "slices" % (t.numel(), per))
        return t.numel() // per

    def trig(self):
        """(sin, cos) per view as the plan computed them (host float32 arrays)."""
        s = np.empty(self.nang, np.float32)
        c = np.empty(self.nang, np.float32)
        L.check(L.lib().itomo_plan_trig(self._h, s.ctypes.data_as(ctypes.c_void_p),
                                        c.ctypes.data_as(ctypes.c_void_p)), "plan_trig")
        return s, c

    # -- launch shapes --------------------------------------------------------------
    @property
    def shapes(self):
        """(forward variant, back variant) of the distance-driven launches."""
        f, b = ctypes.c_int(), ctypes.c_int()
        L.check(L.lib().itomo_plan_get_shapes(self._h, ctypes.byref(f), ctypes.byref(b)),
                "plan_get_shapes")
        return f.value, b.value

    def set_shapes(self, fwd_variant, back_variant):
        """Select launch shapes (see include/itomo.h); results are bit-identical for all."""
        torch.cuda.current_stream().synchronize()
        L.check(L.lib().itomo_plan_set_shapes(self._h, int(fwd_variant), int(back_variant)),
                "plan_set_shapes")

    # -- ray-driven system matrix ----------------------------------------------------
    def rd_matrix(self, mode="query"):
        """Bytes of the cached ray-driven system matrix (0 if none).  mode "build" builds it
        now, "drop" frees it and keeps the walks; both give bit-identical projections."""
        modes = {"query": 0, "build": 1, "drop": 2}
        nbytes = ctypes.c_size_t()
        L.check(L.lib().itomo_plan_rd_matrix(self._h, modes[mode], ctypes.byref(nbytes),
                                             stream_handle()), "plan_rd_matrix")
        return nbytes.value

    # -- projectors -----------------------------------------------------------------
    def forward(self, image, out=None, accumulate=False):
        """A image: (..., N, N) -> (..., V, N)."""
        _check_dev(image, "image")
        nz = self._nslices(image, self.image_numel)
        if out is None:
            out = torch.zeros(image.shape[:-2] + (self.nang, self.npix), dtype=torch.float32,
                              device=image.device)
        _check_dev(out, "out", nz * self.sino_numel)
        L.check(L.lib().itomo_forward(self._h, _p(image), _p(out), nz, int(bool(accumulate)),
                                      stream_handle()), "forward")
        return out

    def back(self, sino, out=None, accumulate=False):
        """A^T sino: (..., V, N) -> (..., N, N)."""
        _check_dev(sino, "sino")
        nz = self._nslices(sino, self.sino_numel)
        if out is None:
            out = torch.zeros(sino.shape[:-2] + (self.npix, self.npix), dtype=torch.float32,
                              device=sino.device)
        _check_dev(out, "out", nz * self.image_numel)
        L.check(L.lib().itomo_back(self._h, _p(sino), _p(out), nz, int(bool(accumulate)),
                                   stream_handle()), "back")
        return out

    # -- SIRT -----------------------------------------------------------------------
    def sirt_weights(self):
        """(R, C) = (1/A(1), 1/A^T(1)) with zero sums replaced by 1e-10 (art.py:226-237)."""
        dev = device()
        R = torch.empty((self.nang, self.npix), dtype=torch.float32, device=dev)
        C = torch.empty((self.npix, self.npix), dtype=torch.float32, device=dev)
        L.check(L.lib().itomo_sirt_weights(self._h, _p(R), _p(C), stream_handle()),
                "sirt_weights")
        return R, C

    def sirt_iterate(self, x, b, R, C, work, pc_beta=-1.0, stats=None):
        """One fused SIRT step in place on x (art.py:212, 314-321).  ``stats``: optional
        float64 GPU tensor of 3*nslices entries (see include/itomo.h)."""
        _check_dev(x, "x")
        nz = self._nslices(x, self.image_numel)
        _check_dev(b, "b", nz * self.sino_numel)
        _check_dev(R, "R", self.sino_numel)
        _check_dev(C, "C", self.image_numel)
        _check_dev(work, "work", nz * self.sino_numel)
        if stats is not None and (stats.dtype != torch.float64 or not stats.is_cuda
                                  or stats.numel() < 3 * nz):
            raise ValueError("stats must be a float64 GPU tensor of >= 3*nslices entries")
        L.check(L.lib().itomo_sirt_iterate(self._h, _p(x), _p(b), _p(R), _p(C), _p(work),
                                           nz, float(pc_beta), _p(stats), stream_handle()),
                "sirt_iterate")
        return x

    def _stats_ok(self, stats, nz):
        if stats is not None and (stats.dtype != torch.float64 or not stats.is_cuda
                                  or stats.numel() < 3 * nz):
            raise ValueError("stats must be a float64 GPU tensor of >= 3*nslices entries")

    def sirt_forward(self, x, b, R, work, stats=None):
        """First half of a SIRT step: work = R * (b - A x); stats[3z] = sum (b-Ax)^2."""
        _check_dev(x, "x")
        nz = self._nslices(x, self.image_numel)
        _check_dev(b, "b", nz * self.sino_numel)
        _check_dev(R, "R", self.sino_numel)
        _check_dev(work, "work", nz * self.sino_numel)
        self._stats_ok(stats, nz)
        L.check(L.lib().itomo_sirt_forward(self._h, _p(x), _p(b), _p(R), _p(work), nz,
                                           _p(stats), stream_handle()), "sirt_forward")
        return work

    def sirt_back(self, work, C, x, pc_beta=-1.0, stats=None):
        """Second half: x = x + C * A^T work, then x[x<0] *= pc_beta when pc_beta >= 0."""
        _check_dev(x, "x")
        nz = self._nslices(x, self.image_numel)
        _check_dev(C, "C", self.image_numel)
        _check_dev(work, "work", nz * self.sino_numel)
        self._stats_ok(stats, nz)
        L.check(L.lib().itomo_sirt_back(self._h, _p(work), _p(C), _p(x), nz, float(pc_beta),
                                        _p(stats), stream_handle()), "sirt_back")
        return x

    # -- MLEM -----------------------------------------------------------------------
    def mlem_forward(self, x, b, work, stats=None):
        _check_dev(x, "x")
        nz = self._nslices(x, self.image_numel)
        _check_dev(b, "b", nz * self.sino_numel)
        _check_dev(work, "work", nz * self.sino_numel)
        self._stats_ok(stats, nz)
        L.check(L.lib().itomo_mlem_forward(self._h, _p(x), _p(b), _p(work), nz, _p(stats),
                                           stream_handle()), "mlem_forward")
        return work

    def mlem_back(self, work, inv_d, x, stats=None):
        _check_dev(x, "x")
        nz = self._nslices(x, self.image_numel)
        _check_dev(inv_d, "inv_d", self.image_numel)
        _check_dev(work, "work", nz * self.sino_numel)
        self._stats_ok(stats, nz)
        L.check(L.lib().itomo_mlem_back(self._h, _p(work), _p(inv_d), _p(x), nz, _p(stats),
                                        stream_handle()), "mlem_back")
        return x

    def mlem_weights(self):
        """1 / A^T(1) with zero sums replaced by 1e-10 (sir.py:215-217, 283)."""
        inv_d = torch.empty((self.npix, self.npix), dtype=torch.float32, device=device())
        L.check(L.lib().itomo_mlem_weights(self._h, _p(inv_d), stream_handle()),
                "mlem_weights")
        return inv_d

    def mlem_iterate(self, x, b, inv_d, work, stats=None):
        _check_dev(x, "x")
        nz = self._nslices(x, self.image_numel)
        _check_dev(b, "b", nz * self.sino_numel)
        _check_dev(inv_d, "inv_d", self.image_numel)
        _check_dev(work, "work", nz * self.sino_numel)
        if stats is not None and (stats.dtype != torch.float64 or not stats.is_cuda
                                  or stats.numel() < 3 * nz):
            raise ValueError("stats must be a float64 GPU tensor of >= 3*nslices entries")
        L.check(L.lib().itomo_mlem_iterate(self._h, _p(x), _p(b), _p(inv_d), _p(work), nz,
                                           _p(stats), stream_handle()), "mlem_iterate")
        return x


_taps = {}


def filter_taps(npix, ftype="ramp"):
    """FP64 taps of the reference's sinogram filter (filters.py:35-113) on the GPU."""
    key = (int(npix), ftype)
    t = _taps.get(key)
    if t is None:
        h = np.empty(npix, dtype=np.float64)
        L.check(L.lib().itomo_filter_taps(int(npix), L.FILTERS[ftype],
                                          h.ctypes.data_as(ctypes.c_void_p)), "filter_taps")
        t = _taps[key] = torch.from_numpy(h).to(device())
    return t


def filter_sino(sino, ftype="ramp"):
    """filters.filter_proj on the GPU: each row convolved with the filter taps."""
    _check_dev(sino, "sino")
    npix = sino.shape[-1]
    taps = filter_taps(npix, ftype)
    out = torch.empty_like(sino)
    L.check(L.lib().itomo_filter_rows(_p(sino), _p(out), sino.numel() // npix, npix, _p(taps),
                                      stream_handle()), "filter_rows")
    return out


_dot_work = {}


def dot_f64(a, b=None):
    """<a, b> accumulated in FP64 on the GPU; returns a 0-d float64 GPU tensor."""
    b = a if b is None else b
    _check_dev(a, "a")
    _check_dev(b, "b", a.numel())
    dev = a.device
    work = _dot_work.get(dev)
    if work is None:
        work = _dot_work[dev] = torch.empty(L.DOT_WORK, dtype=torch.float64, device=dev)
    out = torch.empty((), dtype=torch.float64, device=dev)
    L.check(L.lib().itomo_dot_f64(_p(a), _p(b), a.numel(), _p(out), _p(work),
                                  stream_handle()), "dot_f64")
    return out
