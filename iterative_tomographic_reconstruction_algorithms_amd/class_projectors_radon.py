"""Drop-in for the reference's projector facade
``algorithms/class_projectors_radon.py:44-110`` (class ``projectors``), running the
radon projector pair on the GPU.

Same constructor and attributes (``nang``, ``filt``, ``plot``, ``angles`` = float32
radians, ``oper``).  ``A``/``At`` accept numpy arrays (returned as numpy) or float32 GPU
tensors (kept on the GPU, no host round trip).  The geometry is uploaded once per
object, lazily, so objects stay picklable for reference-style ``multiprocessing`` use.
"""
import numpy as np
import torch

from . import _dropin as D
from .core import Plan, filter_sino

myfloat = np.float32


class projectors:

    def __init__(self, npix, angles, oper="pd", ctr=0.0):
        nang = len(angles)
        angles1 = np.asarray(angles) * np.pi / 180.0       # class_projectors_radon.py:49
        self.nang = nang
        self.filt = "ramp"
        self.plot = True
        self.angles = angles1.astype(myfloat)              # class_projectors_radon.py:54
        self.oper = oper
        self.npix = int(npix)
        self._plan = None

    def __getstate__(self):
        st = dict(self.__dict__)
        st["_plan"] = None
        return st

    # C-order angles: dd and pd fftshift (radon_distance_driven.pyx:27,
    # radon_pixel_driven.pyx:30), rd and ss not.  The wrappers' degree test multiplies by
    # pi/180 (dd, rd) or 2*pi/180 (pd, ss: radon_pixel_driven.pyx:27-28,
    # radon_slant_stacking.pyx:27-28).
    def _c_angles(self):
        a = self.angles.copy()
        if a.size and np.max(a) > 2 * np.pi:
            a *= np.float32((2 * np.pi if self.oper in ("pd", "ss") else np.pi) / 180.0)
        return np.fft.fftshift(a) if self.oper in ("dd", "pd") else a

    def plan(self, npix=None):
        if self.oper not in ("dd", "rd", "pd", "ss"):
            raise NotImplementedError("projector %r is not one of 'pd', 'rd', 'dd', 'ss'"
                                      % self.oper)
        npix = self.npix if npix is None else int(npix)
        if self._plan is None or self._plan.npix != npix:
            # the facade passes method 1 to the pixel-driven and slant-stacking modules
            # (class_projectors_radon.py:62,68,75,81)
            self._plan = Plan(self.oper, npix, self._c_angles(), method=1)
        return self._plan

    @staticmethod
    def _as_f32(x):
        if isinstance(x, torch.Tensor):
            return x.to(torch.float32).contiguous()
        return np.ascontiguousarray(x, dtype=myfloat)

    # Forward projector (class_projectors_radon.py:60-68)
    def A(self, x):
        x = self._as_f32(x)
        p = self.plan(x.shape[-1])
        t, was_np = D.to_device(x)
        return D.from_device(p.forward(t), was_np)

    # Backprojector (class_projectors_radon.py:73-81)
    def At(self, x):
        x = self._as_f32(x)
        p = self.plan(x.shape[-1])
        t, was_np = D.to_device(x)
        return D.from_device(p.back(t), was_np)

    # Filtered backprojection (class_projectors_radon.py:86-110).  The reference raises
    # NameError (`dpc` is undefined at line 92) and back-projects the unfiltered array;
    # this follows the steps it spells out with the filtered array and dpc = False:
    # reverse the bins, filter, back-project, scale by pi/(2 nang), rotate by 180 degrees.
    def fbp(self, x):
        x = self._as_f32(x)
        p = self.plan(x.shape[-1])
        t, was_np = D.to_device(x)
        x0 = torch.flip(t, dims=(-1,)).contiguous()
        x0 = filter_sino(x0, ftype=self.filt)
        reco = p.back(x0)
        reco.mul_(np.float32(np.pi / (2.0 * self.nang)))
        reco = torch.flip(reco, dims=(-2, -1)).contiguous()
        return D.from_device(reco, was_np)
