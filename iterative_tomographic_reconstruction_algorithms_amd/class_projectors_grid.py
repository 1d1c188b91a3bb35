"""Drop-in for the reference's gridding projector facade
``algorithms/class_projectors_grid.py:54-151`` (class ``projectors``), running gridrec on
the GPU (gridrec.py, csrc/gridrec.hip).

Same constructor, defaults and attributes (``lut``, ``deapod`` as the 2-D float32 matrix,
``angles`` in degrees as float32, ``param_nofilt``, ``param_filt``).  ``A``, ``At`` and
``fbp`` accept numpy arrays (returned as numpy) or float32 GPU tensors (kept on the GPU).
The reference prints its settings on construction; this one does not.
"""
import numpy as np
import torch

from .gridrec import FILTERS, GridrecPlan

myfloat = np.float32

# class_projectors_grid.py:43-45
filter_list = np.array([list(f) for f in FILTERS])


class projectors:

    def __init__(self, npix, angles, ctr=0.0, kernel="pswf", oversampl=2.0, interp="nn",
                 radon_degree=0, W=7.0, errs=1e-3, filt=None, args=None):
        if args is not None:                      # class_projectors_grid.py:59-80
            kernel, oversampl = args.kernel, args.oversampl
            interp, W, errs = args.kernel_interp, args.kernel_size, args.errs
            radon_degree = 1 if (getattr(args, "dpc", False) is True
                                 or getattr(args, "dbp", False) is True) else 0
            filt = getattr(args, "filt", None)
        self._plan = GridrecPlan(npix, np.asarray(angles), ctr, kernel, oversampl, interp,
                                 radon_degree, W, errs, filt)
        p = self._plan
        interp1 = 0 if interp == "nn" else 1
        self.lut = p.lut.astype(myfloat)
        self.deapod = np.outer(p.deapod.astype(myfloat), p.deapod.astype(myfloat))
        self.angles = np.asarray(angles).astype(myfloat)
        self.param_nofilt = np.array([ctr, 0, 0, oversampl, interp1, len(p.lut) - 5, p.W,
                                      radon_degree]).astype(myfloat)
        self.param_filt = np.array([ctr, p.filter_type, 0, oversampl, interp1, len(p.lut) - 5,
                                    p.W, radon_degree]).astype(myfloat)

    def plan(self, npix=None):
        """The GPU geometry (gridrec.GridrecPlan); the solvers drive it like core.Plan."""
        if npix is not None and int(npix) != self._plan.npix:
            raise ValueError("this projector is set up for %d pixels, not %d"
                             % (self._plan.npix, npix))
        return self._plan

    @staticmethod
    def _in(x):
        if isinstance(x, torch.Tensor):
            return x.to(torch.float32).contiguous(), True
        return torch.from_numpy(np.ascontiguousarray(x, dtype=myfloat)).cuda(), False

    @staticmethod
    def _out(t, on_gpu):
        return t if on_gpu else t.cpu().numpy()

    def A(self, x):
        """Forward projection (gridrec_v4.pyx:67-94)."""
        xd, g = self._in(x)
        return self._out(self._plan.forward(xd), g)

    def At(self, x):
        """Back projection without filter (gridrec_v4.pyx:32-62, param_nofilt)."""
        xd, g = self._in(x)
        return self._out(self._plan.back(xd, filtered=False), g)

    def fbp(self, x):
        """Filtered back projection with the object's filter (ramp by default)."""
        xd, g = self._in(x)
        return self._out(self._plan.back(xd, filtered=True), g)
