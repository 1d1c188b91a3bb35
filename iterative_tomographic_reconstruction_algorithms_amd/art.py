"""Drop-in for the SIRT functions of the reference's algorithms/art.py (art.py:211-435),
running on the GPU.  Arrays may be numpy (returned as numpy) or float32 GPU tensors.

Supported options: n_iter, pc (phys_constr, art.py:314-321).  TV, masks and the
resolution-circle constraint are outside this path and raise NotImplementedError.
"""
import numpy as np
import torch

from . import solvers
from .class_projectors_radon import projectors
from .params import PROJECTOR_OPER
from .solvers import _dev

myfloat = np.float32
epsc = 1e-10     # art.py:77


def _out(t, like):
    return t if isinstance(like, torch.Tensor) else t.cpu().numpy()


def sirt_step(x, b, tp, R, C):
    """art.py:211-213: x <- x + C * At(R * (b - A x)), one slice, fused on the GPU."""
    plan = tp.plan(x.shape[-1])
    xd, bd = _dev(x), _dev(b)
    Rd, Cd = _dev(R), _dev(C)
    work = torch.empty_like(bd)
    plan.sirt_iterate(xd, bd, Rd, Cd, work)
    if isinstance(x, torch.Tensor):
        if xd.data_ptr() != x.data_ptr():
            x.copy_(xd)
        return x
    x[:] = xd.cpu().numpy()
    return x


def _pc_beta(param):
    pc = getattr(param, "pc", -1.0)
    if pc is None or pc < 0.0:
        return -1.0
    if pc in (1.0, 2.0):
        raise NotImplementedError("pc=%g needs the resolution-circle constraint "
                                  "(utils.resol_circle_constr), outside the GPU path" % pc)
    return float(pc)


def sirt(x, b, tp, param):
    """art.py:217-365: returns (x, info) with info rows [diff, obj]."""
    for opt in ("mask", "mask_add"):
        if getattr(param, opt, None) is not None:
            raise NotImplementedError("SIRT option %r is outside the GPU path" % opt)
    if getattr(param, "tv", 0.0):
        raise NotImplementedError("SIRT with TV denoising is outside the GPU path")
    n_it = 200 if param.n_iter is None else int(param.n_iter)
    xd = _dev(x).clone()
    bd = _dev(b)
    plan = tp.plan(xd.shape[-1])
    xd, info = solvers.sirt(plan, xd, bd, n_it, pc=_pc_beta(param), check_every=10)
    xd = xd.reshape(np.shape(x)) if not isinstance(x, torch.Tensor) else xd.reshape(x.shape)
    return _out(xd, x), info


def alg_rec(b, a, param):
    """art.py:378-435 for the radon projectors: x starts from ones (art.py:408)."""
    oper = PROJECTOR_OPER.get(param.projector, param.projector)
    n = param.npix_op
    tp = projectors(n, a, oper=oper)
    b = np.asarray(b, dtype=myfloat) if not isinstance(b, torch.Tensor) else b
    nz = b.shape[0] if b.ndim == 3 else 1
    x = np.ones((nz, n, n), dtype=myfloat)
    return sirt(x, b.reshape(nz, -1, n) if not isinstance(b, torch.Tensor) else b, tp, param)
