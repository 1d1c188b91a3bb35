"""Drop-in for the MLEM and SPS functions of the reference's algorithms/sir.py
(sir.py:192-679), running on the GPU.  SPS penalties (huber, tikhonov, haar) are outside
this path and raise NotImplementedError."""
import numpy as np
import torch

from . import solvers
from .class_projectors_radon import projectors
from .params import PROJECTOR_OPER
from .solvers import _dev

myfloat = np.float32


def _out(t, like):
    return t if isinstance(like, torch.Tensor) else t.cpu().numpy()


def forward(tp, x):         # sir.py:192-193
    return tp.A(x)


def backward(tp, b):        # sir.py:198-199
    return tp.At(b)


def em(x, b, tp, param):
    """sir.py:204-324: returns (x, info) with info rows [err1, err2]."""
    n_iter = int(param.n_iter) if param.n_iter is not None and param.n_iter < 1e19 else 200
    eps = -1.0 if param.eps is None else float(param.eps)
    xd = _dev(x).clone()
    plan = tp.plan(xd.shape[-1])
    xd, info, _ = solvers.mlem(plan, xd, _dev(b), n_iter, eps=eps)
    xd = xd.reshape(x.shape) if isinstance(x, torch.Tensor) else xd.reshape(np.shape(x))
    return _out(xd, x), np.array(info)


def sps(x, b, tp, param):
    """sir.py:348-600: returns (x, info) with info rows [err, 0].  param.reg in (None,
    'huber', 'tikhonov', 'haar') with param.reg_cost (beta) and param.huber_cost (delta);
    the iterate is clipped to [0, max x] unless the problem is edge-padded
    (sir.py:539-540)."""
    reg = getattr(param, "reg", None)
    reg = None if reg in (None, "none") else reg
    if reg not in (None, "huber", "tikhonov", "haar"):
        raise ValueError("unknown SPS penalty %r" % (reg,))
    n_iter = int(param.n_iter) if param.n_iter is not None and param.n_iter < 1e19 else 200
    clip = not bool(getattr(param, "edge_padding", False))
    eps = -1.0 if getattr(param, "eps", None) is None else float(param.eps)
    xd = _dev(x).clone()
    plan = tp.plan(xd.shape[-1])
    xd, err = solvers.sps(plan, xd, _dev(b), n_iter, penalty=reg,
                          beta=float(getattr(param, "reg_cost", 0.0) or 0.0),
                          delta=float(getattr(param, "huber_cost", 0.1) or 0.1), clip=clip,
                          eps=eps)
    xd = xd.reshape(x.shape) if isinstance(x, torch.Tensor) else xd.reshape(np.shape(x))
    return _out(xd, x), np.stack([err, np.zeros_like(err)], axis=1)


def sir(b, a, param):
    """sir.py:613-679 for the radon projectors: x starts from ones (sir.py:644)."""
    oper = PROJECTOR_OPER.get(param.projector, param.projector)
    n = param.npix_op
    tp = projectors(n, a, oper=oper)
    b = np.asarray(b, dtype=myfloat)
    b = b.reshape(-1, b.shape[-2], b.shape[-1])
    x = np.ones((b.shape[0], n, n), dtype=myfloat)
    if param.algorithm == "em":
        return em(x, b, tp, param)
    if param.algorithm == "sps":
        return sps(x, b, tp, param)
    raise ValueError("algorithm %r" % param.algorithm)
