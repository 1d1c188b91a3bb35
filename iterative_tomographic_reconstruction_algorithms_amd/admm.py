"""Drop-in for the CG, ADMM-L1 and ADMM-TV solvers of the reference's algorithms/admm.py
(admm.py:245-278 conj_grad, 291-447 lasso_l1, 470-646 lasso_tv, 878-953 admm) on the GPU."""
import numpy as np
import torch

from . import cg as _cg
from . import solvers
from .class_projectors_radon import projectors
from .params import PROJECTOR_OPER
from .solvers import _dev

myfloat = np.float32


def conj_grad(x, b, tp, param):
    """admm.py:245-278: CG on slice 0 only (admm.py:276)."""
    x = np.array(x, dtype=myfloat, copy=True) if not isinstance(x, torch.Tensor) else x
    x0, info = _cg.cg(x[0], b[0], tp, param.n_iter)
    x[0] = x0
    return x, info


def lasso_tv(x, b, tp, param):
    """admm.py:470-646: returns (x, err) with err[k] = ||G x - u|| after outer step k."""
    n1 = 200 if param.n_iter is None else int(param.n_iter)
    xd = _dev(x).clone()
    plan = tp.plan(xd.shape[-1])
    xd, err = solvers.admm_tv(plan, xd, _dev(b), n1, n_inner=int(param.n_iter_pcg),
                              lambd1=float(param.lambd1), lambd2=float(param.lambd2),
                              mu=float(param.mu), pc=float(getattr(param, "pc", -1.0)))
    xd = xd.reshape(x.shape) if isinstance(x, torch.Tensor) else xd.reshape(np.shape(x))
    return (xd if isinstance(x, torch.Tensor) else xd.cpu().numpy()), err


def lasso_l1(x, b, tp, param):
    """admm.py:291-447: returns (x, err) with err[k] = ||x - u|| after outer step k."""
    n1 = 200 if param.n_iter is None else int(param.n_iter)
    xd = _dev(x).clone()
    plan = tp.plan(xd.shape[-1])
    xd, err = solvers.admm_l1(plan, xd, _dev(b), n1, n_inner=int(param.n_iter_pcg),
                              lambd1=float(param.lambd1), lambd2=float(param.lambd2),
                              mu=float(param.mu), pc=float(getattr(param, "pc", -1.0)))
    xd = xd.reshape(x.shape) if isinstance(x, torch.Tensor) else xd.reshape(np.shape(x))
    return (xd if isinstance(x, torch.Tensor) else xd.cpu().numpy()), err


def admm(b, a, param):
    """admm.py:878-953 for the radon projectors, regularisers 'cg', 'lasso' and 'lasso-tv';
    x starts from ones (admm.py:911)."""
    oper = PROJECTOR_OPER.get(param.projector, param.projector)
    n = param.npix_op
    tp = projectors(n, a, oper=oper)
    b = np.asarray(b, dtype=myfloat)
    b = b.reshape(-1, b.shape[-2], b.shape[-1])
    x = np.ones((b.shape[0], n, n), dtype=myfloat)
    if param.reg == "cg":
        return conj_grad(x, b, tp, param)
    if param.reg == "lasso":
        return lasso_l1(x, b, tp, param)
    if param.reg == "lasso-tv":
        return lasso_tv(x, b, tp, param)
    raise NotImplementedError("ADMM regulariser %r is outside the GPU path" % param.reg)
