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
    from . import cli
    n = param.npix_op
    tp = cli.make_projector(param.projector, n, a)
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


# ---- command line (admm.py:104-212, 1109-) ------------------------------------------------

def getArgs(argv=None):
    import argparse
    import sys
    from . import cli
    parser = argparse.ArgumentParser(description="ADMM / CG reconstruction",
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    cli.add_common_args(parser, "grid-kb")
    parser.add_argument("-o", "--reco", dest="reco", help="Output name")
    parser.add_argument("-n1", dest="n_iter", type=int, help="Outer iterations")
    parser.add_argument("-n2", dest="n_iter_pcg", type=int, default=4,
                        help="Inner CG iterations")
    parser.add_argument("-pp", dest="param", help="lambd1:lambd2:mu")
    parser.add_argument("-r", dest="reg", default="lasso", help="cg, lasso or lasso-tv")
    parser.add_argument("-pc", dest="pc", type=myfloat, default=-1.0,
                        help="phys. constraint on negative values")
    args = parser.parse_args(argv)
    if args.n_iter is None and args.eps is None:
        parser.print_help()
        sys.exit("\nERROR: Stopping criterion not specified!\n")
    return args


def main(argv=None):
    """admm.py main for the regularisers cg, lasso and lasso-tv, written as
    <input>_cgls_<proj>.<ext> or <input>_admm_<reg>_lambd1_..._mu_..._<proj>.<ext>
    (class_admm_param.py:96-136)."""
    import time
    from . import cli
    from .params import param as make_param
    args = getArgs(argv)
    d = cli.prepare(args)
    nz = len(d["names"])
    if args.reg == "cg":
        l1 = l2 = mu = myfloat(1.0)
        root = "_cgls"
    else:
        pp = args.param if args.param is not None else "1.0:1.0:1.0"
        l1, l2, mu = (myfloat(v) for v in pp.split(":")[:3])
        root = "_admm" + {"lasso": "_lasso", "lasso-tv": "_lassotv"}.get(args.reg, "")
        root += "_lambd1_" + str(l1) + "_lambd2_" + str(l2) + "_mu_" + str(mu)
    root += cli.projector_suffix(args.projector)
    p = make_param(npix_op=d["npix"], nang=d["nang"], nz=nz, n_iter=args.n_iter,
                   n_iter_pcg=args.n_iter_pcg, reg=args.reg, lambd1=float(l1),
                   lambd2=float(l2), mu=float(mu), pc=float(args.pc),
                   projector=args.projector)
    t0 = time.time()
    reco, info = admm(d["sinos"], d["angles"], p)
    t1 = time.time()
    outs = []
    for i in range(nz):
        r = reco[i, d["i1"]:d["i2"], d["i1"]:d["i2"]] if d["edf"] else reco[i]
        outs.append(cli.save_reco(d["pathout"], d["names"][i], args.reco if nz == 1 else None,
                                  root, r))
    print("\nTime elapsed for the reconstruction: ", (t1 - t0) / 60.0)
    return outs, info


if __name__ == "__main__":
    main()
