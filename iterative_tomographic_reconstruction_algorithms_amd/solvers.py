"""GPU iterations on the projector pair: SIRT, MLEM, SPS, CG and ADMM-TV.

Every vector lives in HBM as a float32 tensor; inner products, norms and step sizes are
FP64 and stay on the device (see include/itomo.h).  The host only reads the monitor
scalars, in one transfer per call unless ``check_every`` asks for early stopping.

Reference files: algorithms/art.py:211-365 (SIRT), algorithms/sir.py:204-600 (MLEM,
SPS), algorithms/cg.py:83-310 (CG, CG for ADMM-TV), algorithms/admm.py:245-646.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from .core import Plan, _p, device, stream_handle

__all__ = ["sirt", "mlem", "sps", "cg", "admm_tv", "grad", "grad_adjoint", "VecWork"]


def _dev(a):
    if isinstance(a, torch.Tensor):
        return a.to(device=device(), dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device())


def _as3(t):
    return t if t.dim() == 3 else t.unsqueeze(0)


class VecWork:
    """FP64 scalar block and scratch for the CG/ADMM vector kernels."""

    def __init__(self, nslices):
        dev = device()
        self.nslices = nslices
        self.scal = torch.zeros(4 * nslices, dtype=torch.float64, device=dev)
        self.work = torch.zeros(L.VEC_WORK_PER_SLICE * nslices, dtype=torch.float64, device=dev)


# ---- SIRT -----------------------------------------------------------------------------

def _mask_u8(m, cond, side, dev):
    a = np.asarray(m.cpu().numpy() if isinstance(m, torch.Tensor) else m)
    if a.shape[-2:] != (side, side):
        raise ValueError("mask of shape %s for a %dx%d reconstruction region"
                         % (a.shape, side, side))
    return torch.from_numpy(np.ascontiguousarray(cond(a).astype(np.uint8))).to(dev)


class Constraints:
    """The SIRT constraints of art.py:296-321 applied after each step on the GPU
    (itomo_sirt_constraints): support mask, additional masks, pc (non-negativity,
    resolution circle), on the region [i1, i2)^2."""

    def __init__(self, npix, nslices, pc=-1.0, mask=None, mask_add=None, region=None):
        dev = device()
        self.npix, self.nslices = npix, nslices
        self.i1, self.i2 = (0, npix) if region is None else (int(region[0]), int(region[1]))
        side = self.i2 - self.i1
        self.pc = float(-1.0 if pc is None else pc)
        # supp_constr zeroes where mask == 0, mask_constr averages where mask == 1
        self.mask = None if mask is None else _mask_u8(mask, lambda a: a != 0, side, dev)
        if mask_add is None:
            self.mask_add, self.nmask = None, 0
        else:
            ma = np.asarray(mask_add)
            ma = ma.reshape((-1,) + ma.shape[-2:])
            self.mask_add = _mask_u8(ma, lambda a: a == 1, side, dev)
            self.nmask = ma.shape[0]
        self.work = torch.zeros(max(1, 3 * nslices * ((npix * npix + 2047) // 2048)),
                                dtype=torch.float64, device=dev)

    @property
    def needed(self):
        """Whether the step's fused non-negativity does not cover them."""
        return (self.mask is not None or self.nmask > 0 or self.pc in (1.0, 2.0)
                or self.i1 != 0 or self.i2 != self.npix)

    def apply(self, x, x_old=None, stats=None):
        L.check(L.lib().itomo_sirt_constraints(
            _p(x), self.npix, x.numel() // (self.npix * self.npix), self.i1, self.i2,
            _p(self.mask), _p(self.mask_add), self.nmask, self.pc, _p(x_old), _p(stats),
            _p(self.work), stream_handle()), "sirt_constraints")


def sirt(plan, x, b, n_iter, pc=-1.0, check_every=0, mask=None, mask_add=None, region=None,
         start=0, on_iterate=None):
    """art.py:217-365 on the GPU: R, C weights, then n_iter fused steps
    x <- x + C*A^T(R*(b - A x)) with phys_constr(x, pc) when pc >= 0.  With a support
    ``mask``, additional masks ``mask_add``, pc in {1, 2} (the resolution circle) or a
    reconstruction ``region`` (i1, i2) smaller than the image (edge padding), the
    constraints of art.py:303-321 run after each step instead, in the reference's order.

    Returns (x, info) with info rows [diff, obj] as in art.py:325-345:
    diff = ||x_{k+1} - x_k|| / ||x_k|| over all slices, obj = ||A x_{k+1} - b||^2 of the
    middle slice.  Stops early on a NaN, like art.py:348-350, when check_every > 0.
    ``start`` numbers the first step (a run resumed from iterate ``start``);
    ``on_iterate(k, x)`` is called after step k (k counted from 1 + start), e.g. to write
    the iterate as art.py:328-341 does with -cit.
    """
    x = _as3(x)
    b = _as3(b)
    nz = x.shape[0]
    nzz = 0 if nz == 1 else int(0.5 * nz)
    R, C = plan.sirt_weights()
    work = torch.empty_like(b)
    hist = torch.zeros((n_iter + 1, 3 * nz), dtype=torch.float64, device=x.device)
    cons = Constraints(plan.npix, nz, pc, mask, mask_add, region)
    constrained = cons.needed
    x_old = torch.empty_like(x) if constrained else None
    done = n_iter
    for it in range(n_iter):
        if constrained:
            x_old.copy_(x)
            plan.sirt_iterate(x, b, R, C, work, pc_beta=-1.0, stats=hist[it])
            cons.apply(x, x_old, hist[it])
        else:
            plan.sirt_iterate(x, b, R, C, work, pc_beta=pc, stats=hist[it])
        if on_iterate is not None:
            on_iterate(start + it + 1, x)
        if check_every and (it + 1) % check_every == 0:
            if not torch.isfinite(hist[it]).all():
                done = it + 1
                break
    plan.sirt_forward(x, b, R, work, stats=hist[done])   # residual of the last iterate
    h = hist[: done + 1].cpu().numpy().reshape(done + 1, nz, 3)
    info = []
    for k in range(done):
        dx2 = h[k, :, 1].sum()
        x2 = h[k, :, 2].sum()
        diff = math.sqrt(dx2) / math.sqrt(x2) if x2 > 0 else float("inf")
        obj = h[k + 1, nzz, 0]
        info.append([diff, obj])
        if math.isnan(diff) or math.isnan(obj):
            break
    return x, info


# ---- MLEM -----------------------------------------------------------------------------

def mlem(plan, x, b, n_iter, eps=-1.0):
    """sir.py:204-324 on the GPU.  Returns (x, info, err) where info rows are
    [err1, err2] = [||x_{k+1}-x_k|| (middle slice), 0.5*||A x_{k+1} - b|| (middle slice)]
    and err[k] = sum (A x_k - b)^2 / sum b^2 of the middle slice (sir.py:287)."""
    x = _as3(x)
    b = _as3(b)
    nz = x.shape[0]
    nzz = 0 if nz == 1 else int(nz * 0.5)
    inv_d = plan.mlem_weights()
    work = torch.empty_like(b)
    bsq = float(torch.sum(b[nzz].double() ** 2))
    hist = torch.zeros((n_iter + 1, 3 * nz), dtype=torch.float64, device=x.device)
    for it in range(n_iter):
        plan.mlem_iterate(x, b, inv_d, work, stats=hist[it])
        # sir.py:256,287: `while it < n_iter and err > eps` ends the loop right after the
        # first step whose err (of the iterate entering it) is <= eps
        if eps >= 0 and not float(hist[it, 3 * nzz]) / bsq > eps:
            n_iter = it + 1
            break
    # residual of the final iterate, without changing x: forward epilogue only
    fin = torch.zeros(3 * nz, dtype=torch.float64, device=x.device)
    plan.sirt_forward(x, b, torch.ones_like(b[0]), work, stats=fin)
    hist[n_iter] = fin
    h = hist[: n_iter + 1].cpu().numpy().reshape(n_iter + 1, nz, 3)
    err = h[:n_iter, nzz, 0] / bsq
    info = [[math.sqrt(h[k, nzz, 1]), 0.5 * math.sqrt(h[k + 1, nzz, 0])] for k in range(n_iter)]
    return x, info, err


# ---- SPS ------------------------------------------------------------------------------

def sps(plan, x, b, n_iter, penalty=None, beta=0.0, delta=0.1, clip=True, eps=-1.0):
    """sir.py:348-560: separable paraboloidal surrogate with w = exp(-b/max b),
    den = A^T(w * A(1)) and the FP64 quotient in the update
    x <- x - (A^T(w (A x - b)) + beta R') / (den + beta R'' + eps).  ``penalty`` is None,
    "huber" (on the Sobel magnitude of x, threshold ``delta``), "tikhonov" or "haar"
    (sir.py:455-530); ``clip`` applies np.clip(x, 0, max x) after each step, the
    reference's default when there is no edge padding (sir.py:539-540).
    Returns (x, err) with err[k] = sum ((A x_k - b) w)^2 over the middle slice / sum b^2
    over all slices (sir.py:366, 546).  With ``eps`` >= 0 the loop stops right after the
    first step whose err is <= eps (sir.py:426)."""
    x = _as3(x)
    b = _as3(b)
    nz = x.shape[0]
    w = torch.empty_like(b)
    den = torch.empty_like(x)
    work = torch.empty_like(b)
    bmax = torch.empty(1024, dtype=torch.float32, device=x.device)
    L.check(L.lib().itomo_sps_weights(plan._h, _p(b), nz, _p(w), _p(den), _p(work), _p(bmax),
                                      stream_handle()), "sps_weights")
    kind = L.PENALTY[penalty]
    img_work = torch.empty(((4 if kind else 1) * nz,) + tuple(x.shape[-2:]),
                           dtype=torch.float32, device=x.device)
    nzz = 0 if nz == 1 else int(nz * 0.5)
    hist = torch.zeros((n_iter, 3 * nz), dtype=torch.float64, device=x.device)
    bden = float(torch.sum(b.double() ** 2))
    for it in range(n_iter):
        L.check(L.lib().itomo_sps_step(plan._h, _p(x), _p(b), _p(w), _p(den), _p(work),
                                       _p(img_work), nz, kind, float(beta), float(delta),
                                       int(bool(clip)), _p(hist[it]), stream_handle()),
                "sps_step")
        if eps >= 0 and not float(hist[it, 3 * nzz]) / bden > eps:
            n_iter = it + 1
            break
    h = hist[:n_iter].cpu().numpy().reshape(n_iter, nz, 3)
    return x, h[:, nzz, 0] / bden


def penalty(kind, image, delta=0.1, arith="fast"):
    """(R_num, R_den) of the reference's penalty module (penalty.pyx:16-106) for an
    (..., nr, nc) float32 GPU tensor."""
    img = image.contiguous()
    nr, nc = img.shape[-2:]
    ns = img.numel() // (nr * nc)
    rn = torch.empty_like(img)
    rd = torch.empty_like(img)
    L.check(L.lib().itomo_penalty(_p(img), nr, nc, ns, L.PENALTY[kind], float(delta),
                                  L.ARITH[arith], _p(rn), _p(rd), stream_handle()), "penalty")
    return rn, rd


def sobel_magnitude(x):
    """sir.py:339-343 on an (..., nr, nc) float32 GPU tensor."""
    x = x.contiguous()
    nr, nc = x.shape[-2:]
    g = torch.empty_like(x)
    L.check(L.lib().itomo_sobel_magnitude(_p(x), nr, nc, x.numel() // (nr * nc), _p(g),
                                          stream_handle()), "sobel_magnitude")
    return g


# ---- finite differences -----------------------------------------------------------------

def grad(x):
    """np.gradient of each slice: (..., N, N) -> (..., 2, N, N)."""
    x3 = _as3(x)
    n = x3.shape[-1]
    g = torch.empty((x3.shape[0], 2, n, n), dtype=torch.float32, device=x3.device)
    L.check(L.lib().itomo_grad(_p(x3), n, x3.shape[0], _p(g), stream_handle()), "grad")
    return g if x.dim() == 3 else g[0]


def grad_adjoint(g):
    g4 = g if g.dim() == 4 else g.unsqueeze(0)
    n = g4.shape[-1]
    out = torch.empty((g4.shape[0], n, n), dtype=torch.float32, device=g4.device)
    L.check(L.lib().itomo_grad_adjoint(_p(g4), n, g4.shape[0], _p(out), stream_handle()),
            "grad_adjoint")
    return out if g.dim() == 4 else out[0]


# ---- CG -------------------------------------------------------------------------------

def _cg_loop(plan, x, r, p, n_iter, vw, lambd1=0.0, mu=0.0, tv=False, with_eps=True,
             record=None, b=None):
    """CG iterations on H = A^T A (+ lambd1 I + mu G^T G) given r, p (cg.py:120-160)."""
    nz, n = x.shape[0], x.shape[-1]
    n2 = n * n
    work = torch.empty((nz, plan.nang, n), dtype=torch.float32, device=x.device)
    Hp = torch.empty_like(x)
    g = torch.empty((nz, 2, n, n), dtype=torch.float32, device=x.device) if tv else None
    lib = L.lib()
    st = stream_handle()
    for it in range(n_iter):
        plan.forward(p, out=work)
        plan.back(work, out=Hp)
        if tv:
            L.check(lib.itomo_grad(_p(p), n, nz, _p(g), st), "grad")
        L.check(lib.itomo_cg_hp(_p(Hp), _p(p), _p(g), n, nz, float(lambd1), float(mu),
                                _p(vw.scal), _p(vw.work), st), "cg_hp")
        L.check(lib.itomo_cg_update(_p(x), _p(r), _p(p), _p(Hp), n2, nz, _p(vw.scal),
                                    _p(vw.work), st), "cg_update")
        if record is not None:
            record(it)
        L.check(lib.itomo_cg_direction(_p(p), _p(r), n2, nz, int(with_eps), _p(vw.scal), st),
                "cg_direction")
    return x


def cg(plan, x, b, n_iter):
    """cg.py:83-176: CG on A^T A x = A^T b from x, r0 = A^T b - A^T A x.  Returns (x, info)
    with info rows [||r||, ||x_{k+1} - x_k||, ||A x_{k+1} - b||^2] (cg.py:171-174)."""
    x = _as3(x)
    b = _as3(b)
    nz, n = x.shape[0], x.shape[-1]
    vw = VecWork(nz)
    atb = plan.back(b)
    atax = plan.back(plan.forward(x))
    r = atb - atax                                     # float32, like cg.py:85
    p = r.clone()
    L.check(L.lib().itomo_cg_dot(_p(r), _p(r), n * n, nz, 0, _p(vw.scal), _p(vw.work),
                                 stream_handle()), "cg_dot")
    hist = torch.zeros((n_iter, 4 * nz), dtype=torch.float64, device=x.device)
    x_prev = x.clone()
    diffs = torch.zeros(n_iter, dtype=torch.float64, device=x.device)
    objs = torch.zeros(n_iter, dtype=torch.float64, device=x.device)
    work = torch.empty_like(b)

    def record_full(it):
        hist[it] = vw.scal
        diffs[it] = torch.linalg.vector_norm((x - x_prev).double())
        x_prev.copy_(x)
        objs[it] = torch.linalg.vector_norm((plan.forward(x, out=work) - b).double()) ** 2

    _cg_loop(plan, x, r, p, n_iter, vw, record=record_full)
    h = hist.cpu().numpy().reshape(n_iter, nz, 4)
    info = [[math.sqrt(h[k, 0, 2]), float(diffs[k]), float(objs[k])] for k in range(n_iter)]
    return x, info


# ---- ADMM-TV --------------------------------------------------------------------------

def admm_tv(plan, x, b, n_iter, n_inner=4, lambd1=1.0, lambd2=1.0, mu=1.0, pc=-1.0,
            return_state=False):
    """admm.py:470-646 (lasso_tv): per outer step an inner CG of n_inner iterations on
    (A^T A + lambd1 I + mu G^T G) x = A^T b + G^T(mu u - alpha) (cg.py:256-310), then
    u = soft(G x + alpha/mu, lambd2/mu), alpha += G x - u.  Returns (x, err) with
    err[k] = ||G x - u|| over all slices (admm.py:597), and (x, err, u, alpha), each of
    u and alpha (nz, 2, N, N), when ``return_state``."""
    x = _as3(x)
    b = _as3(b)
    nz, n = x.shape[0], x.shape[-1]
    u = torch.zeros((nz, 2, n, n), dtype=torch.float32, device=x.device)
    al = torch.zeros_like(u)
    v = torch.empty_like(u)
    gx = torch.empty_like(u)
    vw = VecWork(nz)
    lib = L.lib()
    st = stream_handle()
    atb = plan.back(b)
    work = torch.empty_like(b)
    atax = torch.empty_like(x)
    r = torch.empty_like(x)
    errs = torch.zeros(n_iter, dtype=torch.float64, device=x.device)
    for it in range(n_iter):
        L.check(lib.itomo_tv_rhs(_p(v), _p(u), _p(al), v.numel(), float(mu), st), "tv_rhs")
        plan.forward(x, out=work)
        plan.back(work, out=atax)
        L.check(lib.itomo_grad(_p(x), n, nz, _p(gx), st), "grad")
        # r = A^T b + G^T v - (A^T A x + lambd1 x + G^T G x), scal[0] = <r, r>
        L.check(lib.itomo_tv_residual(_p(r), _p(atb), _p(atax), _p(x), _p(gx), _p(v), n, nz,
                                      float(lambd1), _p(vw.scal), _p(vw.work), st),
                "tv_residual")
        p = r.clone()
        _cg_loop(plan, x, r, p, n_inner, vw, lambd1=lambd1, mu=mu, tv=True, with_eps=False)
        if pc >= 0.0:
            x[x < 0] *= pc                                   # admm.py:532-541
        L.check(lib.itomo_admm_tv_update(_p(x), _p(u), _p(al), n, nz, float(mu),
                                         float(lambd2), _p(vw.scal), _p(vw.work), st),
                "admm_tv_update")
        errs[it] = vw.scal.view(nz, 4)[:, 3].sum().sqrt()
    if return_state:
        return x, errs.cpu().numpy(), u, al
    return x, errs.cpu().numpy()


def admm_l1(plan, x, b, n_iter, n_inner=4, lambd1=1.0, lambd2=1.0, mu=1.0, pc=-1.0):
    """admm.py:291-447 (lasso_l1): per outer step an inner CG of n_inner iterations on
    (A^T A + (lambd1 + mu) I) x = A^T b + mu u - alpha (cg.py:189-245), then
    u = soft(x + alpha/mu, lambd2/mu), alpha += x - u.  Returns (x, err) with
    err[k] = ||x - u|| over all slices (admm.py:390)."""
    x = _as3(x)
    b = _as3(b)
    nz, n = x.shape[0], x.shape[-1]
    u = torch.zeros_like(x)
    al = torch.zeros_like(x)
    v = torch.empty_like(x)
    vw = VecWork(nz)
    lib = L.lib()
    st = stream_handle()
    c = float(np.float32(lambd1) + np.float32(mu))        # float32 sum, as in cg.py:191
    atb = plan.back(b)
    work = torch.empty_like(b)
    atax = torch.empty_like(x)
    r = torch.empty_like(x)
    errs = torch.zeros(n_iter, dtype=torch.float64, device=x.device)
    for it in range(n_iter):
        L.check(lib.itomo_tv_rhs(_p(v), _p(u), _p(al), v.numel(), float(mu), st), "tv_rhs")
        plan.forward(x, out=work)
        plan.back(work, out=atax)
        L.check(lib.itomo_l1_residual(_p(r), _p(atb), _p(atax), _p(x), _p(v), n, nz, c,
                                      _p(vw.scal), _p(vw.work), st), "l1_residual")
        p = r.clone()
        _cg_loop(plan, x, r, p, n_inner, vw, lambd1=c, with_eps=False)  # cg.py:236
        if pc >= 0.0:
            x[x < 0] *= pc                                   # admm.py:364-374
        L.check(lib.itomo_admm_l1_update(_p(x), _p(u), _p(al), n, nz, float(mu), float(lambd2),
                                         _p(vw.scal), _p(vw.work), st), "admm_l1_update")
        errs[it] = vw.scal.view(nz, 4)[:, 3].sum().sqrt()
    return x, errs.cpu().numpy()
