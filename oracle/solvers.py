"""numpy restatements of the iterations built on the projector pair (test infrastructure).

Each takes the projector pair as callables ``A(x)``, ``At(y)`` and a dtype: float32
reproduces the reference's float32 numpy arithmetic step by step, float64 is the FP64
oracle.

  sirt : algorithms/art.py:211-237, 278-326 (x <- x + C*At(R*(b - A x)), phys_constr)
  mlem : algorithms/sir.py:204-290
"""
import numpy as np


def sirt_weights(A, At, n, nang, dtype=np.float32):
    R = A(np.ones((n, n), dtype=dtype)).astype(dtype)
    R[R == 0] = R[R == 0] + dtype(1e-10)                     # art.py:229
    R = (dtype(1.0) / R).astype(dtype)
    C = At(np.ones((nang, n), dtype=dtype)).astype(dtype)
    C[C == 0] = C[C == 0] + dtype(1e-10)                     # art.py:236
    C = (dtype(1.0) / C).astype(dtype)
    return R, C


def resol_circle(img):
    """utils.py:229-251 on a 2-D image (a copy)."""
    m, n = img.shape
    ctr_x, ctr_y = m * 0.5 - 0.5, n * 0.5 - 0.5
    d = np.min((ctr_x, ctr_y))
    xx, yy = np.meshgrid(np.arange(m), np.arange(n))
    i = np.argwhere(np.sqrt((xx - ctr_x) ** 2 + (yy - ctr_y) ** 2) > d)
    out = img.copy()
    out[i[:, 0], i[:, 1]] = 0
    return out


def constrain(x, pc=-1.0, mask=None, mask_add=None, region=None, dtype=np.float32):
    """art.py:303-321 after a SIRT step on one slice: support mask (utils.py:264-266),
    additional masks (utils.py:279-281, the float32 numpy mean), then pc (1: x<0 -> 0 and
    the resolution circle, 2: the circle, other >= 0: x<0 *= pc) on [i1, i2)^2."""
    x = x.copy()
    i1, i2 = (0, x.shape[-1]) if region is None else region
    sub = x[i1:i2, i1:i2]
    if mask is not None:
        sub[np.asarray(mask) == 0] = 0.0
    if mask_add is not None:
        for m in np.asarray(mask_add).reshape((-1,) + sub.shape):
            sub[m == 1] = np.mean(sub[m == 1])
    if pc is not None and pc >= 0.0:
        if pc == 1.0:
            neg = sub < 0
            sub[neg] = sub[neg] * dtype(0.0)
            sub[:] = resol_circle(sub)
        if pc == 2.0:
            sub[:] = resol_circle(sub)
        else:
            neg = sub < 0
            sub[neg] = sub[neg] * dtype(pc)                               # utils.py:171-176
    x[i1:i2, i1:i2] = sub
    return x


def sirt(x0, b, A, At, n_iter, pc=-1.0, dtype=np.float32, mask=None, mask_add=None,
         region=None):
    """Returns (x, info, resid) where info rows are [diff, obj] as in art.py:325-345 and
    resid[k] = ||b - A x_k||_2 for the iterate entering step k (in float64).  The
    constraints of art.py:303-321 follow each step (``constrain``)."""
    n = x0.shape[-1]
    nang = b.shape[-2]
    R, C = sirt_weights(A, At, n, nang, dtype)
    x = x0.astype(dtype).copy()
    b = b.astype(dtype)
    info, resid = [], []
    for _ in range(n_iter):
        x_old = x.copy()
        ax = A(x).astype(dtype)
        d = (b - ax).astype(dtype)
        resid.append(np.linalg.norm(d.astype(np.float64)))
        x = (x + C * At((R * d).astype(dtype)).astype(dtype)).astype(dtype)   # art.py:212
        x = constrain(x, pc, mask, mask_add, region, dtype)
        diff = np.linalg.norm((x - x_old).astype(np.float64)) / np.linalg.norm(
            x_old.astype(np.float64))
        obj = np.linalg.norm((A(x) - b).astype(np.float64)) ** 2
        info.append([diff, obj])
    return x, np.array(info), np.array(resid)


def mlem(x0, b, A, At, n_iter, dtype=np.float32, eps=None):
    """Returns (x, err) with err[k] = sum (A x_k - b)^2 / sum b^2 (sir.py:287),
    evaluated in float64 from the float32 arrays.  With ``eps`` the loop is the
    reference's `while it < n_iter and err > eps` (sir.py:251): it ends after the first
    step whose err is <= eps."""
    n = x0.shape[-1]
    nang = b.shape[-2]
    d = At(np.ones((nang, n), dtype=dtype)).astype(dtype)
    d[d == 0] = d[d == 0] + dtype(1e-10)                     # sir.py:216-217
    inv_d = (dtype(1.0) / d).astype(dtype)
    x = x0.astype(dtype).copy()
    b = b.astype(dtype)
    err = []
    for _ in range(n_iter):
        bn = A(x).astype(dtype)
        z = bn == 0
        bn[z] = bn[z] + dtype(np.spacing(1))                  # sir.py:270
        q = (b / bn).astype(dtype)
        err.append(np.sum((bn.astype(np.float64) - b) ** 2) / np.sum(b.astype(np.float64) ** 2))
        xa = At(q).astype(dtype)
        x = (x * (inv_d * xa).astype(dtype)).astype(dtype)   # sir.py:283
        if eps is not None and not err[-1] > eps:
            break
    return x, np.array(err)


def sps(x0, b, A, At, n_iter, dtype=np.float32, penalty=None, beta=0.0, delta=0.1,
        clip=True, eps=None):
    """sir.py:348-560: w = exp(-b/max b); den = At(w*A(1));
    x <- x - (At(w*(A x - b)) + beta R') / (den + beta R'' + spacing(1)), the quotient
    evaluated in float64 (the GPU path's FP64 denominator), R from the restated penalty
    module (oracle/penalty_oracle.c, the reference build's arithmetic) on x, or on the
    scipy Sobel magnitude of x for "huber"; then np.clip(x, 0, max x) when ``clip``
    (sir.py:539-540).  Returns (x, err) with err[k] = sum(w(Ax_k-b))^2/sum b^2 (one
    slice).  With ``eps`` the loop is sir.py:426's `while it < n_iter and err > eps`."""
    n = x0.shape[-1]
    b = b.astype(dtype)
    # the reference's float32 np.exp is not correctly rounded; like the GPU path, take
    # exp in double and round once (differs from numpy's by at most 1 ulp in w)
    w = np.exp((-b / np.max(b)).astype(dtype).astype(np.float64)).astype(dtype)
    den = At((w * A(np.ones((n, n), dtype=dtype))).astype(dtype)).astype(dtype)
    x = x0.astype(dtype).copy()
    bden = np.sum(b.astype(np.float64) ** 2)
    err = []
    for _ in range(n_iter):
        if penalty is not None:
            from . import penalty as P
            src = P.sobel_magnitude(x) if penalty == "huber" else x
            rn, rd = P.penalty(penalty, src, delta)
        bd = ((A(x).astype(dtype) - b) * w).astype(dtype)
        err.append(np.sum(bd.astype(np.float64) ** 2) / bden)
        num = At(bd).astype(dtype)
        p = num.astype(np.float64)
        q = den.astype(np.float64)
        if penalty is not None:     # beta is float32 in the reference (sir.py:128)
            p = p + np.float64(np.float32(beta)) * rn.astype(np.float64)
            q = q + np.float64(np.float32(beta)) * rd.astype(np.float64)
        q = q + np.spacing(1)
        x = (x.astype(np.float64) - p / q).astype(dtype)
        if clip:
            x = np.clip(x, 0, np.max(x))
        if eps is not None and not err[-1] > eps:
            break
    return x, np.array(err)


def gradient(m):
    """cg.py:62-63 (np.gradient, a library routine)."""
    return np.gradient(m)


def gradient_adjoint(g):
    """cg.py:66-71 as written: dot(g1, Gx) + dot(Gy, g0) with dense difference matrices."""
    nx, ny = g[0].shape
    Gy = np.gradient(np.diag(np.ones(nx)))[1]
    Gx = np.gradient(np.diag(np.ones(ny)))[0]
    return np.dot(g[1], Gx) + np.dot(Gy, g[0])


def cg(x0, b, A, At, n_iter, dtype=np.float32, lambd1=0.0, mu=0.0, v=None, tv=False,
       with_eps=True, l1=False):
    """cg.py:83-160 (plain), cg.py:256-306 (tv=True) and cg.py:189-245 (l1=True): CG on
    (At A + lambd1 I + mu Gt G) x = At b (+ Gt v), or (At A + (lambd1 + mu) I) x = At b + v, vectors in `dtype`, inner products and
    step sizes in float64 (the GPU path's FP64 scalars; numpy-1 rounds alpha and beta to
    the vector dtype before scaling).  Returns (x, rr) with rr[k] = <r_{k+1}, r_{k+1}>."""
    dt = dtype
    x = x0.astype(dt).copy()

    c = dt(dt(lambd1) + dt(mu))        # cg.py:191,211: (lambd1 + mu), float32 params

    def H(p):
        h = At(A(p).astype(dt)).astype(dt)
        if l1:
            h = (h + c * p).astype(dt)
        if tv:
            gp = [g.astype(dt) for g in gradient(p)]
            h = (h + dt(lambd1) * p).astype(dt)
            h = (h + dt(mu) * gradient_adjoint(gp).astype(dt)).astype(dt)
        return h

    if tv:
        gx = [g.astype(dt) for g in gradient(x)]
        lhs = (At(b).astype(dt) + gradient_adjoint(v).astype(dt)).astype(dt)
        rhs = ((At(A(x).astype(dt)).astype(dt) + dt(lambd1) * x).astype(dt)
               + gradient_adjoint(gx).astype(dt)).astype(dt)
        r = (lhs - rhs).astype(dt)
    elif l1:
        lhs = (At(b).astype(dt) + v.astype(dt)).astype(dt)
        rhs = (At(A(x).astype(dt)).astype(dt) + (c * x).astype(dt)).astype(dt)
        r = (lhs - rhs).astype(dt)
    else:
        r = (At(b).astype(dt) - At(A(x).astype(dt)).astype(dt)).astype(dt)
    p = r.copy()
    rr = float(np.sum(r.astype(np.float64) ** 2))
    out = []
    eps = np.spacing(1)
    for _ in range(n_iter):
        hp = H(p)
        php = float(np.sum(p.astype(np.float64) * hp.astype(np.float64)))
        a = dt(rr / (php + eps))
        x = (x + (a * p).astype(dt)).astype(dt)
        r = (r - (a * hp).astype(dt)).astype(dt)
        rn = float(np.sum(r.astype(np.float64) ** 2))
        out.append(rn)
        # cg.py:126-142, 211-231, 274-292: the loops stop once ||r|| <= 1e-12 (only an
        # exactly solved system gets there; float32 vectors never do)
        if np.sqrt(rn) <= CG_STOP:
            break
        bta = dt(rn / (rr + (eps if with_eps else 0.0)))
        p = (r + (bta * p).astype(dt)).astype(dt)
        rr = rn
    return x, np.array(out)


CG_STOP = 1e-12     # cg.py:101, 203, 271: eps of the inner CG loops


def admm_l1(x0, b, A, At, n_iter, n_inner=4, lambd1=1.0, lambd2=1.0, mu=1.0,
            dtype=np.float32):
    """admm.py:291-447 for one slice: inner CG (cg.py:189-245, no eps in beta) then
    u = soft(x + alpha/mu, lambd2/mu), alpha += x - u.  Returns (x, err) with
    err[k] = ||x - u|| after outer step k."""
    dt = dtype
    x = x0.astype(dt).copy()
    u = np.zeros_like(x)
    al = np.zeros_like(x)
    err = []
    for _ in range(n_iter):
        v = (dt(mu) * u - al).astype(dt)
        x, _ = cg(x, b, A, At, n_inner, dtype=dt, lambd1=lambd1, mu=mu, v=v, l1=True,
                  with_eps=False)
        u = soft((x + (al / dt(mu)).astype(dt)).astype(dt), dt(lambd2) / dt(mu)).astype(dt)
        r = (x - u).astype(dt)
        al = (al + r).astype(dt)
        err.append(np.sqrt(np.sum(r.astype(np.float64) ** 2)))
    return x, np.array(err)


def soft(y, t):
    """utils.py:214-216 (hard_thres is a soft threshold)."""
    return np.maximum(0, y - t) - np.maximum(0, -y - t)


def admm_tv(x0, b, A, At, n_iter, n_inner=4, lambd1=1.0, lambd2=1.0, mu=1.0, dtype=np.float32,
            return_state=False):
    """admm.py:470-604 for one slice: inner CG (cg.py:256-310) on
    (At A + lambd1 I + mu Gt G) x = At b + Gt v with v = mu u - alpha (admm.py:532-534),
    then u = soft(G x + alpha/mu, lambd2/mu) (admm.py:460-461, 565-571) and
    alpha += G x - u (admm.py:464-465, 576-593).  Returns (x, err) with
    err[k] = ||G x - u|| after outer step k, and (x, err, u, alpha) when return_state.

    The three steps are the module functions admm_tv_rhs / admm_tv_shrink /
    admm_tv_dual, looked up at call time."""
    dt = dtype
    n = x0.shape[-1]
    x = x0.astype(dt).copy()
    u = np.zeros((2, n, n), dtype=dt)
    al = np.zeros((2, n, n), dtype=dt)
    err = []
    for _ in range(n_iter):
        v = admm_tv_rhs(u, al, mu, dt)
        x, _ = cg(x, b, A, At, n_inner, dtype=dt, lambd1=lambd1, mu=mu, v=v, tv=True,
                  with_eps=False)
        g = np.array([gg.astype(dt) for gg in gradient(x)])
        u = admm_tv_shrink(g, al, mu, lambd2, dt)
        al, r = admm_tv_dual(al, g, u, dt)
        err.append(np.sqrt(np.sum(r.astype(np.float64) ** 2)))
    if return_state:
        return x, np.array(err), u, al
    return x, np.array(err)


def admm_tv_rhs(u, al, mu, dt):
    """v = mu*u - alpha, the TV data term of the inner CG (admm.py:532-534)."""
    return (dt(mu) * u - al).astype(dt)


def admm_tv_shrink(g, al, mu, lambd2, dt):
    """u = soft(G x + alpha/mu, lambd2/mu) (admm.py:460-461)."""
    return soft((g + (al / dt(mu)).astype(dt)).astype(dt), dt(lambd2) / dt(mu)).astype(dt)


def admm_tv_dual(al, g, u, dt):
    """alpha += G x - u, without a factor mu (admm.py:464-465, 576-593).  Returns
    (alpha, G x - u)."""
    r = (g - u).astype(dt)
    return (al + r).astype(dt), r


def calc_filter(nfreq, ftype="ramp"):
    """filters.py:35-76 with dpc = False."""
    if ftype == "none":
        return np.ones(2 * nfreq)
    filtarr = 2 * np.arange(nfreq + 1) / np.float32(2 * nfreq)
    w = 2 * np.pi * np.arange(nfreq + 1) / np.float32(2 * nfreq)
    d = 1.0
    if ftype == "shepp-logan":
        filtarr[1:] *= np.sin(w[1:]) / (2.0 * d * 2.0 * d * w[1:])
    elif ftype == "cosine":
        filtarr[1:] *= np.cos(w[1:]) / (2.0 * d * w[1:])
    elif ftype == "hamming":
        filtarr[1:] *= (0.54 + 0.46 * np.cos(w[1:]) / d)
    elif ftype == "hanning":
        filtarr[1:] *= (1.0 + np.cos(w[1:] / d) / 2.0)
    return np.concatenate((filtarr, filtarr[nfreq - 1:0:-1]), axis=0)


def filter_proj(sino, ftype="ramp"):
    """filters.py:89-113: zero-pad each row to 2*nfreq, multiply by the filter in Fourier
    space, keep the first npix samples (float64 arithmetic, float32 result)."""
    nang, npix = sino.shape
    nfreq = 2 * int(2 ** (int(np.ceil(np.log2(npix)))))
    filtarr = calc_filter(nfreq, ftype)
    padded = np.concatenate((sino, np.zeros((nang, 2 * nfreq - npix))), axis=1)
    out = np.real(np.fft.ifft(np.fft.fft(padded, axis=1) * filtarr, axis=1))
    return out[:, :npix].astype(np.float32)


def fbp(sino, back, nang, ftype="ramp"):
    """class_projectors_radon.py:86-110 with its two defects repaired (dpc = False, the
    filtered array is the one back-projected)."""
    x0 = sino[:, ::-1].copy()
    x0 = filter_proj(x0, ftype)
    reco = back(x0).astype(np.float32) * np.float32(np.pi / (2.0 * nang))
    return reco[::-1, ::-1].copy()
