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


def sirt(x0, b, A, At, n_iter, pc=-1.0, dtype=np.float32):
    """Returns (x, info, resid) where info rows are [diff, obj] as in art.py:325-345 and
    resid[k] = ||b - A x_k||_2 for the iterate entering step k (in float64)."""
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
        if pc >= 0.0:
            neg = x < 0
            x[neg] = x[neg] * dtype(pc)                                       # utils.py:171-176
        diff = np.linalg.norm((x - x_old).astype(np.float64)) / np.linalg.norm(
            x_old.astype(np.float64))
        obj = np.linalg.norm((A(x) - b).astype(np.float64)) ** 2
        info.append([diff, obj])
    return x, np.array(info), np.array(resid)


def mlem(x0, b, A, At, n_iter, dtype=np.float32):
    """Returns (x, err) with err[k] = sum (A x_k - b)^2 / sum b^2 (sir.py:287),
    evaluated in float64 from the float32 arrays."""
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
    return x, np.array(err)
