"""CPU pins of the ADMM-TV restatement (oracle/solvers.admm_tv, after admm.py:460-604 and
cg.py:256-310) by what the mathematics fixes, on a tiny dense system.

The recurrence is: inner CG on (A^T A + l1 I + mu G^T G) x = A^T b + G^T v with
v = mu u - alpha, u = soft(G x + alpha/mu, l2/mu), alpha += G x - u.  Its fixed points:

* alpha stationary gives G x = u;
* u = soft(u + alpha/mu, l2/mu) gives alpha in l2 * d|u|_1: |alpha| <= l2, and
  alpha = l2 sign(G x) wherever G x != 0;
* x stationary needs the inner CG's first residual to vanish.  That residual is
  A^T b + G^T v - (A^T A x + l1 x + G^T G x) (cg.py:258: the G^T G x term carries no mu),
  so A^T A x - A^T b + l1 x + G^T alpha + (1 - mu) G^T G x = 0.  At mu = 1 (the reference's
  default, class_admm_param.py:55-59) these are the KKT conditions of
  1/2 |A x - b|^2 + l1/2 |x|^2 + l2 |G x|_1.

With l2 = 0 and mu = 1 the fixed point is the ridge solution (A^T A + l1 I)^-1 A^T b.
Each plausible slip in the three steps (the sign of alpha in v, the mu in v, alpha/mu in
the shrink, its threshold l2/mu, the sign of the multiplier step) is planted in the
oracle's step functions and must fail these checks; a mis-scaled multiplier step keeps
the fixed point, so it is pinned by the closed form of one step instead.
"""
import numpy as np
import pytest

from oracle import solvers as S


def _dense(n, v, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((v * n, n * n))
    A = lambda x: (M @ x.reshape(-1)).reshape(v, n)
    At = lambda y: (M.T @ y.reshape(-1)).reshape(n, n)
    return M, A, At


def _gmat(n):
    """np.gradient (a library routine) applied to unit images, as a dense matrix."""
    G = np.zeros((2 * n * n, n * n))
    for k in range(n * n):
        e = np.zeros(n * n)
        e[k] = 1.0
        g = np.gradient(e.reshape(n, n))
        G[:, k] = np.concatenate([g[0].reshape(-1), g[1].reshape(-1)])
    return G


N, V, L1, L2 = 4, 6, 0.3, 0.5
M, A, At = _dense(N, V, 11)
B = np.random.default_rng(12).standard_normal((V, N))
G = _gmat(N)


def _kkt(x, u, al, mu, l1=L1, l2=L2):
    """Residuals of the fixed-point conditions above (relative stationarity, G x - u,
    the excess of |alpha| over l2, and alpha - l2 sign(G x) on the support of G x)."""
    xv, a = x.reshape(-1), al.reshape(-1)
    gx = G @ xv
    rx = M.T @ (M @ xv) - M.T @ B.reshape(-1) + l1 * xv + G.T @ a + (1.0 - mu) * (G.T @ gx)
    act = np.abs(gx) > 1e-6
    return {
        "stat": np.linalg.norm(rx) / np.linalg.norm(M.T @ B.reshape(-1)),
        "feas": np.linalg.norm(gx - u.reshape(-1)),
        "box": max(0.0, float(np.max(np.abs(a))) - l2),
        "sign": float(np.max(np.abs(a[act] - l2 * np.sign(gx[act])))) if act.any() else 0.0,
        "active": int(act.sum()),
    }


def _run(n_iter, mu, l2=L2):
    with np.errstate(all="ignore"):
        return S.admm_tv(np.zeros((N, N)), B, A, At, n_iter, 4, L1, l2, mu, dtype=np.float64,
                         return_state=True)


def _kkt_holds(k, tol=1e-8):
    return (np.isfinite(k["stat"]) and k["stat"] <= tol and k["feas"] <= tol
            and k["box"] <= tol and k["sign"] <= tol)


def test_lambda2_zero_converges_to_ridge_solution():
    x, err, u, al = _run(800, 1.0, l2=0.0)
    ridge = np.linalg.solve(M.T @ M + L1 * np.eye(N * N), M.T @ B.reshape(-1))
    np.testing.assert_allclose(x.reshape(-1), ridge, rtol=0, atol=1e-8 * np.abs(ridge).max())
    assert not np.any(al)          # with l2 = 0 and mu = 1 the multiplier stays zero


@pytest.mark.parametrize("mu,n_iter", [(1.0, 1000), (2.0, 2000)])
def test_fixed_point_satisfies_kkt(mu, n_iter):
    x, err, u, al = _run(n_iter, mu)
    k = _kkt(x, u, al, mu)
    assert _kkt_holds(k), k
    # the solution has both zero and nonzero gradient entries, so the box and the
    # sign conditions are both exercised
    assert 0 < k["active"] < G.shape[0], k


def test_mu_two_fixed_point_carries_the_reference_residual_term():
    # cg.py:258 computes G^T G x without mu: at mu != 1 the textbook KKT (no (1 - mu)
    # term) does not hold, the reference's does
    x, err, u, al = _run(2000, 2.0)
    assert _kkt_holds(_kkt(x, u, al, 2.0))
    textbook = _kkt(x, u, al, 1.0)
    assert textbook["stat"] > 1e-3, textbook


def test_multiplier_step_has_no_mu_factor():
    # one outer step from alpha = 0 with a threshold so large that u = 0: the step
    # alpha += G x - u (admm.py:576-593) leaves alpha = G x exactly (mu alpha would not)
    x, err, u, al = _run(1, 2.0, l2=1e6)
    assert not np.any(u)
    gx = np.array(np.gradient(x))
    np.testing.assert_array_equal(al, gx)


PLANTED = {
    # v = mu u + alpha (sign of alpha in the inner right-hand side)
    "rhs_sign": ("admm_tv_rhs", lambda u, al, mu, dt: (dt(mu) * u + al).astype(dt), (1.0, 2.0)),
    # v = u - alpha (mu dropped)
    "rhs_no_mu": ("admm_tv_rhs", lambda u, al, mu, dt: (u - al).astype(dt), (2.0,)),
    # soft(G x + alpha * mu, l2/mu)
    "shrink_alpha_times_mu": ("admm_tv_shrink",
                              lambda g, al, mu, l2, dt: S.soft(g + al * dt(mu), dt(l2) / dt(mu)),
                              (2.0,)),
    # soft(G x + alpha/mu, l2)
    "shrink_threshold": ("admm_tv_shrink",
                         lambda g, al, mu, l2, dt: S.soft(g + al / dt(mu), dt(l2)), (2.0,)),
    # alpha -= G x - u
    "dual_sign": ("admm_tv_dual", lambda al, g, u, dt: (al - (g - u), g - u), (1.0, 2.0)),
}


@pytest.mark.parametrize("name", sorted(PLANTED))
def test_planted_mistakes_fail_the_fixed_point_checks(name, monkeypatch):
    fn, bad, mus = PLANTED[name]
    monkeypatch.setattr(S, fn, bad)
    for mu in mus:
        x, err, u, al = _run(600, mu)
        k = _kkt(x, u, al, mu)
        assert not _kkt_holds(k, tol=1e-6), (name, mu, k)
