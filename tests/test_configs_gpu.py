"""The benchmark suite's other workloads, at their own sizes, on the GPU.

* MLEM at 2048^2 (Poisson counts, I0 = 1e4, from ones): two steps on a plan over 15 of the
  1800 views, chosen so that every forward path of the distance-driven kernel runs
  (LDS-staged with 3-6 candidate pixels per bin and line, gathered with 7-8, staged in
  the wide launch with 9-32, gathered beyond 32; the MLEM epilogue is fused into each):
  bitwise equal to the restated sir.py:264-283 over the CPU oracle.
* The 1024^2 x 900 stack: a persistent batch of 8 slices, and one mixed launch holding
  2048^2 x 1440, 1024^2 x 900, 512^2 x 360 and 256^2 x 180 members: every slice bitwise
  equal to its per-plan run; in each member one view of the forward half (the SIRT
  update R (b - A x)) bitwise equal to the oracle.
* CG and ADMM-TV at 1024^2 x 720 with Gaussian noise of 1% of the largest projection:
  every FP64 scalar the kernels keep on the device equals torch's FP64 dot of the same
  float32 vectors to 1e-12; the ADMM-TV stationarity residual
  ||A^T A x - A^T b + l1 x + G^T alpha|| / ||A^T b|| falls from step 2 to step 10, and the
  split residual ||G x - u|| is past its peak; on a 45-view subset at 1024^2 the float32
  GPU results after 5 iterations are compared with the FP64 CPU oracle (error printed,
  bound 1e-3).
* The warm-start network at 512^2, batch 64: 32 x 32 output tiles of the full-batch
  GPU forward against the FP64 restatement (BF16 rounding at the same points) run on a
  crop with a 64-pixel halo, beyond the network's receptive field.
"""
import numpy as np
import pytest
import torch

from oracle import radon as O
from oracle import solvers as S
from oracle import unet_ref as U
from iterative_tomographic_reconstruction_algorithms_amd import _lib as L
from iterative_tomographic_reconstruction_algorithms_amd import solvers
from iterative_tomographic_reconstruction_algorithms_amd.batch import Batch, Member
from iterative_tomographic_reconstruction_algorithms_amd.core import Plan, _p, stream_handle
from iterative_tomographic_reconstruction_algorithms_amd.phantom import (
    angles_deg, ellipses, ellipsoid_slices, textured)
from iterative_tomographic_reconstruction_algorithms_amd.unet import UNet

pytestmark = pytest.mark.gpu


def c_angles(v):
    """The fftshift-ed list the distance-driven C receives for v views on [0, 180)."""
    return np.fft.fftshift((angles_deg(v) * np.pi / 180).astype(np.float32))


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---- MLEM at 2048^2 ---------------------------------------------------------------------

# views of the 1800-view list (fftshift-ed order) and the forward path each takes at 2048^2
MLEM_VIEWS = [0, 1, 450, 600, 755, 1500,     # staged, 3-6 candidates
              804,                            # gathered, 7-8
              1740, 843, 859, 876, 881,       # wide launch, staged 9-32
              890,                            # wide launch, gathered beyond 32
              899, 900]                       # next to the column-sum pole


def test_mlem_2048_every_forward_path_bitwise():
    n = 2048
    c = np.ascontiguousarray(c_angles(1800)[MLEM_VIEWS])
    plan = Plan("dd", n, c)
    given = np.fft.ifftshift(c)                 # the oracle applies the wrapper's fftshift
    A = lambda z: O.dd_forward(z, given, O.FAST)
    At = lambda z: O.dd_back(z, given, O.FAST)
    p = A(ellipses(n, 40, seed=2))
    b = np.random.default_rng(2).poisson(1e4 * p / p.max()).astype(np.float32)
    inv_d = plan.mlem_weights()
    x = torch.ones((n, n), dtype=torch.float32, device="cuda")
    bd = torch.from_numpy(b).cuda()
    work = torch.empty_like(bd)
    stats = torch.zeros(3, dtype=torch.float64, device="cuda")
    bsq = float(np.sum(b.astype(np.float64) ** 2))
    errs = []
    for _ in range(2):
        plan.mlem_iterate(x, bd, inv_d, work, stats=stats)
        errs.append(float(stats[0]) / bsq)
    x_ref, err_ref = S.mlem(np.ones((n, n), np.float32), b, A, At, 2)
    np.testing.assert_array_equal(x.cpu().numpy(), x_ref)
    np.testing.assert_allclose(errs, err_ref, rtol=1e-12)


# ---- the 1024^2 x 900 stack and the mixed launch -------------------------------------------

def dd_plan(n, v):
    return Plan("dd", n, c_angles(v))


def eager_sirt(plan, b, iters):
    R, C = plan.sirt_weights()
    x = torch.zeros((b.shape[0], plan.npix, plan.npix), dtype=torch.float32, device="cuda")
    w = torch.empty_like(b)
    for _ in range(iters):
        plan.sirt_iterate(x, b, R, C, w, pc_beta=0.0)
    return x


def check_forward_view(plan, member, x_prev, view):
    """Row of member.work for `view` after a step from x_prev: R (b - A x_prev) with the
    oracle's projection of that one view (art.py:212, 227-230)."""
    n = plan.npix
    ang = np.array([plan.angles[view]], np.float32)
    row = plan.nang - 1 - view                   # radon_dd.c writes view v to row V-1-v
    a1 = O.dd_forward(np.ones((n, n), np.float32), ang, O.FAST)[0]
    a1[a1 == 0] += np.float32(1e-10)
    R = (np.float32(1.0) / a1).astype(np.float32)
    b = member.b.cpu().numpy()
    work = member.work.cpu().numpy()
    xp = x_prev.cpu().numpy()
    for z in range(member.nslices):
        ax = O.dd_forward(xp[z], ang, O.FAST)[0]
        want = (R * (b[z, row] - ax).astype(np.float32)).astype(np.float32)
        np.testing.assert_array_equal(work[z, row], want, err_msg="n=%d view %d slice %d"
                                      % (n, view, z))


def test_stack_batch_of_8_slices_1024x900():
    n, v, nz = 1024, 900, 8
    plan = dd_plan(n, v)
    vol = torch.from_numpy(ellipsoid_slices(n, 1024, 60, seed=3000, first=508, count=nz))
    b = plan.forward(vol.cuda())
    m = Member(plan, b)
    batch = Batch([m], pc_beta=0.0)
    batch.iterate()
    batch.iterate()
    torch.cuda.synchronize()
    assert torch.equal(m.x, eager_sirt(plan, b, 2))
    check_forward_view(plan, m, eager_sirt(plan, b, 1), 317)
    batch.close()


def test_mixed_launch_with_2048x1440_and_256x180_members():
    geoms = [(2048, 1440, 1), (1024, 900, 2), (512, 360, 2), (256, 180, 3)]
    members = []
    for k, (n, v, nz) in enumerate(geoms):
        plan = dd_plan(n, v)
        imgs = np.stack([ellipses(n, 10, seed=40 + 7 * k + z) for z in range(nz)])
        members.append(Member(plan, plan.forward(torch.from_numpy(imgs).cuda())))
    batch = Batch(members, pc_beta=0.0)
    batch.iterate()
    batch.iterate()
    torch.cuda.synchronize()
    for m in members:
        assert torch.equal(m.x, eager_sirt(m.plan, m.b, 2)), m.plan.npix
        check_forward_view(m.plan, m, eager_sirt(m.plan, m.b, 1), m.plan.nang // 3)
    batch.close()


# ---- CG and ADMM-TV at 1024^2 x 720 ----------------------------------------------------------

def noisy_sino(plan, seed):
    b = plan.forward(torch.from_numpy(ellipses(plan.npix, 10, seed=seed)).cuda())
    gen = torch.Generator(device="cuda").manual_seed(seed)
    noise = torch.randn(b.shape, device="cuda", generator=gen)
    return (b + 0.01 * float(b.max()) * noise).contiguous()


def ddot(a, b):
    return float(torch.dot(a.double().reshape(-1), b.double().reshape(-1)))


def test_cg_fp64_scalars_at_1024x720():
    n, v = 1024, 720
    plan = dd_plan(n, v)
    b = noisy_sino(plan, 4)
    lib, st = L.lib(), stream_handle()
    vw = solvers.VecWork(1)
    x = torch.zeros((1, n, n), dtype=torch.float32, device="cuda")
    r = plan.back(b).reshape(1, n, n)           # A^T b - A^T A x at x = 0 (cg.py:85)
    p = r.clone()
    work = torch.empty((1, v, n), dtype=torch.float32, device="cuda")
    Hp = torch.empty_like(x)
    L.check(lib.itomo_cg_dot(_p(r), _p(r), n * n, 1, 0, _p(vw.scal), _p(vw.work), st), "dot")
    assert abs(float(vw.scal[0]) - ddot(r, r)) <= 1e-12 * ddot(r, r)
    rr = [float(vw.scal[0])]
    for it in range(6):
        plan.forward(p, out=work)
        plan.back(work, out=Hp)
        L.check(lib.itomo_cg_hp(_p(Hp), _p(p), None, n, 1, 0.0, 0.0, _p(vw.scal),
                                _p(vw.work), st), "cg_hp")
        php = ddot(p, Hp)
        assert abs(float(vw.scal[1]) - php) <= 1e-12 * abs(php), it
        L.check(lib.itomo_cg_update(_p(x), _p(r), _p(p), _p(Hp), n * n, 1, _p(vw.scal),
                                    _p(vw.work), st), "cg_update")
        rn = ddot(r, r)
        rn_dev = float(vw.scal[2])
        assert abs(rn_dev - rn) <= 1e-12 * rn, it
        rr.append(rn)
        L.check(lib.itomo_cg_direction(_p(p), _p(r), n * n, 1, 1, _p(vw.scal), st), "dir")
        assert float(vw.scal[0]) == rn_dev          # the new <r, r> becomes the old
    # CG on the normal equations: ||A x - b|| falls every step
    assert rr[-1] < 0.05 * rr[0]


def stationarity(plan, x, b, alpha, lambd1):
    """||A^T A x - A^T b + l1 x + G^T alpha|| / ||A^T b|| (mu = 1)."""
    atb = plan.back(b)
    s = plan.back(plan.forward(x)) - atb + np.float32(lambd1) * x + solvers.grad_adjoint(alpha)
    return float(torch.linalg.vector_norm(s.double()) / torch.linalg.vector_norm(atb.double()))


def test_admm_tv_residuals_fall_at_1024x720():
    n, v = 1024, 720
    plan = dd_plan(n, v)
    b = noisy_sino(plan, 4)
    out = {}
    for k in (2, 10):
        x0 = torch.zeros((1, n, n), dtype=torch.float32, device="cuda")
        x, err, u, al = solvers.admm_tv(plan, x0, b, k, n_inner=4, return_state=True)
        assert torch.isfinite(x).all()
        out[k] = (stationarity(plan, x[0], b, al[0], 1.0), err)
    print("ADMM-TV 1024^2 x 720: stationarity %.3e after 2, %.3e after 10; ||Gx-u|| %s"
          % (out[2][0], out[10][0], np.array2string(out[10][1], precision=3)))
    assert out[10][0] < out[2][0]
    # ||G x - u|| first grows while u fills in, then falls: past its peak by step 10
    assert out[10][1][-1] < out[10][1].max()


@pytest.mark.parametrize("which", ["cg", "admm_tv"])
def test_error_against_fp64_oracle_1024(which):
    n, v = 1024, 45
    facade = (angles_deg(v) * np.pi / 180).astype(np.float32)
    plan = Plan("dd", n, np.fft.fftshift(facade))
    img = ellipses(n, 10, seed=4)
    b32 = O.dd_forward(img, facade, O.FAST)
    b32 = (b32 + 0.01 * b32.max() * np.random.default_rng(4).standard_normal(b32.shape)
           ).astype(np.float32)
    A64 = lambda z: O.forward("dd", z, facade, O.F64)
    At64 = lambda z: O.back("dd", z, facade, O.F64)
    x0 = np.zeros((n, n), np.float32)
    bd = torch.from_numpy(b32).cuda()
    if which == "cg":
        xg, _ = solvers.cg(plan, torch.from_numpy(x0).cuda(), bd, 5)
        x64, _ = S.cg(x0.astype(np.float64), b32.astype(np.float64), A64, At64, 5,
                      dtype=np.float64)
    else:
        xg, _ = solvers.admm_tv(plan, torch.from_numpy(x0).cuda(), bd, 5, n_inner=4)
        x64, _ = S.admm_tv(x0.astype(np.float64), b32.astype(np.float64), A64, At64, 5, 4,
                           dtype=np.float64)
    e = rel(xg.cpu().numpy().reshape(n, n), x64)
    print("%s, 1024^2 x 45 views, 5 iterations: relative L2 to the FP64 oracle %.2e"
          % (which, e))
    assert e <= 1e-3


# ---- the warm-start network at 512^2, batch 64 ---------------------------------------------

def _rand_net(seed):
    net = UNet(seed=seed)
    g = torch.Generator().manual_seed(seed + 1)
    net.hw.copy_((torch.randn(32, generator=g) * 0.2).cuda())
    net.hb.fill_(0.1)
    for name in net.b:
        net.b[name].copy_((torch.randn(net.b[name].numel(), generator=g) * 0.05).cuda())
    net.refresh_bf16()
    return net


def test_unet_forward_512_batch64_tiles_match_oracle():
    net = _rand_net(31)
    imgs = np.stack([textured(512, 10, seed=900 + k) for k in range(64)])
    out = net(torch.from_numpy(imgs).cuda()).cpu().double()
    params = {}
    for name, t, tbf in net.params():
        params[name] = (tbf if tbf is not None else t).detach().double().cpu()
    halo, tile = 64, 32
    worst, worst_exact = 0.0, 0.0
    for z, oy, ox in [(0, 0, 0), (17, 192, 256), (40, 480, 96), (63, 352, 480)]:
        y0, x0 = max(0, oy - halo), max(0, ox - halo)
        y1, x1 = min(512, oy + tile + halo), min(512, ox + tile + halo)
        crop = torch.from_numpy(imgs[z, y0:y1, x0:x1]).double()[None]
        sl = (slice(oy - y0, oy - y0 + tile), slice(ox - x0, ox - x0 + tile))
        img_t = crop[0][sl]
        got = out[z, oy:oy + tile, ox:ox + tile] - img_t
        ref = U.forward(params, crop, bf16=True)[0][sl] - img_t
        exact = U.forward(params, crop, bf16=False)[0][sl] - img_t
        worst = max(worst, rel(got, ref))
        worst_exact = max(worst_exact, rel(got, exact))
    print("U-Net 512^2 x 64 tiles: rel L2 of the correction vs BF16-faithful oracle %.2e, "
          "vs FP64 network %.2e" % (worst, worst_exact))
    assert worst <= 2e-2
    assert worst_exact <= 1.5e-1
