"""Learned warm start for SIRT: training pairs made by the method itself, training of
the U-Net (unet.py) on them, and the time-to-residual comparison.

A pair is (input, target) for one random phantom: the phantom is projected with the
distance-driven projector, optionally made noisy (Poisson counts with an incident
intensity I0 drawn log-uniformly from a range, after scaling the largest line integral
to ``mu_max``), the input is the filtered back projection of the sinogram and the
target is the plain SIRT iterate after ``sirt_iters`` iterations from zero
(non-negativity on).  The warm start is then x0 = U-Net(FBP(b)); plain SIRT continues
from x0 and is stopped at the residual ||A x - b|| / ||b|| that plain SIRT from zero
reached after a given number of iterations.  Plain SIRT stays the fallback and decides
the result: the network only moves the starting point.

Everything runs on the GPU through libitomo; torch allocates and, for the synthetic
Poisson noise only, draws random numbers.
"""
import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .core import Plan, device, filter_sino
from .phantom import angles_deg, ellipses, textured
from .unet import UNet

__all__ = ["dd_plan", "fbp", "sirt", "residual", "make_pairs", "train", "iterations_to",
           "warm_start", "compare_starts", "save_weights", "load_weights"]


def dd_plan(npix, nang):
    """Distance-driven plan over nang views uniform on [0, 180) degrees (angles in the
    order the reference C receives them)."""
    facade = (angles_deg(nang) * np.pi / 180.0).astype(np.float32)
    return Plan("dd", npix, np.fft.fftshift(facade))


def fbp(plan, sino, ftype="ramp"):
    """Filtered back projection of (B, V, N) sinograms, the steps of
    class_projectors_radon.py:86-110: reverse the bins, filter, back-project, scale by
    pi / (2 V), rotate by 180 degrees."""
    x0 = filter_sino(torch.flip(sino, dims=(-1,)).contiguous(), ftype)
    reco = plan.back(x0)
    reco.mul_(np.float32(np.pi / (2.0 * plan.nang)))
    return torch.flip(reco, dims=(-2, -1)).contiguous()


_GRIDREC = {}


def gridrec_fbp(plan, sino):
    """Gridding (gridrec) filtered back projection of the distance-driven plan's (B, V, N)
    sinograms: their rows are the views of angles_deg(V) in order, with the bins in the
    gridding pair's orientation, and the rotation axis lies between the two middle bins,
    i.e. at N/2 - 1/2 in gridrec's centre coordinate (the PSWF defaults of
    class_projectors_grid.py:50)."""
    from .gridrec import GridrecPlan
    key = (plan.npix, plan.nang, torch.cuda.current_device())
    if key not in _GRIDREC:
        _GRIDREC[key] = GridrecPlan(plan.npix, angles_deg(plan.nang), ctr=plan.npix * 0.5 - 0.5)
    return _GRIDREC[key].back(sino.contiguous(), filtered=True)


def fbp_start(plan, sino, kind="dd"):
    """The FBP start: "dd" (fbp above) or "gridrec" (gridrec_fbp)."""
    if kind == "gridrec":
        return gridrec_fbp(plan, sino)
    if kind != "dd":
        raise ValueError("FBP kind %r (have 'dd', 'gridrec')" % kind)
    return fbp(plan, sino)


def sirt(plan, b, iters, x0=None, pc_beta=0.0):
    """Plain SIRT (art.py:212, 314-321) on (B, V, N) sinograms, all slices in each
    launch; returns the (B, N, N) iterate."""
    nz = b.shape[0]
    R, C = plan.sirt_weights()
    x = torch.zeros((nz, plan.npix, plan.npix), device=b.device) if x0 is None \
        else x0.clone().contiguous()
    work = torch.empty_like(b)
    for _ in range(iters):
        plan.sirt_iterate(x, b, R, C, work, pc_beta=pc_beta)
    return x


def residual(plan, x, b):
    """||A x - b|| / ||b|| per slice (float64 CPU numpy array)."""
    r = plan.forward(x.contiguous()) - b
    num = torch.linalg.vector_norm(r.double().reshape(r.shape[0], -1), dim=1)
    den = torch.linalg.vector_norm(b.double().reshape(b.shape[0], -1), dim=1)
    return (num / den).cpu().numpy()


def _noisy(plan, x_true, i0, mu_max, gen):
    b = plan.forward(x_true)
    peak = b.reshape(b.shape[0], -1).amax(dim=1).clamp_min(1e-6)
    scale = (mu_max / peak)[:, None, None]
    lam = i0[:, None, None] * torch.exp(-b * scale)
    counts = torch.poisson(lam, generator=gen)
    return (-torch.log(counts.clamp_min(1.0) / i0[:, None, None]) / scale).contiguous()


def make_pairs(plan, count, sirt_iters, seed=0, noise=(1e3, 1e5), mu_max=3.0,
               n_ellipses=10, texture=0.15, chunk=64, input_iters=0, fbp_kind="dd",
               verbose=False, target="sirt"):
    """(inputs, targets, sinograms, phantoms), each (count, ...) float32 on the GPU.
    ``noise`` = (I0_lo, I0_hi) or None for noiseless data.  The input is the FBP of
    ``fbp_kind`` (``input_iters`` = 0) or the plain SIRT iterate after ``input_iters``
    iterations; the target is the SIRT iterate after ``sirt_iters`` iterations
    (``target`` = "sirt") or the phantom itself ("truth")."""
    if target not in ("sirt", "truth"):
        raise ValueError("target %r (have 'sirt', 'truth')" % target)
    dev = device()
    gen = torch.Generator(device=dev).manual_seed(seed)
    rng = np.random.default_rng(seed)
    ins, outs, sinos, truths = [], [], [], []
    for k0 in range(0, count, chunk):
        nz = min(chunk, count - k0)
        ph = np.stack([textured(plan.npix, n_ellipses, seed=seed * 100003 + k0 + k,
                                texture=texture) if texture > 0
                       else ellipses(plan.npix, n_ellipses, seed=seed * 100003 + k0 + k)
                       for k in range(nz)])
        x_true = torch.from_numpy(ph).to(dev)
        if noise is None:
            b = plan.forward(x_true)
        else:
            i0 = torch.from_numpy(np.exp(rng.uniform(math.log(noise[0]), math.log(noise[1]),
                                                     nz)).astype(np.float32)).to(dev)
            b = _noisy(plan, x_true, i0, mu_max, gen)
        if input_iters > 0:
            xi = sirt(plan, b, input_iters)
            ins.append(xi)
            outs.append(x_true if target == "truth" else
                        sirt(plan, b, sirt_iters - input_iters, x0=xi))
        else:
            ins.append(fbp_start(plan, b, fbp_kind))
            outs.append(x_true if target == "truth" else sirt(plan, b, sirt_iters))
        sinos.append(b)
        truths.append(x_true)
        if verbose:
            torch.cuda.synchronize()
            print("pairs %d/%d" % (k0 + nz, count), flush=True)
    return torch.cat(ins), torch.cat(outs), torch.cat(sinos), torch.cat(truths)


def train(net, inputs, targets, steps, batch=16, lr=1e-3, seed=0, crop=None, log_every=0):
    """Adam on random mini-batches (optionally random crops of side ``crop``, a multiple
    of 8); returns the per-step losses (numpy)."""
    g = torch.Generator().manual_seed(seed)
    n, h, w = inputs.shape
    losses = torch.empty(steps, dtype=torch.float64, device=inputs.device)
    for s in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), generator=g).to(inputs.device)
        xi, yi = inputs[idx], targets[idx]
        if crop is not None and crop < h:
            oy = int(torch.randint(0, h - crop + 1, (1,), generator=g))
            ox = int(torch.randint(0, w - crop + 1, (1,), generator=g))
            xi = xi[:, oy:oy + crop, ox:ox + crop]
            yi = yi[:, oy:oy + crop, ox:ox + crop]
        lr_s = lr * 0.5 * (1.0 + math.cos(math.pi * s / steps))   # cosine decay
        losses[s] = net.train_step(xi.contiguous(), yi.contiguous(), lr=lr_s)
        if log_every and (s + 1) % log_every == 0:
            print("step %d loss %.4e" % (s + 1, float(losses[s])), flush=True)
    return losses.cpu().numpy()


def iterations_to(plan, b, x0, target, max_iters, pc_beta=0.0):
    """Plain SIRT from x0 until every slice's residual is <= target[slice] (residual of
    the iterate after k steps, evaluated after every step).  Returns (iterations per
    slice, iterate per slice after exactly that many steps); a slice that never gets
    there counts max_iters + 1 and returns its iterate after max_iters steps."""
    nz = b.shape[0]
    R, C = plan.sirt_weights()
    x = x0.clone().contiguous()
    out = x0.clone().contiguous()
    prev = torch.empty_like(x)
    work = torch.empty_like(b)
    stats = torch.zeros(3 * nz, dtype=torch.float64, device=b.device)
    bn = torch.linalg.vector_norm(b.double().reshape(nz, -1), dim=1)
    tgt = torch.as_tensor(np.asarray(target, np.float64), device=b.device)
    need = torch.full((nz,), max_iters + 1, dtype=torch.int64, device=b.device)
    r0 = torch.as_tensor(residual(plan, x, b), device=b.device)
    need[r0 <= tgt] = 0
    for k in range(1, max_iters + 2):
        prev.copy_(x)                                   # x_{k-1}
        plan.sirt_iterate(x, b, R, C, work, pc_beta=pc_beta, stats=stats)
        # stats[3z] = ||b - A x_{k-1}||^2: the residual of the iterate entering step k
        rk = stats[0::3].sqrt() / bn
        newly = (rk <= tgt) & (need > max_iters)
        need[newly] = k - 1
        out = torch.where(newly[:, None, None], prev, out)
        if k % 16 == 0 and bool((need <= max_iters).all()):
            break
    # slices that never got there: prev holds x_{max_iters} after the last step
    out = torch.where((need > max_iters)[:, None, None], prev, out)
    return need.cpu().numpy(), out


def warm_start(plan, net, b, x_in=None, pc_beta=0.0):
    """x0 = net(x_in) (x_in defaults to FBP(b)) per slice, with plain SIRT as the
    fallback: a slice whose network output is not finite, or whose residual
    ||A x0 - b|| / ||b|| is not below that of the zero start (1) and of x_in, starts
    from zero instead.  Non-negativity (pc_beta >= 0) is applied to x0 as SIRT applies
    it to every iterate.  Returns (x0, mask of slices that fell back)."""
    if x_in is None:
        x_in = fbp(plan, b)
    x0 = net(x_in).contiguous()
    if pc_beta >= 0:
        x0 = torch.where(x0 < 0, x0 * np.float32(pc_beta), x0)
    finite = torch.isfinite(x0.reshape(x0.shape[0], -1)).all(dim=1)
    x_safe = torch.where(finite[:, None, None], x0, torch.zeros_like(x0))
    r_net = torch.as_tensor(residual(plan, x_safe, b), device=b.device)
    r_in = torch.as_tensor(residual(plan, x_in.contiguous(), b), device=b.device)
    ok = finite & (r_net < 1.0) & (r_net < r_in)
    fell_back = ~ok
    x0 = torch.where(ok[:, None, None], x_safe, torch.zeros_like(x_safe)).contiguous()
    return x0, fell_back.cpu().numpy()


def _time_ms(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def compare_starts(plan, net, b, plain_iters, max_iters, input_iters=0, fbp_kind="dd"):
    """Time for plain SIRT from zero, from FBP and from the network's output to reach the
    residual plain SIRT from zero has after ``plain_iters`` iterations, on the (B, V, N)
    sinograms b.  The network's input is FBP(b) of ``fbp_kind`` (``input_iters`` = 0) or
    the plain SIRT iterate after ``input_iters`` iterations, whose cost is then part of
    its time.  Wall time = start-up (FBP, input iterations, network) + iterations x
    measured time per iteration for the whole batch."""
    nz = b.shape[0]
    x_plain = sirt(plan, b, plain_iters)
    target = residual(plan, x_plain, b) * (1.0 + 1e-9)
    zero = torch.zeros((nz, plan.npix, plan.npix), device=b.device)
    x_fbp = fbp_start(plan, b, fbp_kind)
    x_in = sirt(plan, b, input_iters) if input_iters > 0 else x_fbp
    x_net = net(x_in)
    k_zero, x_z = iterations_to(plan, b, zero, target, max_iters)
    k_fbp, _ = iterations_to(plan, b, x_fbp, target, max_iters)
    k_net, x_w = iterations_to(plan, b, x_net, target, max_iters)
    R, C = plan.sirt_weights()
    xx = zero.clone()
    work = torch.empty_like(b)
    t_iter = _time_ms(lambda: plan.sirt_iterate(xx, b, R, C, work, pc_beta=0.0), reps=10)
    t_fbp = _time_ms(lambda: fbp_start(plan, b, fbp_kind))
    t_net = _time_ms(lambda: net(x_in))
    t_in = t_fbp if input_iters == 0 else input_iters * t_iter
    kz, kf, kn = int(k_zero.max()), int(k_fbp.max()), int(k_net.max())
    # each slice's warm iterate at its own stopping step against plain's at its own
    diff = torch.linalg.vector_norm((x_w - x_z).double().reshape(nz, -1), dim=1)
    ref = torch.linalg.vector_norm(x_z.double().reshape(nz, -1), dim=1)
    rel_slices = (diff / ref).cpu().numpy()
    rel_to_plain = float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(ref))
    return {
        "slices": nz, "npix": plan.npix, "nang": plan.nang, "plain_iters": plain_iters,
        "target_residual_max": float(target.max()),
        "iters_zero": kz, "iters_fbp": kf, "iters_unet": kn,
        "ms_per_iter": t_iter, "ms_fbp": t_fbp, "ms_unet": t_net,
        "seconds_zero": kz * t_iter / 1e3,
        "seconds_fbp": (t_fbp + kf * t_iter) / 1e3,
        "input_iters": input_iters, "fbp": fbp_kind,
        "seconds_unet": (t_in + t_net + kn * t_iter) / 1e3,
        "rel_l2_unet_vs_plain": rel_to_plain,
        "rel_l2_unet_vs_plain_max_slice": float(rel_slices.max()),
        "residual_fbp": float(residual(plan, x_fbp, b).max()),
        "residual_unet": float(residual(plan, x_net, b).max()),
    }


def save_weights(net, path):
    """FP32 master weights as an .npz of plain arrays (no pickled objects)."""
    np.savez(path, **{k: v.numpy().astype(np.float32) for k, v in net.state_dict().items()})


def load_weights(path, seed=0):
    """A UNet with the weights of an .npz written by save_weights."""
    net = UNet(seed=seed)
    with np.load(path, allow_pickle=False) as z:
        net.load_state_dict({k: torch.from_numpy(z[k]) for k in z.files})
    return net


def main(argv=None):
    ap = argparse.ArgumentParser(description="train and evaluate the SIRT warm start")
    ap.add_argument("--npix", type=int, default=256)
    ap.add_argument("--nang", type=int, default=180)
    ap.add_argument("--pairs", type=int, default=1024)
    ap.add_argument("--sirt-iters", type=int, default=500)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=0)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--noiseless", action="store_true")
    ap.add_argument("--test", type=int, default=32)
    ap.add_argument("--max-iters", type=int, default=2000)
    ap.add_argument("--ellipses", type=int, default=10)
    ap.add_argument("--texture", type=float, default=0.15)
    ap.add_argument("--chunk", type=int, default=64, help="slices per SIRT launch")
    ap.add_argument("--input-iters", type=int, default=0,
                    help="network input: 0 = FBP, k = plain SIRT after k iterations")
    ap.add_argument("--plain-iters", type=int, default=0,
                    help="target residual: plain SIRT after this many iterations "
                         "(default: --sirt-iters)")
    ap.add_argument("--fbp", default="dd", choices=("dd", "gridrec"),
                    help="FBP of the network input and of the FBP start")
    ap.add_argument("--target", default="sirt", choices=("sirt", "truth"),
                    help="training target: the SIRT iterate after --sirt-iters, or the phantom")
    ap.add_argument("--save", default="")
    args = ap.parse_args(argv)
    plan = dd_plan(args.npix, args.nang)
    noise = None if args.noiseless else (1e3, 1e5)
    t0 = time.perf_counter()
    xin, xout, _, _ = make_pairs(plan, args.pairs, args.sirt_iters, seed=1, noise=noise,
                                 n_ellipses=args.ellipses, texture=args.texture,
                                 chunk=args.chunk, input_iters=args.input_iters,
                                 fbp_kind=args.fbp, verbose=True, target=args.target)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    net = UNet(seed=0)
    t0 = time.perf_counter()
    losses = train(net, xin, xout, args.steps, args.batch, args.lr,
                   crop=args.crop or None, log_every=max(1, args.steps // 10))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del xin, xout
    _, _, b_test, _ = make_pairs(plan, args.test, 1, seed=2, noise=noise,
                                 n_ellipses=args.ellipses, texture=args.texture,
                                 chunk=args.chunk)
    res = compare_starts(plan, net, b_test, args.plain_iters or args.sirt_iters,
                         args.max_iters, input_iters=args.input_iters, fbp_kind=args.fbp)
    res.update({"pairs": args.pairs, "train_steps": args.steps, "target": args.target,
                "seconds_data": t_data,
                "seconds_train": t_train, "loss_first": float(losses[:10].mean()),
                "loss_last": float(losses[-10:].mean()), "parameters": net.num_parameters()})
    if args.save:
        save_weights(net, args.save)
    print(json.dumps(res), flush=True)
    return res


if __name__ == "__main__":
    main(sys.argv[1:])
