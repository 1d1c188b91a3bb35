"""Benchmark: plain SIRT on the distance-driven projector pair at 2048^2 x 1800 views.

One step = one fused SIRT iteration  x <- x + C * A^T(R * (b - A x)), x >= 0, with the
FP64 monitors (residual, step size, norm) computed inside the kernels -- the work of one
iteration of algorithms/art.py:sirt (art.py:212, 314-326) on one slice.

Workload: 2048 x 2048 float32 random-ellipse phantom (40 ellipses, seed 2), 1800 views
uniform over [0, 180) degrees, 2048 bins, noiseless b = A x_true computed on the GPU,
SIRT from x = 0.  Inputs are resident in HBM before timing starts.

Metric: Gupdates/s = 2 * N^2 * V * iterations / seconds / 1e9 (one pixel-view update per
projector application).  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                  [--residual-target 1e-3] [--max-residual-iters M]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, V, ELLIPSES, SEED = 2048, 1800, 40, 2
HBM_PEAK_GBS = 8000.0


def alg_bytes_forward(n, v):
    # gathers (one image value per pixel-view) + x read + b, R read + work write
    return 4.0 * n * n * v + 4.0 * n * n + 3 * 4.0 * v * n


def alg_bytes_back(n, v):
    # gathers (one sinogram value per pixel-view) + work read + x, C read + x write
    return 4.0 * n * n * v + 4.0 * v * n + 3 * 4.0 * n * n


def cpu_baseline(img, c_angles, budget_views=20):
    """Reference C (radon_dd.c built with its own flags into oracle/_ref) timed on a
    bounded sample: forward + back of `budget_views` evenly spaced views at full N,
    single thread (the reference C is single-threaded)."""
    import numpy as np
    from oracle import ref, radon as O
    sub = np.ascontiguousarray(c_angles[:: max(1, len(c_angles) // budget_views)][:budget_views])
    # the wrapper fftshifts; hand it the list that shifts back to `sub`
    given = np.fft.ifftshift(sub)
    if ref.available("dd", "fast"):
        kind = "reference"
        fwd = lambda: ref.forward("dd", img, given, "fast")
        back = lambda s: ref.back("dd", s, given, "fast")
    else:
        kind = "port"
        fwd = lambda: O.dd_forward(img, given, O.FAST)
        back = lambda s: O.dd_back(s, given, O.FAST)
    t0 = time.perf_counter()
    s = fwd()
    t1 = time.perf_counter()
    back(s)
    t2 = time.perf_counter()
    secs = t2 - t0
    gups = 2.0 * img.shape[0] ** 2 * len(sub) / secs / 1e9
    return {"value": round(gups, 6), "unit": "Gupdates/s", "cores": 1, "kind": kind,
            "sample": "forward+back of %d of the %d views at %dx%d (%.1f s fwd, %.1f s back), "
                      "single thread" % (len(sub), len(c_angles), img.shape[0], img.shape[0],
                                          t1 - t0, t2 - t1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--residual-target", type=float, default=1e-3)
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="after the timed steps, copy a 256 MiB float32 array 3 times "
                         "(a kernel of known traffic for scaling FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--max-residual-iters", type=int, default=0,
                    help="also time plain SIRT from zero to the residual target (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d "
              "requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from iterative_tomographic_reconstruction_algorithms_amd import autotune
    from iterative_tomographic_reconstruction_algorithms_amd.core import Plan
    from iterative_tomographic_reconstruction_algorithms_amd.phantom import angles_deg, ellipses

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    img = ellipses(N, ELLIPSES, seed=SEED)
    facade = (angles_deg(V) * np.pi / 180).astype(np.float32)
    c_angles = np.fft.fftshift(facade)
    plan = Plan("dd", N, c_angles)
    # launch shapes from the autotuner (timed here, not read from a cache)
    shapes = autotune.tune(plan, nslices=1, reps=2, use_cache=False)
    x_true = torch.from_numpy(img).to(dev)
    b = plan.forward(x_true)
    R, C = plan.sirt_weights()
    x = torch.zeros((N, N), dtype=torch.float32, device=dev)
    work = torch.empty((V, N), dtype=torch.float32, device=dev)
    stats = torch.zeros(3, dtype=torch.float64, device=dev)
    bnorm = float(torch.linalg.vector_norm(b.double()))

    for _ in range(args.warmup):
        plan.sirt_iterate(x, b, R, C, work, pc_beta=0.0, stats=stats)
    torch.cuda.synchronize()

    K = args.steps
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3 * K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        ev[3 * k].record()
        plan.sirt_forward(x, b, R, work, stats=stats)
        ev[3 * k + 1].record()
        plan.sirt_back(work, C, x, pc_beta=0.0, stats=stats)
        ev[3 * k + 2].record()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    fwd_ms = sum(ev[3 * k].elapsed_time(ev[3 * k + 1]) for k in range(K)) / K
    back_ms = sum(ev[3 * k + 1].elapsed_time(ev[3 * k + 2]) for k in range(K)) / K
    ms_per_step = secs * 1e3 / K
    gups = 2.0 * N * N * V * K / secs / 1e9
    resid_rel = float(stats[0].sqrt()) / bnorm

    if back_ms >= fwd_ms:
        dom, dom_ms, dom_bytes = "dd_back_kernel<FAST,SIRT_BACK,%d,%d>" % tuple(
            [(64, 4), (32, 2), (32, 4)][shapes["back"]]), back_ms, alg_bytes_back(N, V)
    else:
        dom, dom_ms, dom_bytes = "dd_fwd_kernel<FAST,SIRT_FWD,%d>" % (
            [256, 128, 64][shapes["fwd"]]), fwd_ms, alg_bytes_forward(N, V)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    line = {
        "metric": "SIRT Gupdates/s & s-to-residual 1e-3 (plain vs warm-start) @2048²×1800 views",
        "value": round(gups, 3),
        "unit": "Gupdates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 monitors)",
        "data": "synthetic",
        "config": {"workload": "plain SIRT, distance-driven pair, 2048x2048 image, 1800 views "
                               "over [0,180), 2048 bins, 40-ellipse phantom (seed 2), noiseless, "
                               "x0=0, non-negativity on", "npix": N, "nang": V, "slices": 1,
                   "shapes": {"fwd": shapes["fwd"], "back": shapes["back"],
                              "fwd_ms": shapes["fwd_ms"], "back_ms": shapes["back_ms"]}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "kernels_ms": {"sirt_forward": round(fwd_ms, 4), "sirt_back": round(back_ms, 4)},
        "residual_after_timed_steps": resid_rel,
    }

    if args.pmc_calibration:
        src = torch.rand(64 << 20, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        line["pmc_calibration"] = {"copy_bytes_read": src.numel() * 4,
                                   "copy_bytes_written": dst.numel() * 4}

    if args.max_residual_iters > 0:
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, rel = 0, 1.0
        while it < args.max_residual_iters:
            plan.sirt_iterate(x, b, R, C, work, pc_beta=0.0, stats=stats)
            it += 1
            if it % 5 == 0:
                rel = float(stats[0].sqrt()) / bnorm  # residual of the iterate entering step
                if rel <= args.residual_target:
                    break
        torch.cuda.synchronize()
        line["plain_sirt_to_residual"] = {"target": args.residual_target, "iterations": it,
                                          "seconds": round(time.perf_counter() - t0, 4),
                                          "reached": rel <= args.residual_target,
                                          "residual": rel}

    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(img, c_angles)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
