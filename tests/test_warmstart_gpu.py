"""The learned warm start gains over the method without it, on a small geometry.

Pairs (FBP input, SIRT-100 target) are made by the library from noisy random textured
phantoms (64 x 64, 60 views, I0 log-uniform in [1e3, 1e5]); a U-Net is trained on them
on the GPU; on held-out sinograms, plain SIRT is continued from zero, from FBP and from
U-Net(FBP) until it reaches the residual plain SIRT from zero has after 100 iterations.
The network must need fewer SIRT iterations than both starts without it."""
import numpy as np
import pytest
import torch

from iterative_tomographic_reconstruction_algorithms_amd import warmstart as W
from iterative_tomographic_reconstruction_algorithms_amd.unet import UNet

pytestmark = pytest.mark.gpu


def test_fbp_matches_facade():
    from iterative_tomographic_reconstruction_algorithms_amd.class_projectors_radon import (
        projectors)
    from iterative_tomographic_reconstruction_algorithms_amd.phantom import angles_deg
    plan = W.dd_plan(64, 60)
    tp = projectors(64, angles_deg(60), oper="dd")
    b = torch.rand((2, 60, 64), device="cuda")
    got = W.fbp(plan, b)
    for k in range(2):
        assert torch.equal(got[k], tp.fbp(b[k].contiguous()))


def test_unet_warm_start_needs_fewer_iterations():
    torch.manual_seed(0)
    plan = W.dd_plan(64, 60)
    xin, xout, _, _ = W.make_pairs(plan, 384, 100, seed=1, noise=(1e3, 1e5), n_ellipses=6)
    net = UNet(seed=0)
    losses = W.train(net, xin, xout, steps=400, batch=16, lr=2e-3)
    assert losses[-20:].mean() < 0.5 * losses[:20].mean()
    _, _, b, _ = W.make_pairs(plan, 16, 1, seed=2, noise=(1e3, 1e5), n_ellipses=6)
    r = W.compare_starts(plan, net, b, plain_iters=100, max_iters=400)
    print(r)
    assert r["iters_unet"] < r["iters_fbp"]
    assert r["iters_unet"] < r["iters_zero"]
    assert r["iters_zero"] <= 101
