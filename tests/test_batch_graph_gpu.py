"""Batched persistent SIRT over mixed geometries and hipGraph replay: both must equal the
per-plan eager path bit for bit (the work schedule never changes an output element's
accumulation order); the FP64 monitors agree to 1e-12 (their partial sums are grouped
differently).  Graph replays are interleaved with eager work and every replay is
checked against the eager result."""
import numpy as np
import pytest
import torch

from iterative_tomographic_reconstruction_algorithms_amd.batch import Batch, Graph, Member
from iterative_tomographic_reconstruction_algorithms_amd.core import Plan
from iterative_tomographic_reconstruction_algorithms_amd.phantom import angles_deg, ellipses

pytestmark = pytest.mark.gpu


def dd_plan(n, v):
    return Plan("dd", n, np.fft.fftshift((angles_deg(v) * np.pi / 180).astype(np.float32)))


def sinos(plan, nz, seed):
    imgs = np.stack([ellipses(plan.npix, 8, seed=seed + k) for k in range(nz)])
    return plan.forward(torch.from_numpy(imgs).cuda())


def eager_sirt(plan, b, iters, pc=0.0):
    R, C = plan.sirt_weights()
    x = torch.zeros((b.shape[0], plan.npix, plan.npix), dtype=torch.float32, device="cuda")
    w = torch.empty_like(b)
    st = torch.zeros(3 * b.shape[0], dtype=torch.float64, device="cuda")
    for _ in range(iters):
        plan.sirt_iterate(x, b, R, C, w, pc_beta=pc, stats=st)
    return x, st


def test_batch_of_mixed_geometries_equals_per_plan():
    geoms = [(64, 45, 2), (96, 30, 1), (128, 90, 3), (48, 180, 2)]
    members = []
    for k, (n, v, nz) in enumerate(geoms):
        p = dd_plan(n, v)
        members.append(Member(p, sinos(p, nz, 10 * k)))
    batch = Batch(members, pc_beta=0.0)
    for _ in range(4):
        batch.iterate()
    for m in members:
        x_ref, st_ref = eager_sirt(m.plan, m.b, 4)
        assert torch.equal(m.x, x_ref)
        # FP64 monitors: the batch reduces the residual per wave inside the sweep, the
        # per-plan path per block of its epilogue pass; same terms, other grouping
        torch.testing.assert_close(m.stats, st_ref, rtol=1e-12, atol=0)


def test_graph_replay_equals_eager():
    p = dd_plan(96, 60)
    b = sinos(p, 2, 5)
    R, C = p.sirt_weights()
    x = torch.zeros((2, 96, 96), dtype=torch.float32, device="cuda")
    w = torch.empty_like(b)
    st = torch.zeros(6, dtype=torch.float64, device="cuda")
    g = Graph.sirt(p, x, b, R, C, w, pc_beta=0.0, stats=st, iters=2)
    for rep in range(1, 4):
        g.launch()
        # eager work between replays, then the check against an eager run of 2*rep steps
        junk = torch.ones(1000, device="cuda").sum()
        torch.cuda.synchronize()
        x_ref, st_ref = eager_sirt(p, b, 2 * rep)
        assert torch.equal(x, x_ref), "replay %d" % rep
        assert torch.equal(st, st_ref), "replay %d" % rep
        assert float(junk) == 1000.0
    g.close()


def test_graph_of_batch_equals_eager():
    members = [Member(dd_plan(64, 40), sinos(dd_plan(64, 40), 2, 1)),
               Member(dd_plan(80, 72), sinos(dd_plan(80, 72), 1, 2))]
    batch = Batch(members, pc_beta=0.0)
    g = batch.graph(iters=3)
    for rep in range(1, 3):
        g.launch()
        torch.cuda.synchronize()
        for m in members:
            x_ref, _ = eager_sirt(m.plan, m.b, 3 * rep)
            assert torch.equal(m.x, x_ref), "replay %d" % rep
    g.close()


def test_graph_keeps_its_scratch_when_the_plan_buffer_grows():
    # The graph replays into its own scratch: calls that make the plan grow its buffer
    # (more slices) or lay it out for other launch shapes between replays must not
    # change what a replay computes.
    p = dd_plan(96, 60)
    b = sinos(p, 1, 7)
    R, C = p.sirt_weights()
    x = torch.zeros((1, 96, 96), dtype=torch.float32, device="cuda")
    w = torch.empty_like(b)
    st = torch.zeros(3, dtype=torch.float64, device="cuda")
    g = Graph.sirt(p, x, b, R, C, w, pc_beta=0.0, stats=st, iters=2)
    g.launch()
    big = sinos(p, 6, 11)                       # plan scratch grows for 6 slices
    xb = torch.zeros((6, 96, 96), dtype=torch.float32, device="cuda")
    wb = torch.empty_like(big)
    p.sirt_iterate(xb, big, R, C, wb, pc_beta=0.0,
                   stats=torch.zeros(18, dtype=torch.float64, device="cuda"))
    fwd0, back0 = p.shapes
    p.set_shapes((fwd0 + 1) % 3, (back0 + 1) % 3)   # another partial-sum layout
    p.forward(xb)
    g.launch()
    torch.cuda.synchronize()
    p.set_shapes(fwd0, back0)
    x_ref, st_ref = eager_sirt(p, b, 4)
    assert torch.equal(x, x_ref)
    assert torch.equal(st, st_ref)
    g.close()
