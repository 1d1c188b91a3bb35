"""Batched SIRT over several geometries in one persistent launch, and hipGraph replay of
SIRT iterations (include/itomo.h: itomo_batch_*, itomo_graph_*).

The reference reconstructs a stack slice by slice through a multiprocessing.Pool that is
rebuilt every iteration (art.py:289-295); here slices of any mix of image sizes and view
counts run together on the GPU, and a whole iteration (or several) replays as one graph.
"""
import ctypes

import torch

from . import _lib as L
from .core import Plan, _check_dev, _p, device, stream_handle


class Member:
    """One geometry of a batch: a plan and its stacked slices, all on the GPU.
    ``b`` holds the slices' sinograms, shape (V, N) or (nz, V, N) of the plan."""

    def __init__(self, plan, b):
        # the persistent kernels read and write nang*npix entries per slice of b
        _check_dev(b, "b")
        if b.dim() not in (2, 3) or tuple(b.shape[-2:]) != (plan.nang, plan.npix):
            raise ValueError("b has shape %s; a member of this plan needs (%d, %d) or "
                             "(nz, %d, %d)" % (tuple(b.shape), plan.nang, plan.npix,
                                               plan.nang, plan.npix))
        if plan.oper != "dd" or plan.arith != "fast":
            raise ValueError("batch members are distance-driven plans with the reference "
                             "build's arithmetic (got %s/%s)" % (plan.oper, plan.arith))
        self.plan = plan
        self.b = b if b.dim() == 3 else b.unsqueeze(0)
        nz = self.b.shape[0]
        dev = self.b.device
        self.x = torch.zeros((nz, plan.npix, plan.npix), dtype=torch.float32, device=dev)
        self.R, self.C = plan.sirt_weights()
        self.work = torch.empty_like(self.b)
        self.stats = torch.zeros(3 * nz, dtype=torch.float64, device=dev)

    @property
    def nslices(self):
        return self.b.shape[0]


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


class Batch:
    """All members advance together with :meth:`iterate` (one persistent forward and one
    persistent back launch per iteration)."""

    def __init__(self, members, pc_beta=0.0):
        self.members = list(members)
        self.pc_beta = float(pc_beta)
        n = len(self.members)
        plans = (ctypes.c_void_p * n)(*[m.plan._h.value for m in self.members])
        nsl = (ctypes.c_int * n)(*[m.nslices for m in self.members])
        h = ctypes.c_void_p()
        L.check(L.lib().itomo_batch_create(
            ctypes.byref(h), n, ctypes.cast(plans, ctypes.c_void_p),
            ctypes.cast(nsl, ctypes.c_void_p),
            ctypes.cast(_ptrs([m.x for m in self.members]), ctypes.c_void_p),
            ctypes.cast(_ptrs([m.b for m in self.members]), ctypes.c_void_p),
            ctypes.cast(_ptrs([m.R for m in self.members]), ctypes.c_void_p),
            ctypes.cast(_ptrs([m.C for m in self.members]), ctypes.c_void_p),
            ctypes.cast(_ptrs([m.work for m in self.members]), ctypes.c_void_p),
            ctypes.cast(_ptrs([m.stats for m in self.members]), ctypes.c_void_p)),
            "batch_create")
        self._h = h

    def iterate(self):
        L.check(L.lib().itomo_batch_sirt_iterate(self._h, self.pc_beta, stream_handle()),
                "batch_sirt_iterate")

    def graph(self, iters=1):
        return Graph.batch(self, iters)

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and h.value:
            torch.cuda.current_stream().synchronize()
            L.lib().itomo_batch_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured sequence of SIRT iterations; :meth:`launch` replays it."""

    def __init__(self, h, keep):
        self._h = h
        self._keep = keep   # arrays the graph reads and writes must outlive it

    @classmethod
    def sirt(cls, plan, x, b, R, C, work, pc_beta=-1.0, stats=None, iters=1):
        nz = x.numel() // plan.image_numel
        h = ctypes.c_void_p()
        L.check(L.lib().itomo_graph_sirt(ctypes.byref(h), plan._h, _p(x), _p(b), _p(R), _p(C),
                                         _p(work), nz, float(pc_beta), _p(stats), int(iters),
                                         stream_handle()), "graph_sirt")
        return cls(h, (plan, x, b, R, C, work, stats))

    @classmethod
    def batch(cls, batch, iters=1):
        h = ctypes.c_void_p()
        L.check(L.lib().itomo_graph_batch_sirt(ctypes.byref(h), batch._h, batch.pc_beta,
                                               int(iters), stream_handle()), "graph_batch")
        return cls(h, (batch,))

    def launch(self):
        L.check(L.lib().itomo_graph_launch(self._h, stream_handle()), "graph_launch")

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and h.value:
            torch.cuda.current_stream().synchronize()
            L.lib().itomo_graph_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
