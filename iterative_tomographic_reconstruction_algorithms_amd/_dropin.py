"""Shared plumbing of the drop-in projector modules.

The reference's Cython wrappers (radon_distance_driven.pyx, radon_ray_driven.pyx,
radon_pixel_driven.pyx) take float32 C-contiguous numpy arrays; a wrong dtype, rank or layout raises
``ValueError`` from the buffer check and ``None`` raises ``TypeError``.  The drop-ins
keep that contract for numpy input (the result comes back as numpy) and additionally
accept float32 GPU tensors, which stay on the GPU.
"""
import collections

import numpy as np
import torch

from .core import Plan, device

_PLANS = collections.OrderedDict()
_MAX_PLANS = 16


def plan_for(oper, npix, c_angles, arith="fast", method=1):
    """Cached plan for (projector, npix, angles as the C receives them[, method])."""
    a = np.ascontiguousarray(c_angles, dtype=np.float32)
    key = (oper, int(npix), arith, int(method), a.tobytes())
    p = _PLANS.get(key)
    if p is None:
        p = Plan(oper, npix, a, arith, method)
        _PLANS[key] = p
        if len(_PLANS) > _MAX_PLANS:
            _PLANS.popitem(last=False)[1].close()
    else:
        _PLANS.move_to_end(key)
    return p


def check_buffer(arr, ndim, name):
    """Cython `np.ndarray[float, ndim=k, mode="c"] not None` semantics."""
    if arr is None:
        raise TypeError("Argument '%s' must not be None" % name)
    if isinstance(arr, torch.Tensor):
        if arr.dim() != ndim:
            raise ValueError("Buffer has wrong number of dimensions (expected %d, got %d)"
                             % (ndim, arr.dim()))
        if arr.dtype != torch.float32:
            raise ValueError("Buffer dtype mismatch, expected 'float' but got %s" % arr.dtype)
        if not arr.is_contiguous():
            raise ValueError("ndarray is not C-contiguous")
        return arr
    if not isinstance(arr, np.ndarray):
        raise TypeError("Argument '%s' has incorrect type (expected numpy.ndarray, got %s)"
                        % (name, type(arr).__name__))
    if arr.ndim != ndim:
        raise ValueError("Buffer has wrong number of dimensions (expected %d, got %d)"
                         % (ndim, arr.ndim))
    if arr.dtype != np.float32:
        raise ValueError("Buffer dtype mismatch, expected 'float' but got '%s'" % arr.dtype)
    if not arr.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")
    return arr


def degrees_to_radians_inplace(angles, factor):
    """`if np.max(angles) > 2*np.pi: angles *= factor` -- in place on the caller's
    array, as the reference wrappers do (radon_distance_driven.pyx:25-26)."""
    if isinstance(angles, torch.Tensor):
        if angles.numel() and float(angles.max()) > 2 * np.pi:
            angles.mul_(factor)
        return angles.detach().cpu().numpy()
    if angles.size and np.max(angles) > 2 * np.pi:
        angles *= factor
    return angles


def to_device(arr):
    if isinstance(arr, torch.Tensor):
        return arr, False
    return torch.from_numpy(arr).to(device()), True


def from_device(t, was_numpy):
    if was_numpy:
        return t.cpu().numpy()
    return t
