"""Launch-shape autotuner for the distance-driven projector pair.

The forward kernel can run 256, 128 or 64 bins of a view per block and the back kernel
64x64 tiles (4x4 pixels per thread), 32x32 tiles (2x2) or 32x32 tiles (4x4).  Which is
fastest depends on the image size, the view count and how many slices share a launch
(a 256^2 single slice has only 16 of the large back tiles -- far too few blocks for 256
compute units).  Every shape accumulates each output element in the same order, so the
choice never changes a projection or an iterate.

:func:`tune` times each forward and each back shape of a plan on the GPU with HIP
events on the current stream, keeps the fastest and records it in a JSON cache keyed by
(device, projector, arithmetic, npix, nang, nslices); the bucket widths make one entry
serve nearby geometries.  The cache lives at ``$ITOMO_SHAPE_CACHE`` or
``~/.cache/itomo/shapes.json``.
"""
import json
import os
import tempfile

from . import _lib as L

__all__ = ["tune", "bucket_key", "load_cache", "store_entry", "cache_path", "fastest"]


def cache_path(path=None):
    if path:
        return path
    env = os.environ.get("ITOMO_SHAPE_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "itomo", "shapes.json")


def _pow2_ceil(n):
    k = 1
    while k < n:
        k <<= 1
    return k


def bucket_key(device_name, oper, arith, npix, nang, nslices):
    """Cache key: npix rounded up to a multiple of 64, nang and nslices to powers of 2."""
    nb = ((int(npix) + 63) // 64) * 64
    return "%s|%s|%s|n%d|v%d|z%d" % (device_name, oper, arith, nb, _pow2_ceil(int(nang)),
                                     _pow2_ceil(int(nslices)))


def load_cache(path=None):
    p = cache_path(path)
    try:
        with open(p) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_entry(key, entry, path=None):
    """Add one entry, writing the file atomically (concurrent tuners may race; the last
    complete write wins and the file is never left half-written)."""
    p = cache_path(path)
    data = load_cache(p)
    data[key] = entry
    d = os.path.dirname(p) or "."
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d, prefix=".shapes.", suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def fastest(times):
    """Index of the smallest time (first one on ties)."""
    best = 0
    for k, t in enumerate(times):
        if t < times[best]:
            best = k
    return best


def _time(fn, reps):
    import torch
    fn()  # warm-up (first launch, scratch growth)
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / reps


def tune(plan, nslices=1, reps=3, use_cache=True, path=None):
    """Pick the fastest launch shapes for ``plan`` running ``nslices`` slices per call and
    set them on the plan.  Returns ``{"fwd": f, "back": b, "fwd_ms": [...],
    "back_ms": [...], "cached": bool}``."""
    import torch
    from .core import device

    if plan.oper != "dd":
        return {"fwd": 0, "back": 0, "fwd_ms": [], "back_ms": [], "cached": False}
    dev = device()
    key = bucket_key(torch.cuda.get_device_name(dev), plan.oper, plan.arith, plan.npix,
                     plan.nang, nslices)
    if use_cache:
        hit = load_cache(path).get(key)
        if (isinstance(hit, dict) and 0 <= hit.get("fwd", -1) < L.FWD_VARIANTS
                and 0 <= hit.get("back", -1) < L.BACK_VARIANTS):
            plan.set_shapes(hit["fwd"], hit["back"])
            return dict(hit, cached=True)

    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.rand((nslices, plan.npix, plan.npix), device=dev, generator=g)
    y = torch.rand((nslices, plan.nang, plan.npix), device=dev, generator=g)
    sino = torch.empty_like(y)
    img = torch.empty_like(x)
    fwd_ms, back_ms = [], []
    for f in range(L.FWD_VARIANTS):
        plan.set_shapes(f, 0)
        fwd_ms.append(_time(lambda: plan.forward(x, out=sino), reps))
    for b in range(L.BACK_VARIANTS):
        plan.set_shapes(0, b)
        back_ms.append(_time(lambda: plan.back(y, out=img), reps))
    entry = {"fwd": fastest(fwd_ms), "back": fastest(back_ms),
             "fwd_ms": [round(t, 4) for t in fwd_ms], "back_ms": [round(t, 4) for t in back_ms]}
    plan.set_shapes(entry["fwd"], entry["back"])
    if use_cache:
        store_entry(key, entry, path)
    return dict(entry, cached=False)
