"""Drop-in for the reference's ``radon_distance_driven`` Cython module
(common/operators/radon_module/pymodule_radon_distance_driven/radon_distance_driven.pyx:14-62),
running the distance-driven projector on the GPU.

``forwproj(image, angles)`` -> sinogram (nang x npix); ``backproj(sino, angles)`` ->
image (npix x npix).  As in the reference: ``angles`` in radians (multiplied in place by
pi/180 when their maximum exceeds 2*pi), ``np.fft.fftshift``-ed before the projector,
and sinogram row nang-1-v holds view v of the shifted list.
"""
import numpy as np

from . import _dropin as D


def _c_angles(angles):
    a = D.degrees_to_radians_inplace(D.check_buffer(angles, 1, "angles"), np.pi / 180.0)
    return np.fft.fftshift(np.asarray(a, dtype=np.float32))


def forwproj(image, angles):
    image = D.check_buffer(image, 2, "image")
    npix = image.shape[0]
    plan = D.plan_for("dd", npix, _c_angles(angles))
    img, was_np = D.to_device(image)
    return D.from_device(plan.forward(img), was_np)


def backproj(sino, angles):
    sino = D.check_buffer(sino, 2, "sino")
    npix = sino.shape[1]
    plan = D.plan_for("dd", npix, _c_angles(angles))
    if sino.shape[0] != plan.nang:
        raise ValueError("sinogram has %d rows but %d angles were given"
                         % (sino.shape[0], plan.nang))
    s, was_np = D.to_device(sino)
    return D.from_device(plan.back(s), was_np)
