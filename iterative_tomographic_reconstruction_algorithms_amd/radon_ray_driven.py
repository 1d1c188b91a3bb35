"""Drop-in for the reference's ``radon_ray_driven`` Cython module
(common/operators/radon_module/pymodule_radon_ray_driven/radon_ray_driven.pyx:15-61),
running the ray-driven projector on the GPU.

``forwproj(image, angles)`` -> sinogram (nang x npix); ``backproj(sino, angles)`` ->
image.  Angles in radians (multiplied in place by pi/180 when their maximum exceeds
2*pi); no fftshift; sinogram row v holds view v.
"""
import numpy as np

from . import _dropin as D


def _c_angles(angles):
    a = D.degrees_to_radians_inplace(D.check_buffer(angles, 1, "angles"), np.pi / 180.0)
    return np.asarray(a, dtype=np.float32)


def forwproj(image, angles):
    image = D.check_buffer(image, 2, "image")
    npix = image.shape[0]
    plan = D.plan_for("rd", npix, _c_angles(angles))
    img, was_np = D.to_device(image)
    return D.from_device(plan.forward(img), was_np)


def backproj(sino, angles):
    sino = D.check_buffer(sino, 2, "sino")
    npix = sino.shape[1]
    plan = D.plan_for("rd", npix, _c_angles(angles))
    if sino.shape[0] != plan.nang:
        raise ValueError("sinogram has %d rows but %d angles were given"
                         % (sino.shape[0], plan.nang))
    s, was_np = D.to_device(sino)
    return D.from_device(plan.back(s), was_np)
