"""Drop-in for the reference's ``radon_slant_stacking`` Cython module
(common/operators/radon_module/pymodule_radon_slant_stacking/radon_slant_stacking.pyx:15-73),
running the slant-stacking projector on the GPU.

``forwproj(image, angles, method)`` -> sinogram (nang x npix); ``backproj(sino, angles,
method)`` -> image (npix x npix).  As in the reference: when ``max(angles) > 2*pi`` the
caller's array is multiplied in place by ``2*pi/180`` (radon_slant_stacking.pyx:27-28);
no fftshift (commented out there, line 30); sinogram row nang-1-v holds view v
(radon_ss.c:111); ``method`` None means 0 (nearest pixel), any other nonzero value the
linear interpolation (radon_ss.c:123 tests ``method == 0`` only).  Lines that miss the
image square are sampled only inside the image (the reference indexes outside it).
"""
import numpy as np

from . import _dropin as D


def _c_angles(angles):
    a = D.degrees_to_radians_inplace(D.check_buffer(angles, 1, "angles"), 2 * np.pi / 180.0)
    return np.asarray(a, dtype=np.float32)


def _method(method):
    return 0 if method is None or int(method) == 0 else 1


def forwproj(image, angles, method):
    image = D.check_buffer(image, 2, "image")
    npix = image.shape[0]
    plan = D.plan_for("ss", npix, _c_angles(angles), method=_method(method))
    img, was_np = D.to_device(image)
    return D.from_device(plan.forward(img), was_np)


def backproj(sino, angles, method):
    sino = D.check_buffer(sino, 2, "sino")
    npix = sino.shape[1]
    plan = D.plan_for("ss", npix, _c_angles(angles), method=_method(method))
    if sino.shape[0] != plan.nang:
        raise ValueError("sinogram has %d rows but %d angles were given"
                         % (sino.shape[0], plan.nang))
    s, was_np = D.to_device(sino)
    return D.from_device(plan.back(s), was_np)
