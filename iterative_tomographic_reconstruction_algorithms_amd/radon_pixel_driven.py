"""Drop-in for the reference's ``radon_pixel_driven`` Cython module
(common/operators/radon_module/pymodule_radon_pixel_driven/radon_pixel_driven.pyx:15-73),
running the pixel-driven projector on the GPU.

``forwproj(image, angles, method)`` -> sinogram (nang x npix); ``backproj(sino, angles,
method)`` -> image (npix x npix).  As in the reference: when ``max(angles) > 2*pi`` the
caller's array is multiplied in place by ``2*pi/180`` (sic, radon_pixel_driven.pyx:27-28,
twice the degree factor); the angles are ``np.fft.fftshift``-ed before the projector and
sinogram row v holds view v of the shifted list; ``method`` None means 0 (nearest bin),
any other nonzero value the linear split (radon_pd.c:57 tests ``method == 0`` only).
"""
import numpy as np

from . import _dropin as D


def _c_angles(angles):
    a = D.degrees_to_radians_inplace(D.check_buffer(angles, 1, "angles"), 2 * np.pi / 180.0)
    return np.fft.fftshift(np.asarray(a, dtype=np.float32))


def _method(method):
    return 0 if method is None or int(method) == 0 else 1


def forwproj(image, angles, method):
    image = D.check_buffer(image, 2, "image")
    npix = image.shape[0]
    plan = D.plan_for("pd", npix, _c_angles(angles), method=_method(method))
    img, was_np = D.to_device(image)
    return D.from_device(plan.forward(img), was_np)


def backproj(sino, angles, method):
    sino = D.check_buffer(sino, 2, "sino")
    npix = sino.shape[1]
    plan = D.plan_for("pd", npix, _c_angles(angles), method=_method(method))
    if sino.shape[0] != plan.nang:
        raise ValueError("sinogram has %d rows but %d angles were given"
                         % (sino.shape[0], plan.nang))
    s, was_np = D.to_device(sino)
    return D.from_device(plan.back(s), was_np)
