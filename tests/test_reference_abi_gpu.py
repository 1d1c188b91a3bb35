"""The reference-signature entry points itomo_radon_dd / _rd / _pd / _ss, bound with
ctypes the way the replacement of the Cython modules binds them: a CDLL of libitomo.so,
the argument list of radon_dd.c:54 / radon_rd.c:76 (radon_pd.c:20 / radon_ss.c:93 add
`method`) plus a stream, device pointers, angles on the host as the C receives them
(after the wrapper's degree conversion and fftshift, radon_distance_driven.pyx:25-27,
radon_pixel_driven.pyx:27-30, radon_slant_stacking.pyx:27-30).

Bars:
  * into zeroed outputs, forward and back are bitwise equal to the reference's own C
    built with its flags (the committed golden vectors), on every golden case;
  * into non-zero outputs they accumulate like the C, which adds every term to whatever
    the output holds (radon_dd.c:79-96): bitwise equal to the CPU restatement called the
    same way on the same non-zero array.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

from oracle import radon as O
from iterative_tomographic_reconstruction_algorithms_amd import _lib

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
Z = {p: np.load(os.path.join(GOLD, f)) for p, f in
     (("dd", "radon_ref.npz"), ("pd", "radon_pd_ref.npz"), ("ss", "radon_ss_ref.npz"))}
Z["rd"] = Z["dd"]


def keys(proj):
    return sorted({k.split("/")[0] for k in Z[proj].files if k.startswith(proj + "_")})


# --- the binding, as a maintainer would write it in place of the .pyx --------------------
_BOUND = []


def binding():
    if not _BOUND:
        L = ctypes.CDLL(_lib.LIB_PATH)
        vp, i = ctypes.c_void_p, ctypes.c_int
        for name in ("itomo_radon_dd", "itomo_radon_rd"):
            getattr(L, name).argtypes = [vp, i, vp, i, i, vp, vp]
            getattr(L, name).restype = i
        for name in ("itomo_radon_pd", "itomo_radon_ss"):
            getattr(L, name).argtypes = [vp, i, vp, i, i, i, vp, vp]
            getattr(L, name).restype = i
        L.itomo_last_error.restype = ctypes.c_char_p
        _BOUND.append(L)
    return _BOUND[0]


def c_angles(proj, angles):
    """The angles as the C receives them from its wrapper."""
    a = np.array(angles, np.float32).reshape(-1)
    if a.size and a.max() > 2 * np.pi:
        # radon_distance_driven.pyx:25-26, radon_ray_driven.pyx; pd/ss use 2*pi/180
        a *= np.float32((2 * np.pi if proj in ("pd", "ss") else np.pi) / 180.0)
    if proj in ("dd", "pd"):
        a = np.fft.fftshift(a)
    return np.ascontiguousarray(a, np.float32)


def radon(proj, image, sino, angles_c, oper, method=1):
    """One call of the reference-signature entry point on device tensors (in place)."""
    npix = image.shape[-1]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    a = angles_c.ctypes.data_as(ctypes.c_void_p)
    args = [ctypes.c_void_p(image.data_ptr()), npix, a, angles_c.size, oper]
    if proj in ("pd", "ss"):
        args.append(method)
    args += [ctypes.c_void_p(sino.data_ptr()), stream]
    L = binding()
    rc = getattr(L, "itomo_radon_" + proj)(*args)
    assert rc == 0, L.itomo_last_error()
    torch.cuda.synchronize()


def gpu(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()


def bits_equal(a, b, mask=None):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape
    diff = a.view(np.int32) != b.view(np.int32)
    if mask is not None:
        diff &= ~mask
    assert int(diff.sum()) == 0, "%d of %d entries differ" % (int(diff.sum()), a.size)


CASES = ([("dd", k) for k in keys("dd")] + [("rd", k) for k in keys("rd")]
         + [("pd", k) for k in keys("pd")] + [("ss", k) for k in keys("ss")])


@pytest.mark.parametrize("proj,key", CASES)
def test_reference_signature_matches_reference_build(proj, key):
    z = Z[proj]
    method = int(key.rsplit("_m", 1)[1]) if proj in ("pd", "ss") else 1
    img, sino, ang = z[key + "/image"], z[key + "/sino"], z[key + "/angles"]
    a = c_angles(proj, ang)
    n, v = img.shape[0], a.size
    s_out = torch.zeros((v, n), dtype=torch.float32, device="cuda")
    radon(proj, gpu(img), s_out, a, 0, method)
    edge = z[key + "/edge"] if proj == "ss" else None
    bits_equal(s_out.cpu().numpy(), z[key + "/fwd_fast"], edge)
    i_out = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    radon(proj, i_out, gpu(sino), a, 1, method)
    bits_equal(i_out.cpu().numpy(), z[key + "/back_fast"])


@pytest.mark.parametrize("proj", ["dd", "rd", "pd", "ss"])
def test_reference_signature_accumulates_into_nonzero_outputs(proj):
    rng = np.random.default_rng({"dd": 1, "rd": 2, "pd": 3, "ss": 4}[proj])
    n, v = 48, 37
    facade = (np.arange(v) * 180.0 / v * np.pi / 180).astype(np.float32)
    a = c_angles(proj, facade)
    img = (rng.random((n, n)) - 0.2).astype(np.float32)
    sino = rng.random((v, n)).astype(np.float32)
    s0 = rng.random((v, n)).astype(np.float32)        # non-zero outputs
    i0 = rng.random((n, n)).astype(np.float32)
    s_out, i_out = gpu(s0), gpu(i0)
    radon(proj, gpu(img), s_out, a, 0)
    radon(proj, i_out, gpu(sino), a, 1)
    if proj in ("dd", "rd"):
        ref_s, ref_i = s0.copy(), i0.copy()
        fn = getattr(O.lib(), "oracle_radon_" + proj)
        fn(img.copy(), n, a, v, 0, ref_s, O.FAST)
        fn(ref_i, n, a, v, 1, sino.copy(), O.FAST)
    else:
        fwd, back = (O.pd_forward, O.pd_back) if proj == "pd" else (O.ss_forward, O.ss_back)
        ref_s = fwd(img, facade, O.FAST, 1, init=s0)
        ref_i = back(sino, facade, O.FAST, 1, init=i0)
    bits_equal(s_out.cpu().numpy(), ref_s)
    bits_equal(i_out.cpu().numpy(), ref_i)
