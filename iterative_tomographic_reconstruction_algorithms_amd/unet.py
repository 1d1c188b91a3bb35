"""Learned warm start for SIRT: a 4-level U-Net run and trained on the library's HIP
kernels (BF16 on the matrix cores, FP32 accumulation; see include/itomo.h).

Architecture (base width 32, 3x3 convolutions, ReLU, 2x2 max pooling, nearest
upsampling with skip concatenation)::

    e0: 1 -> 32 -> 32        (H)          d0: 64+32 -> 32 -> 32   (H)
    e1: 32 -> 64 -> 64       (H/2)        d1: 128+64 -> 64 -> 64  (H/2)
    e2: 64 -> 128 -> 128     (H/4)        d2: 256+128 -> 128 -> 128 (H/4)
    e3: 128 -> 256 -> 256    (H/8, bottleneck)
    out = image + head(d0),  head = 1x1 convolution 32 -> 1 (FP32)

The head starts at zero, so an untrained network returns its input (the FBP image) and
training learns a correction towards the SIRT iterate.  Images are (B, H, W) float32
GPU tensors with H and W multiples of 8.  Nothing here runs on the CPU: every layer is
a libitomo kernel on torch's current stream, torch only allocates.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from .core import device, stream_handle

__all__ = ["UNet", "CONVS"]

# name: (input channels, output channels)
CONVS = [
    ("e0a", 1, 32), ("e0b", 32, 32),
    ("e1a", 32, 64), ("e1b", 64, 64),
    ("e2a", 64, 128), ("e2b", 128, 128),
    ("e3a", 128, 256), ("e3b", 256, 256),
    ("d2a", 384, 128), ("d2b", 128, 128),
    ("d1a", 192, 64), ("d1b", 64, 64),
    ("d0a", 96, 32), ("d0b", 32, 32),
]
HEAD_C = 32


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _ck(rc, what):
    L.check(rc, what)


class UNet:
    """Parameters (FP32 masters + BF16 copies), gradients and Adam state on the GPU."""

    def __init__(self, seed=0):
        dev = device()
        g = torch.Generator().manual_seed(seed)
        self.shapes = {name: (cin, cout) for name, cin, cout in CONVS}
        self.w, self.b = {}, {}
        for name, cin, cout in CONVS:
            std = (2.0 / (9 * cin)) ** 0.5           # He initialisation for ReLU
            self.w[name] = (torch.randn((cout, 9, cin), generator=g) * std).to(dev)
            self.b[name] = torch.zeros(cout, device=dev)
        self.hw = torch.zeros(HEAD_C, device=dev)
        self.hb = torch.zeros(1, device=dev)
        self.wbf = {k: torch.empty(v.shape, dtype=torch.bfloat16, device=dev)
                    for k, v in self.w.items()}
        self.refresh_bf16()
        self._zero_like_state()
        self.step = 0
        self._mse_work = torch.empty(L.UNET_MSE_WORK, dtype=torch.float64, device=dev)

    # -- parameter bookkeeping ------------------------------------------------------
    def params(self):
        """(name, fp32 master, bf16 copy or None) for every parameter tensor."""
        out = []
        for name, _, _ in CONVS:
            out.append(("w." + name, self.w[name], self.wbf[name]))
            out.append(("b." + name, self.b[name], None))
        out.append(("hw", self.hw, None))
        out.append(("hb", self.hb, None))
        return out

    def _zero_like_state(self):
        self.grad = {n: torch.zeros_like(t) for n, t, _ in self.params()}
        self.m = {n: torch.zeros_like(t) for n, t, _ in self.params()}
        self.v = {n: torch.zeros_like(t) for n, t, _ in self.params()}

    def refresh_bf16(self):
        for name, t in self.w.items():
            _ck(L.lib().itomo_unet_to_bf16(_p(t), t.numel(), _p(self.wbf[name]),
                                           stream_handle()), "unet_to_bf16")

    def state_dict(self):
        return {n: t.detach().cpu().clone() for n, t, _ in self.params()}

    def load_state_dict(self, sd):
        for n, t, _ in self.params():
            t.copy_(sd[n].to(t.device, torch.float32).reshape(t.shape))
        self.refresh_bf16()

    def num_parameters(self):
        return sum(t.numel() for _, t, _ in self.params())

    # -- layers ---------------------------------------------------------------------
    @staticmethod
    def _empty(shape, dtype=torch.bfloat16):
        return torch.empty(shape, dtype=dtype, device=device())

    def _conv(self, x, name, relu=True, mask=None, wt=None, bias=True, cout=None):
        n, h, w, cin = x.shape
        if wt is None:
            wt = self.wbf[name]
            cout = self.shapes[name][1]
        y = self._empty((n, h, w, cout))
        _ck(L.lib().itomo_unet_conv3x3(_p(x), n, h, w, cin, _p(wt),
                                       _p(self.b[name]) if bias else None, cout, _p(mask),
                                       int(relu), _p(y), stream_handle()), "unet_conv3x3")
        return y

    def _pool(self, x):
        n, h, w, c = x.shape
        y = self._empty((n, h // 2, w // 2, c))
        idx = torch.empty((n, h // 2, w // 2, c), dtype=torch.uint8, device=x.device)
        _ck(L.lib().itomo_unet_pool_fwd(_p(x), n, h, w, c, _p(y), _p(idx), stream_handle()),
            "unet_pool_fwd")
        return y, idx

    def _upcat(self, a, b):
        n, h, w, cb = b.shape
        ca = a.shape[-1]
        y = self._empty((n, h, w, ca + cb))
        _ck(L.lib().itomo_unet_upcat_fwd(_p(a), _p(b), n, h, w, ca, cb, _p(y),
                                         stream_handle()), "unet_upcat_fwd")
        return y

    # -- forward / backward ---------------------------------------------------------
    def forward(self, img, keep=False):
        """img: (B, H, W) float32 GPU tensor, H and W multiples of 8 -> (B, H, W)."""
        if img.dim() == 2:
            return self.forward(img[None], keep)[0]
        if img.dtype != torch.float32 or not img.is_cuda or not img.is_contiguous():
            raise ValueError("img must be a C-contiguous float32 GPU tensor")
        n, h, w = img.shape
        if h % 8 or w % 8:
            raise ValueError("image sides must be multiples of 8 (got %dx%d)" % (h, w))
        x0 = self._empty((n, h, w, 1))
        _ck(L.lib().itomo_unet_to_bf16(_p(img), img.numel(), _p(x0), stream_handle()),
            "unet_to_bf16")
        s = {"x0": x0, "img": img}
        s["a_e0"] = self._conv(x0, "e0a")
        s["e0"] = self._conv(s["a_e0"], "e0b")
        s["p0"], s["i0"] = self._pool(s["e0"])
        s["a_e1"] = self._conv(s["p0"], "e1a")
        s["e1"] = self._conv(s["a_e1"], "e1b")
        s["p1"], s["i1"] = self._pool(s["e1"])
        s["a_e2"] = self._conv(s["p1"], "e2a")
        s["e2"] = self._conv(s["a_e2"], "e2b")
        s["p2"], s["i2"] = self._pool(s["e2"])
        s["a_e3"] = self._conv(s["p2"], "e3a")
        s["e3"] = self._conv(s["a_e3"], "e3b")
        s["u2"] = self._upcat(s["e3"], s["e2"])
        s["a_d2"] = self._conv(s["u2"], "d2a")
        s["d2"] = self._conv(s["a_d2"], "d2b")
        s["u1"] = self._upcat(s["d2"], s["e1"])
        s["a_d1"] = self._conv(s["u1"], "d1a")
        s["d1"] = self._conv(s["a_d1"], "d1b")
        s["u0"] = self._upcat(s["d1"], s["e0"])
        s["a_d0"] = self._conv(s["u0"], "d0a")
        s["d0"] = self._conv(s["a_d0"], "d0b")
        out = torch.empty_like(img)
        _ck(L.lib().itomo_unet_head_fwd(_p(s["d0"]), _p(self.hw), _p(self.hb), _p(img),
                                        img.numel(), HEAD_C, _p(out), stream_handle()),
            "unet_head_fwd")
        self._saved = s if keep else None
        return out

    def _conv_back(self, name, x_in, dpre, mask_in):
        """Layer `name` with input x_in and output gradient dpre (already multiplied by
        its ReLU mask): accumulate dW, db; return the input gradient multiplied by
        (mask_in > 0), or None when mask_in is None (first layer)."""
        n, h, w, cin = x_in.shape
        cout = self.shapes[name][1]
        _ck(L.lib().itomo_unet_wgrad(_p(x_in), _p(dpre), n, h, w, cin, cout,
                                     _p(self.grad["w." + name]), _p(self.grad["b." + name]),
                                     stream_handle()), "unet_wgrad")
        if mask_in is None:
            return None
        wf = self._empty((cin, 9, cout))
        _ck(L.lib().itomo_unet_flip(_p(self.wbf[name]), cin, cout, _p(wf), stream_handle()),
            "unet_flip")
        return self._conv(dpre, name, relu=False, mask=mask_in, wt=wf, bias=False, cout=cin)

    def _upcat_back(self, du, a, cb):
        n, h, w, c = du.shape
        ca = a.shape[-1]
        da = self._empty(a.shape)
        db = self._empty((n, h, w, cb))
        _ck(L.lib().itomo_unet_upcat_bwd(_p(du), _p(a), n, h, w, ca, cb, _p(da), _p(db),
                                         stream_handle()), "unet_upcat_bwd")
        return da, db

    def _pool_back(self, dp, idx, skip, e):
        n, h, w, c = e.shape
        de = self._empty(e.shape)
        _ck(L.lib().itomo_unet_pool_bwd(_p(dp), _p(idx), _p(skip), _p(e), n, h, w, c, _p(de),
                                        stream_handle()), "unet_pool_bwd")
        return de

    def backward(self, dout):
        """Accumulate parameter gradients for d loss / d out = dout (B, H, W float32),
        using the activations kept by forward(keep=True)."""
        s = self._saved
        if s is None:
            raise RuntimeError("backward needs forward(img, keep=True) first")
        dout = dout.contiguous()
        dd0 = self._empty(s["d0"].shape)
        _ck(L.lib().itomo_unet_head_bwd(_p(s["d0"]), _p(self.hw), _p(dout), dout.numel(),
                                        HEAD_C, _p(dd0), _p(self.grad["hw"]),
                                        _p(self.grad["hb"]), stream_handle()), "unet_head_bwd")
        d = self._conv_back("d0b", s["a_d0"], dd0, s["a_d0"])
        du0 = self._conv_back("d0a", s["u0"], d, s["u0"])
        dd1, ds0 = self._upcat_back(du0, s["d1"], 32)
        d = self._conv_back("d1b", s["a_d1"], dd1, s["a_d1"])
        du1 = self._conv_back("d1a", s["u1"], d, s["u1"])
        dd2, ds1 = self._upcat_back(du1, s["d2"], 64)
        d = self._conv_back("d2b", s["a_d2"], dd2, s["a_d2"])
        du2 = self._conv_back("d2a", s["u2"], d, s["u2"])
        de3, ds2 = self._upcat_back(du2, s["e3"], 128)
        d = self._conv_back("e3b", s["a_e3"], de3, s["a_e3"])
        dp2 = self._conv_back("e3a", s["p2"], d, s["p2"])
        de2 = self._pool_back(dp2, s["i2"], ds2, s["e2"])
        d = self._conv_back("e2b", s["a_e2"], de2, s["a_e2"])
        dp1 = self._conv_back("e2a", s["p1"], d, s["p1"])
        de1 = self._pool_back(dp1, s["i1"], ds1, s["e1"])
        d = self._conv_back("e1b", s["a_e1"], de1, s["a_e1"])
        dp0 = self._conv_back("e1a", s["p0"], d, s["p0"])
        de0 = self._pool_back(dp0, s["i0"], ds0, s["e0"])
        d = self._conv_back("e0b", s["a_e0"], de0, s["a_e0"])
        self._conv_back("e0a", s["x0"], d, None)
        self._saved = None

    def zero_grad(self):
        for t in self.grad.values():
            t.zero_()

    def loss_and_grad(self, out, target):
        """(mean squared error as a 0-d float64 GPU tensor, d loss / d out)."""
        grad = torch.empty_like(out)
        loss = torch.empty((), dtype=torch.float64, device=out.device)
        _ck(L.lib().itomo_unet_mse(_p(out), _p(target.contiguous()), out.numel(), _p(grad),
                                   _p(loss), _p(self._mse_work), stream_handle()), "unet_mse")
        return loss, grad

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        for name, t, tbf in self.params():
            _ck(L.lib().itomo_unet_adam(_p(t), _p(self.grad[name]), _p(self.m[name]),
                                        _p(self.v[name]), _p(tbf), t.numel(), float(lr),
                                        float(beta1), float(beta2), float(eps), self.step,
                                        stream_handle()), "unet_adam")

    def train_step(self, img, target, lr=1e-3):
        """One Adam step on a batch; returns the batch loss (0-d float64 GPU tensor)."""
        self.zero_grad()
        out = self.forward(img, keep=True)
        loss, grad = self.loss_and_grad(out, target)
        self.backward(grad)
        self.adam_step(lr)
        return loss

    def __call__(self, img):
        return self.forward(img)
