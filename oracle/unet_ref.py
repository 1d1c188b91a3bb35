"""CPU float64 restatement of the learned warm-start network (test infrastructure only:
the GPU U-Net of iterative_tomographic_reconstruction_algorithms_amd/unet.py is
checked against it; nothing in the product imports this file).

The network has no counterpart in the reference; it is restated here with torch CPU
float64 functional ops (conv2d, max_pool2d, nearest upsampling, ReLU) and autograd for
the gradients.  With ``bf16=True`` every tensor the GPU stores in BF16 is rounded to
BF16 at the same place -- the input image, each activation after its ReLU, and in the
backward pass the gradient of each of those -- so that the GPU and this restatement
differ only by FP32 versus FP64 accumulation.  With ``bf16=False`` it is the exact FP64
network, against which the BF16 error itself is measured.
"""
import torch
import torch.nn.functional as F

CONVS = [
    ("e0a", 1, 32), ("e0b", 32, 32),
    ("e1a", 32, 64), ("e1b", 64, 64),
    ("e2a", 64, 128), ("e2b", 128, 128),
    ("e3a", 128, 256), ("e3b", 256, 256),
    ("d2a", 384, 128), ("d2b", 128, 128),
    ("d1a", 192, 64), ("d1b", 64, 64),
    ("d0a", 96, 32), ("d0b", 32, 32),
]


def bf16_round(x):
    """Round float64 values to the nearest BF16 (ties to even), as float64."""
    return x.to(torch.float32).to(torch.bfloat16).to(torch.float64)


class _Round(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return bf16_round(x)

    @staticmethod
    def backward(ctx, g):
        return bf16_round(g)


def conv_weight(w_cout_9_cin):
    """[cout][9][cin] (the GPU layout) -> [cout][cin][3][3] (torch's)."""
    cout, _, cin = w_cout_9_cin.shape
    return w_cout_9_cin.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)


def forward(params, img, bf16=True):
    """params: dict name -> float64 CPU tensor ('w.<conv>' [cout][9][cin], 'b.<conv>',
    'hw', 'hb'); img: (B, H, W) float64.  Returns out (B, H, W)."""
    r = _Round.apply if bf16 else (lambda t: t)

    def conv(x, name, relu=True):
        y = F.conv2d(x, conv_weight(params["w." + name]), params["b." + name], padding=1)
        return r(F.relu(y)) if relu else r(y)

    x0 = r(img[:, None])
    e0 = conv(conv(x0, "e0a"), "e0b")
    p0 = r(F.max_pool2d(e0, 2))
    e1 = conv(conv(p0, "e1a"), "e1b")
    p1 = r(F.max_pool2d(e1, 2))
    e2 = conv(conv(p1, "e2a"), "e2b")
    p2 = r(F.max_pool2d(e2, 2))
    e3 = conv(conv(p2, "e3a"), "e3b")

    def upcat(a, b):
        return r(torch.cat([F.interpolate(a, scale_factor=2, mode="nearest"), b], dim=1))

    d2 = conv(conv(upcat(e3, e2), "d2a"), "d2b")
    d1 = conv(conv(upcat(d2, e1), "d1a"), "d1b")
    d0 = conv(conv(upcat(d1, e0), "d0a"), "d0b")
    head = torch.einsum("bchw,c->bhw", d0, params["hw"]) + params["hb"][0]
    return img + head


def loss_and_grads(params, img, target, bf16=True):
    """(mean squared error, dict of d loss / d param) in float64 on the CPU."""
    leaves = {k: v.detach().clone().requires_grad_(True) for k, v in params.items()}
    out = forward(leaves, img, bf16)
    loss = ((out - target) ** 2).mean()
    loss.backward()
    return loss.item(), {k: v.grad for k, v in leaves.items()}
