"""The U-Net preconditioner on the MI355X matrix cores.

Every 3x3 convolution (forward, the gradient at its input, the gradient at
its weights) runs in the hand-written bfloat16 MFMA kernels of
`csrc/unet.hip` with float32 accumulation, through the C ABI
(`cw_conv3x3_bf16_fwd` / `_bwd_data` / `_bwd_weight`), and the optimiser step
is the library's AdamW kernel (`cw_adamw`). Activations are NHWC bfloat16,
master weights float32. The pooling, the 2x2 up-convolutions (plain GEMMs
on the library's bfloat16 matrix multiply), the concatenations and the 1x1
output stay in torch.

No reference counterpart: the reference has no learned preconditioner.
There is no CPU path; without the library or a GPU the network raises.
"""

from __future__ import annotations

import ctypes
import math

import torch
from torch import nn

from . import _abi

CIN_ALIGN = 16  # the kernels take channel counts in multiples of 16


def _args(n, h, w, cin, cout) -> _abi.Conv3x3Args:
    a = _abi.Conv3x3Args()
    a.n, a.h, a.w, a.cin, a.cout = int(n), int(h), int(w), int(cin), int(cout)
    return a


def _bf16_weights(w: torch.Tensor, cout: int, cin: int, flipped: bool):
    wt = torch.empty((cout, 9 * cin), dtype=torch.bfloat16, device=w.device)
    wd = torch.empty((cin, 9 * cout), dtype=torch.bfloat16, device=w.device) if flipped else None
    _abi.call("cw_conv3x3_weights", cout, cin, _abi.ptr(w), _abi.ptr(wt), _abi.ptr(wd),
              _abi.stream())
    return wt, wd


def _splits(pixels: int) -> int:
    """Pixel runs a weight gradient sums apart: about 2 k pixels each, 16 .. 256 runs."""
    return max(16, min(256, pixels // 2048))


class _Conv3x3(torch.autograd.Function):
    """y = act(conv3x3(x, w) + b) on NHWC bfloat16, w float32 (cout, 9 cin)."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, w: torch.Tensor, b: torch.Tensor, relu: bool):
        n, h, wd_, cin = x.shape
        cout = w.shape[0]
        if w.shape[1] != 9 * cin:
            raise ValueError(f"weights {tuple(w.shape)} do not fit {cin} input channels")
        x = x.contiguous()
        wt, wflip = _bf16_weights(w.detach().contiguous(), cout, cin, flipped=cin in (32, 64, 128))
        y = torch.empty((n, h, wd_, cout), dtype=torch.bfloat16, device=x.device)
        a = _args(n, h, wd_, cin, cout)
        a.x, a.weight, a.bias, a.relu, a.y = _abi.ptr(x), _abi.ptr(wt), _abi.ptr(b.detach()), int(relu), _abi.ptr(y)
        _abi.call("cw_conv3x3_bf16_fwd", ctypes.byref(a), _abi.stream())
        ctx.save_for_backward(x, y)
        ctx.wflip, ctx.relu = wflip, relu
        return y

    @staticmethod
    def backward(ctx, dy: torch.Tensor):
        x, y = ctx.saved_tensors
        n, h, w, cin = x.shape
        cout = y.shape[-1]
        dy = dy.to(torch.bfloat16)
        if ctx.relu:
            dy = torch.where(y > 0, dy, torch.zeros((), dtype=dy.dtype, device=dy.device))
        dy = dy.contiguous()
        a = _args(n, h, w, cin, cout)
        a.x, a.dy = _abi.ptr(x), _abi.ptr(dy)
        dx = None
        if ctx.needs_input_grad[0]:
            if ctx.wflip is None:
                raise NotImplementedError("the input gradient is built for 32, 64 or 128 channels")
            dx = torch.empty_like(x)
            a.weight_flipped, a.dx = _abi.ptr(ctx.wflip), _abi.ptr(dx)
            _abi.call("cw_conv3x3_bf16_bwd_data", ctypes.byref(a), _abi.stream())
        splits = _splits(n * h * w)
        work = torch.empty((splits, 9 * cin, cout), dtype=torch.float32, device=x.device)
        dweight = torch.empty((cout, 9 * cin), dtype=torch.float32, device=x.device)
        a.workspace, a.splits, a.dweight = _abi.ptr(work), splits, _abi.ptr(dweight)
        _abi.call("cw_conv3x3_bf16_bwd_weight", ctypes.byref(a), _abi.stream())
        dbias = dy.float().sum((0, 1, 2))
        return dx, dweight, dbias, None


class Conv3x3(nn.Module):
    """A 3x3 convolution (padding 1) on NHWC bfloat16, its weights (cout, 9 cin) in
    float32, row k = tap cin + ci with tap = 3 kh + kw; torch's default initialisation."""

    def __init__(self, cin: int, cout: int, relu: bool = True):
        super().__init__()
        self.cin, self.cout, self.relu = cin, cout, relu
        bound = 1.0 / math.sqrt(9 * cin)
        self.weight = nn.Parameter(torch.empty(cout, 9 * cin).uniform_(-bound, bound))
        self.bias = nn.Parameter(torch.empty(cout).uniform_(-bound, bound))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        return _Conv3x3.apply(x, self.weight, self.bias, self.relu)


class UpConv(nn.Module):
    """A 2x2, stride-2 transposed convolution on NHWC: one GEMM to (cin -> 4 cout), then a
    scatter of each pixel's four outputs onto its 2x2 block."""

    def __init__(self, cin: int, cout: int):
        super().__init__()
        bound = 1.0 / math.sqrt(4 * cout)
        self.cout = cout
        self.weight = nn.Parameter(torch.empty(cin, 4 * cout).uniform_(-bound, bound))
        self.bias = nn.Parameter(torch.empty(cout).uniform_(-bound, bound))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        n, h, w, _ = x.shape
        y = (x @ self.weight.to(x.dtype)).reshape(n, h, w, 2, 2, self.cout)
        y = y.permute(0, 1, 3, 2, 4, 5).reshape(n, 2 * h, 2 * w, self.cout)
        return y + self.bias.to(y.dtype)


def _pool(x: torch.Tensor) -> torch.Tensor:
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean((2, 4))


class UNet(nn.Module):
    """Three levels, 32 / 64 / 128 channels: (B, H, W, cin) -> (B, H, W, cout) NHWC,
    H and W divisible by 4; the input's channels padded to 16 inside."""

    def __init__(self, cin: int = 2, cout: int = 1, base: int = 32):
        super().__init__()
        c1, c2, c3 = base, 2 * base, 4 * base
        self.cin = cin
        self.cin_k = -(-cin // CIN_ALIGN) * CIN_ALIGN
        self.enc1 = nn.Sequential(Conv3x3(self.cin_k, c1), Conv3x3(c1, c1))
        self.enc2 = nn.Sequential(Conv3x3(c1, c2), Conv3x3(c2, c2))
        self.mid = nn.Sequential(Conv3x3(c2, c3), Conv3x3(c3, c3))
        self.up2 = UpConv(c3, c2)
        self.dec2 = nn.Sequential(Conv3x3(2 * c2, c2), Conv3x3(c2, c2))
        self.up1 = UpConv(c2, c1)
        self.dec1 = nn.Sequential(Conv3x3(2 * c1, c1), Conv3x3(c1, c1))
        bound = 1.0 / math.sqrt(c1)
        self.out_w = nn.Parameter(torch.empty(c1, cout).uniform_(-bound, bound))
        self.out_b = nn.Parameter(torch.empty(cout).uniform_(-bound, bound))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        if x.device.type != "cuda":
            raise RuntimeError("the U-Net runs on the GPU's matrix cores only")
        _abi.load()
        pad = self.cin_k - x.shape[-1]
        x = torch.nn.functional.pad(x.to(torch.bfloat16), (0, pad)).contiguous()
        e1 = self.enc1(x)
        e2 = self.enc2(_pool(e1))
        m = self.mid(_pool(e2))
        d2 = self.dec2(torch.cat([self.up2(m), e2], -1).contiguous())
        d1 = self.dec1(torch.cat([self.up1(d2), e1], -1).contiguous())
        return d1.float() @ self.out_w + self.out_b


class AdamW:
    """AdamW over float32 parameters through the library's kernel (cw_adamw): the
    update torch.optim.AdamW makes, decoupled weight decay and bias-corrected moments."""

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 1e-2):
        self.params = [p for p in params if p.requires_grad]
        self.lr, self.betas, self.eps, self.weight_decay = lr, betas, eps, weight_decay
        self.state = [(torch.zeros_like(p), torch.zeros_like(p)) for p in self.params]
        self.t = 0

    def zero_grad(self) -> None:
        for p in self.params:
            p.grad = None

    @torch.no_grad()
    def step(self) -> None:
        self.t += 1
        for p, (m, v) in zip(self.params, self.state):
            if p.grad is None:
                continue
            if p.dtype != torch.float32 or not p.is_contiguous():
                raise ValueError("AdamW takes contiguous float32 parameters")
            g = p.grad.contiguous()
            a = _abi.AdamWArgs()
            a.count = p.numel()
            a.param, a.grad, a.m, a.v = _abi.ptr(p), _abi.ptr(g), _abi.ptr(m), _abi.ptr(v)
            a.lr, a.beta1, a.beta2 = self.lr, self.betas[0], self.betas[1]
            a.eps, a.weight_decay, a.step = self.eps, self.weight_decay, self.t
            _abi.call("cw_adamw", ctypes.byref(a), _abi.stream())
