"""CPU oracle of the U-Net preconditioner: TEST INFRASTRUCTURE ONLY.

Only tests/ import this; the product (`cuwave_amd.unet`) never does. It
restates the product's network in float32 with torch's CPU library ops, on
the product's own parameters (the weight layouts translated), so the
hand-written matrix-core kernels are checked against a library computation:

- a 3x3 convolution's (cout, 9 cin) weights, row k = tap cin + ci with
  tap = 3 kh + kw, are the (cout, cin, 3, 3) tensor W[co, ci, kh, kw] of
  torch's cross-correlation `conv2d(padding=1)`;
- the 2x2 up-convolution's (cin, 4 cout) GEMM, column (2 a + b) cout + co,
  is `conv_transpose2d` with weight (cin, cout, 2, 2) [ci, co, a, b];
- average pooling, concatenation and the 1x1 output as the product has them.

The product rounds its activations to bfloat16 between layers; the oracle
keeps float32 throughout, so the two agree to bfloat16's precision.
"""

from __future__ import annotations

import torch
import torch.nn.functional as F


def conv3x3_weight(w: torch.Tensor, cin: int) -> torch.Tensor:
    """(cout, 9 cin) -> (cout, cin, 3, 3)."""
    cout = w.shape[0]
    return w.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2).contiguous()


def conv3x3(x: torch.Tensor, w: torch.Tensor, b: torch.Tensor | None, relu: bool) -> torch.Tensor:
    """NHWC in and out."""
    cin = x.shape[-1]
    y = F.conv2d(x.permute(0, 3, 1, 2), conv3x3_weight(w, cin), b, padding=1)
    if relu:
        y = torch.relu(y)
    return y.permute(0, 2, 3, 1)


def upconv(x: torch.Tensor, w: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    cin, cout = w.shape[0], w.shape[1] // 4
    wt = w.reshape(cin, 2, 2, cout).permute(0, 3, 1, 2).contiguous()
    y = F.conv_transpose2d(x.permute(0, 3, 1, 2), wt, b, stride=2)
    return y.permute(0, 2, 3, 1)


def unet(params: dict, x: torch.Tensor, cin_k: int) -> torch.Tensor:
    """The network on NHWC float32 `x`, from the product's named parameters."""
    x = F.pad(x, (0, cin_k - x.shape[-1]))

    def block(name, h):
        for i in range(2):
            h = conv3x3(h, params[f"{name}.{i}.weight"], params[f"{name}.{i}.bias"], True)
        return h

    def pool(h):
        return F.avg_pool2d(h.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)

    e1 = block("enc1", x)
    e2 = block("enc2", pool(e1))
    m = block("mid", pool(e2))
    d2 = block("dec2", torch.cat([upconv(m, params["up2.weight"], params["up2.bias"]), e2], -1))
    d1 = block("dec1", torch.cat([upconv(d2, params["up1.weight"], params["up1.bias"]), e1], -1))
    return d1 @ params["out_w"] + params["out_b"]
