"""The U-Net's matrix-core kernels against torch's CPU library, on the GPU.

Inputs and weights are rounded to bfloat16 first, so both sides see the same
values: bfloat16 products are exact in float32, and the kernels accumulate in
float32, so before the output's own rounding the two agree to float32 sums.
- forward (bias and ReLU fused): within bfloat16's rounding of the output;
- the input gradient: within bfloat16's rounding of the output;
- the weight gradient (float32 out): within 1e-4 relative L2;
- AdamW against torch.optim.AdamW over several steps: 1e-6 relative;
- the whole network against the oracle network (`oracle/unet.py`, float32
  throughout): the output within 3 % relative L2, every parameter gradient
  within 10 % relative L2 at a cosine above 0.99, the drift of bfloat16
  activations and gradients through fourteen layers.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from cuwave_amd import unet  # noqa: E402
from oracle import unet as O  # noqa: E402


def _bf(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.bfloat16).float()


def _rel(a: torch.Tensor, b: torch.Tensor) -> float:
    a, b = a.detach().float().cpu(), b.detach().float().cpu()
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("cin, cout, relu", [(16, 32, True), (32, 32, True), (64, 128, False),
                                             (128, 64, True), (128, 128, True)])
def test_a_convolution_and_its_gradients_match_the_library(cin, cout, relu):
    print(f"case conv3x3 {cin}->{cout} relu={relu}", flush=True)
    g = torch.Generator().manual_seed(cin * 1000 + cout)
    n, h, w = 2, 37, 41  # an odd size: tiles that run past the last pixel
    x = _bf(torch.randn(n, h, w, cin, generator=g))
    conv = unet.Conv3x3(cin, cout, relu)
    with torch.no_grad():
        conv.weight.copy_(_bf(conv.weight))
    xg = x.cuda().to(torch.bfloat16).requires_grad_(cin in (32, 64, 128))
    convg = conv.cuda()
    y = convg(xg)
    ref_x = x.clone().requires_grad_(True)
    ref_w = conv.weight.detach().cpu().clone().requires_grad_(True)
    ref_b = conv.bias.detach().cpu().clone().requires_grad_(True)
    ref = O.conv3x3(ref_x, ref_w, ref_b, relu)
    # within bfloat16's rounding of the output (relative 2^-8 per element)
    err = (y.detach().float().cpu() - ref.detach()).abs()
    assert float((err - ref.detach().abs() * 2.0**-8).max()) <= 1e-5
    dy = _bf(torch.randn(n, h, w, cout, generator=g))
    y.backward(dy.cuda().to(torch.bfloat16))
    ref.backward(dy)
    # the kernels take the gradient before the activation in bfloat16: the same
    # rounded values on the reference side give the weight gradient in float32
    mask = (y.float().cpu() > 0) if relu else torch.ones_like(dy, dtype=torch.bool)
    dpre = torch.where(mask, dy, torch.zeros_like(dy))
    dw_ref = torch.einsum("nhwo,nhwk->ok", dpre.double(),
                          torch.nn.functional.unfold(x.permute(0, 3, 1, 2).double(), 3, padding=1)
                          .reshape(n, cin, 9, h, w).permute(0, 3, 4, 2, 1).reshape(n, h, w, 9 * cin))
    assert _rel(convg.weight.grad, dw_ref.float()) < 1e-4
    assert _rel(convg.bias.grad, dpre.sum((0, 1, 2))) < 1e-5
    if xg.requires_grad:
        assert _rel(xg.grad, ref_x.grad) < 1e-2


def test_adamw_makes_torch_s_update():
    torch.manual_seed(0)
    p = torch.randn(1000)
    grads = [torch.randn(1000) for _ in range(5)]
    ours = p.clone().cuda()
    theirs = p.clone().requires_grad_(True)
    opt = unet.AdamW([ours.requires_grad_(True)], lr=1e-2, weight_decay=0.05)
    ref = torch.optim.AdamW([theirs], lr=1e-2, weight_decay=0.05)
    for g in grads:
        ours.grad = g.cuda()
        opt.step()
        theirs.grad = g.clone()
        ref.step()
    assert _rel(ours, theirs) < 1e-6


def test_the_network_matches_the_oracle_forward_and_backward():
    print("case unet 2x28x36", flush=True)
    torch.manual_seed(1)
    net = unet.UNet().cuda()
    with torch.no_grad():
        for p in net.parameters():
            p.copy_(_bf(p))
    x = torch.randn(2, 28, 36, 2)
    y = net(x.cuda())
    params = {k: v.detach().cpu().clone().requires_grad_(True) for k, v in net.named_parameters()}
    ref = O.unet(params, x, net.cin_k)
    assert y.shape == (2, 28, 36, 1)
    assert _rel(y, ref) < 3e-2
    target = torch.randn_like(ref)
    torch.mean((y - target.cuda()) ** 2).backward()
    torch.mean((ref - target) ** 2).backward()
    # each layer rounds its activations and the gradient it passes back to
    # bfloat16; the deepest parameters see that fourteen times over
    for name, p in net.named_parameters():
        a, b = p.grad.detach().float().cpu().ravel(), params[name].grad.ravel()
        rel, cos = _rel(a, b), float(a @ b / (a.norm() * b.norm()))
        print(f"{name}: rel {rel:.3e} cos {cos:.6f}", flush=True)
        assert rel < 1e-1 and cos > 0.99, name
    count = sum(p.numel() for p in net.parameters())
    assert 440_000 < count < 500_000


def test_training_fits_a_target_on_the_gpu():
    torch.manual_seed(0)
    x = torch.randn(4, 16, 16, 2).cuda()
    y = 0.5 * x[..., :1] - 0.25 * x[..., 1:]
    net = unet.UNet(base=32).cuda()
    opt = unet.AdamW(net.parameters(), lr=3e-3, weight_decay=1e-4)
    losses = []
    for _ in range(60):
        opt.zero_grad()
        loss = torch.mean((net(x) - y) ** 2)
        loss.backward()
        opt.step()
        losses.append(loss.item())
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.3 * losses[0]
