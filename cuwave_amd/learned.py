"""Learned FWI: a U-Net that preconditions the illumination-normalised gradient.

This has no reference counterpart. The reference's `cuwave/nn.py` is a design
generator, not a preconditioner. The pieces:

- `layered_model`: the velocity models, a layered background (3-8 layers,
  1500-2500 m/s) plus 1-6 elliptical inclusions, seeded.
- `UNet` (`unet.py`): three levels of 3x3 convolutions, 32/64/128 channels,
  about 0.47 M parameters, NHWC bfloat16 activations; every 3x3 convolution
  (forward and both gradients) runs in the library's hand-written MFMA
  kernels with float32 accumulation, and training steps with its AdamW
  kernel.
- `normalised_gradient`: the misfit gradient divided by the illumination map
  (`wave.illumination`), floored.
- `training_sample` / `train`: inputs (the normalised gradient and the
  normalised illumination at the start model) against the target (the
  indicator step from the start model to the true one), MSE, AdamW.
- `fwi`: the plain baseline (Adam on the indicator, as the reference's
  example FWI runs it) and the preconditioned iteration. The preconditioned
  iteration accepts a step only when the exact solver's misfit decreases,
  halving it otherwise. Both record (seconds, misfit) for time-to-misfit.
"""

from __future__ import annotations

import time
from typing import Sequence

import numpy as np
import torch

from . import utils
from .unet import AdamW, UNet
from .wave import Simulation, Source, illumination


# ------------------------------------------------------------------ models
def layered_model(seed: int, shape: tuple[int, int], vmin: float = 1500.0,
                  vmax: float = 2500.0) -> np.ndarray:
    """A layered background of 3-8 flat layers plus 1-6 ellipses (m/s)."""
    rng = np.random.default_rng(seed)
    n0, n1 = shape
    layers = int(rng.integers(3, 9))
    bounds = np.sort(rng.choice(np.arange(1, n0), size=layers - 1, replace=False))
    speeds = np.sort(rng.uniform(vmin, vmax, layers))
    v = np.empty(shape)
    top = 0
    for k, bottom in enumerate(list(bounds) + [n0]):
        v[top:bottom] = speeds[k]
        top = bottom
    a0, a1 = np.meshgrid(np.arange(n0), np.arange(n1), indexing="ij")
    for _ in range(int(rng.integers(1, 7))):
        c = rng.uniform(vmin, vmax)
        ax = rng.uniform(0.04, 0.15, 2) * min(shape)
        ang = rng.uniform(0, np.pi)
        cen = rng.uniform(0, 1, 2) * np.array(shape)
        x, y = a0 - cen[0], a1 - cen[1]
        u = x * np.cos(ang) + y * np.sin(ang)
        w = -x * np.sin(ang) + y * np.cos(ang)
        v[(u / ax[0]) ** 2 + (w / ax[1]) ** 2 <= 1.0] = c
    return v


# --------------------------------------------------------------- features
def normalised_gradient(gradient: torch.Tensor, illum: torch.Tensor,
                        floor: float = 1e-3) -> torch.Tensor:
    """gradient / (illum + floor * max(illum)): the illumination-normalised gradient."""
    illum = torch.as_tensor(illum, device=gradient.device, dtype=gradient.dtype)
    return gradient / (illum + floor * illum.max())


def _scaled(x: torch.Tensor) -> torch.Tensor:
    m = x.abs().max()
    return x / m if m > 0 else x


def _fit(x: torch.Tensor, mult: int = 4):
    """Pad the last two axes up to a multiple of `mult` (edge values); returns the crop."""
    h, w = x.shape[-2:]
    ph, pw = (-h) % mult, (-w) % mult
    if ph or pw:
        x = torch.nn.functional.pad(x, (0, pw, 0, ph), mode="replicate")
    return x, (h, w)


def features(gradient: torch.Tensor, illum: torch.Tensor, floor: float = 1e-3) -> torch.Tensor:
    """The (2, H, W) network input: the normalised gradient and the illumination, each
    scaled to a unit maximum."""
    g = _scaled(normalised_gradient(gradient, illum, floor))
    il = _scaled(torch.as_tensor(illum, device=g.device, dtype=g.dtype))
    return torch.stack([g, il]).float()


def _nhwc(x: torch.Tensor) -> torch.Tensor:
    """(B, C, H, W) -> (B, H, W, C)."""
    return x.permute(0, 2, 3, 1).contiguous()


def direction(net: UNet, gradient: torch.Tensor, illum: torch.Tensor,
              floor: float = 1e-3) -> torch.Tensor:
    """The network's model step for a gradient, on the gradient's grid and device."""
    x, (h, w) = _fit(features(gradient, illum, floor)[None])
    with torch.no_grad():
        y = net(_nhwc(x))
    return y[0, :h, :w, 0].to(gradient.dtype)


# ---------------------------------------------------------------- training
def training_sample(sim: Simulation, sources: Sequence[Source], sensors: utils.Sensors,
                    true_indicator, start_indicator, batch: int | None = None):
    """(input, target) for one model: the features at the start model against the
    indicator step to the true one (unit-max scaled)."""
    observed = utils.measure(sim, sources, true_indicator, sensors, batch=batch)
    _, gradient = utils.misfit_gradient(sim, sources, start_indicator, sensors, observed,
                                        batch=batch)
    illum = illumination(sim, sources, start_indicator, batch=batch)
    step = torch.as_tensor(np.asarray(true_indicator) - np.asarray(start_indicator),
                           device=gradient.device, dtype=torch.float32)
    return features(gradient, illum), _scaled(step)[None]


def train(net: UNet, samples, epochs: int = 50, lr: float = 1e-3,
          weight_decay: float = 1e-4, batch: int = 16, seed: int = 0) -> list[float]:
    """Fit the network to (input, target) samples: MSE, the library's AdamW, shuffled
    minibatches of `batch`; returns the mean loss per epoch."""
    xs = _nhwc(torch.stack([_fit(x)[0] for x, _ in samples]))
    ys = _nhwc(torch.stack([_fit(y)[0] for _, y in samples]))
    net = net.to(xs.device)
    opt = AdamW(net.parameters(), lr=lr, weight_decay=weight_decay)
    order = torch.Generator().manual_seed(seed)
    losses = []
    for _ in range(epochs):
        perm = torch.randperm(len(xs), generator=order)
        total = 0.0
        for i in range(0, len(xs), batch):
            idx = perm[i:i + batch].to(xs.device)
            opt.zero_grad()
            loss = torch.mean((net(xs[idx]) - ys[idx]) ** 2)
            loss.backward()
            opt.step()
            total += loss.item() * len(idx)
        losses.append(total / len(xs))
    return losses


# --------------------------------------------------------------------- FWI
def fwi(sim: Simulation, sources: Sequence[Source], sensors: utils.Sensors, observed,
        start_indicator, iterations: int, step: float, net: UNet | None = None,
        batch: int | None = None, halvings: int = 4, floor: float = 1e-3):
    """Plain (net None: Adam on the indicator) or preconditioned FWI.

    The preconditioned iteration moves along the network's direction, scaled so
    its largest change is `step`, and accepts the step only when the exact
    misfit decreases, halving it up to `halvings` times. Returns the final
    indicator and the (seconds, misfit) history, the first entry the start.
    """
    gamma = torch.tensor(np.asarray(start_indicator, dtype=np.float64))
    t0 = time.perf_counter()
    history = []
    adam = None
    if net is None:
        gamma.requires_grad_(True)
        adam = torch.optim.Adam([gamma], lr=step)
    for _ in range(iterations):
        current = gamma.detach().numpy().astype(sim.np_dtype)
        cost, gradient = utils.misfit_gradient(sim, sources, current, sensors, observed,
                                               batch=batch)
        history.append((time.perf_counter() - t0, cost))
        if net is None:
            adam.zero_grad(set_to_none=True)
            gamma.grad = gradient.detach().cpu().double()
            adam.step()
            continue
        illum = illumination(sim, sources, current, batch=batch)
        d = direction(net, gradient, illum, floor).detach().cpu().double()
        d = d / max(float(d.abs().max()), 1e-30) * step
        for _ in range(halvings + 1):
            trial = (gamma + d).numpy().astype(sim.np_dtype)
            trial_cost = utils.misfit(sim, sources, trial, sensors, observed, batch=batch)
            if trial_cost < cost:
                gamma = torch.tensor(trial.astype(np.float64))
                break
            d = d / 2
    final = gamma.detach().numpy().astype(sim.np_dtype)
    history.append((time.perf_counter() - t0,
                    utils.misfit(sim, sources, final, sensors, observed, batch=batch)))
    return final, history
