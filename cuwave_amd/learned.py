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
- `fwi`: the plain baseline (the illumination-normalised gradient) and the
  network-led iteration (the network's step first, then the same baseline),
  both a line search that accepts a step only when the exact solver's misfit
  decreases. Both record (seconds, misfit) for time-to-misfit.
"""

from __future__ import annotations

import time
from typing import Sequence

import numpy as np
import torch

from . import utils
from .unet import AdamW, UNet
from .wave import Simulation, Source


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
    """The (2, H, W) network input: the normalised gradient over its 99th percentile of
    magnitude, clipped to +-3 (spikes at the sources would otherwise set the scale),
    and the illumination on a log scale from its floor to its maximum, in [0, 1]."""
    g = normalised_gradient(gradient, illum, floor)
    scale = torch.quantile(g.abs().flatten().float(), 0.99).to(g.dtype)
    g = torch.clamp(g / scale, -3.0, 3.0) if scale > 0 else g
    il = torch.as_tensor(illum, device=g.device, dtype=g.dtype)
    top = il.max()
    if top > 0:
        lo, hi = np.log(floor), np.log(1.0 + floor)
        il = (torch.log(il / top + floor) - lo) / (hi - lo)
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
def _accumulator(sim: Simulation) -> torch.Tensor:
    return torch.zeros(tuple(sim.Nx_padded), dtype=sim.dtype, device=torch.device("cuda"))


def training_sample(sim: Simulation, sources: Sequence[Source], sensors: utils.Sensors,
                    true_indicator, start_indicator, batch: int | None = None):
    """(input, target) for one model: the features at the start model against the
    indicator step to the true one (unit-max scaled)."""
    observed = utils.measure(sim, sources, true_indicator, sensors, batch=batch)
    illum = _accumulator(sim)
    _, gradient = utils.misfit_gradient(sim, sources, start_indicator, sensors, observed,
                                        batch=batch, illum=illum)
    step = torch.as_tensor(np.asarray(true_indicator) - np.asarray(start_indicator),
                           device=gradient.device, dtype=torch.float32)
    return features(gradient, illum), _scaled(step)[None]


def train(net: UNet, samples, epochs: int = 50, lr: float = 1e-3,
          weight_decay: float = 1e-4, batch: int = 16, seed: int = 0,
          flip: bool = True, progress=None) -> list[float]:
    """Fit the network to (input, target) samples: MSE, the library's AdamW, shuffled
    minibatches of `batch`, each sample also mirrored left to right (`flip`); returns
    the mean loss per epoch (each also handed to `progress(epoch, loss)` if given)."""
    xs = _nhwc(torch.stack([_fit(x)[0] for x, _ in samples]))
    ys = _nhwc(torch.stack([_fit(y)[0] for _, y in samples]))
    if flip:  # the acquisition is symmetric under a left-right flip: twice the data
        xs, ys = torch.cat([xs, xs.flip(2)]), torch.cat([ys, ys.flip(2)])
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
        if progress is not None:
            progress(len(losses), losses[-1])
    return losses


# --------------------------------------------------------------------- FWI
def fwi(sim: Simulation, sources: Sequence[Source], sensors: utils.Sensors, observed,
        start_indicator, iterations: int, step: float, net: UNet | None = None,
        batch: int | None = None, halvings: int = 3, floor: float = 1e-3, lead: int = 1):
    """Illumination-normalised gradient FWI, or the same led by the network.

    The misfit gradient and the illumination come from one sweep at the current
    model. The direction is the negative normalised gradient (`net` None), or,
    for the first `lead` iterations, the network's model step with the
    normalised gradient as the fallback: the network learned the step from the
    start model, which later iterates leave behind (its later directions
    measured as tiny accepted steps), so after the lead, or once no step along
    its direction lowers the misfit, the run goes on with the normalised
    gradient alone, from the first step size. A direction is scaled so its
    largest change is the current step and tried at it and `halvings` halvings;
    each trial is the forward half of the gradient sweep at the trial model,
    whose adjoint half runs only if the misfit came out lower. The first such
    trial is accepted (its sweep has the next gradient) and the step grows by
    half, else the next direction (or none) is tried and the step halves.
    Returns the final indicator and the (seconds, misfit) history: the start
    model, then the model after each iteration.
    """
    gamma = np.asarray(start_indicator, dtype=np.float64)
    t0 = time.perf_counter()
    illum = _accumulator(sim)
    cost, gradient = utils.misfit_gradient(sim, sources, gamma.astype(sim.np_dtype), sensors,
                                           observed, batch=batch, illum=illum)
    history = [(time.perf_counter() - t0, cost)]
    first = step
    for it in range(iterations):
        if it == lead and net is not None:
            net, step = None, first  # the gradient starts from the first step's size
        plain = -normalised_gradient(gradient, illum, floor)
        directions = [plain] if net is None else [direction(net, gradient, illum, floor), plain]
        accepted = False
        for which, d in enumerate(directions):
            d = d.detach().cpu().double().numpy()
            d = d / max(float(np.abs(d).max()), 1e-30)
            for k in range(halvings + 1):
                trial = (gamma + d * step / 2**k).astype(sim.np_dtype)
                il = _accumulator(sim)
                c, g = utils.misfit_gradient(sim, sources, trial, sensors, observed,
                                             batch=batch, illum=il, below=cost)
                if g is not None:
                    gamma, cost, gradient, illum = trial.astype(np.float64), c, g, il
                    step = step / 2**k * 1.5
                    accepted = True
                    break
            if accepted:
                break
            if net is not None and which == 0:
                net = None  # the network's direction failed: plain from here on
        if not accepted:
            step /= 2.0
        history.append((time.perf_counter() - t0, cost))
    return gamma.astype(sim.np_dtype), history


def time_to_misfit(plain: list, learned: list) -> dict:
    """The learned run's seconds to reach the plain run's final misfit, against the
    plain run's own seconds to it (its whole run)."""
    target, t_plain = plain[-1][1], plain[-1][0]
    reached = next((t for t, c in learned if c <= target), None)
    return {"target": target, "plain_s": t_plain, "learned_s": reached,
            "ratio": None if reached is None else reached / t_plain}


def evaluate(sim: Simulation, sources, sensors: utils.Sensors, indicator_of, models: dict,
             net: UNet, iterations: int, step: float, batch: int | None = None,
             progress=None) -> list[dict]:
    """Plain against learned FWI on each held-out (true, start) velocity pair of
    `models`: the histories, the time-to-misfit and the final model errors (each
    model's result also handed to `progress(result)` if given)."""
    out = []
    for seed, (true, start) in models.items():
        t_ind, s_ind = indicator_of(true), indicator_of(start)
        observed = utils.measure(sim, sources, t_ind, sensors, batch=batch)
        runs = {}
        for name, n in (("plain", None), ("learned", net)):
            final, history = fwi(sim, sources, sensors, observed, s_ind, iterations, step,
                                 net=n, batch=batch)
            err = float(np.linalg.norm(final - t_ind) / np.linalg.norm(s_ind - t_ind))
            runs[name] = {"history": history, "model_error": err}
        out.append({"seed": seed, **runs, **time_to_misfit(runs["plain"]["history"],
                                                           runs["learned"]["history"])})
        if progress is not None:
            progress(out[-1])
    return out


