"""The solver as a differentiable torch operation: materials in, traces out.

`traces(indicator, sim, sources, sensors)` runs every shot forward and
returns the (B, N, count) receiver traces as a tensor that carries a
gradient back to `indicator`. A loss of any form can then be written in
torch on the traces. On `backward()` the incoming gradient of the traces is
the adjoint source, as the reference's objective callables return it
(`objective(traces) -> (cost, dcost/dtraces)`, cuwave/sensitivity.py:93-100).
One adjoint sweep over the batch then rebuilds the forward field backwards,
as `reconstruction_sensitivity` does. The mass and stiffness densities it
sums are chained onto the indicator by the simulation's
`parametrization_jacobian`, as the reference's caller does by hand
(cuwave/utils.py:19-24). That jacobian may be constant, when the kernels
chain it into one accumulator as they sum (and skip a density whose factor
is zero), or field-valued, as for `BulkDensityWave`.

The forward half runs inside the autograd forward and keeps the boundary
strip and the checkpoints. The adjoint half runs only if a gradient is asked
for. Double backward is not supported.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .sensitivity import _finalize, _Reconstruction, _zero_outside
from .utils import MAX_BATCH, Sensors
from .wave import Simulation, Source, as_numpy


def _chain(sim: Simulation, indicator: np.ndarray, g_mass: torch.Tensor,
           g_stiff: torch.Tensor) -> torch.Tensor:
    """dJ/d(indicator) from the mass and stiffness densities."""
    d_mass, d_stiff = sim.parametrization_jacobian(indicator)

    def field(d):
        return torch.as_tensor(d, dtype=g_mass.dtype, device=g_mass.device)

    if tuple(sim.indicator_shape()) != tuple(sim.Nx_padded):
        # one indicator field per material: (mass field, stiffness field)
        return torch.stack([field(d_mass) * g_mass, field(d_stiff) * g_stiff])
    return field(d_mass) * g_mass + field(d_stiff) * g_stiff


class Traces(torch.autograd.Function):
    """indicator -> (B, N, count) traces of one batch of at most 64 shots."""

    @staticmethod
    def forward(ctx, indicator: torch.Tensor, sim: Simulation, sources: Sequence[Source],
                sensors: Sensors, checkpoints: int | None) -> torch.Tensor:
        ind = np.asarray(as_numpy(indicator.detach()), dtype=sim.np_dtype)
        for s in sources:
            sim.require_interior(s.position, "source")
        sim.require_interior(sensors.nodes, "sensor")
        mat = sim.build_materials(ind)
        run = _Reconstruction(sim, mat, list(sources), sensors.nodes, checkpoints=checkpoints,
                              indicator=ind)
        jm, js = sim.parametrization_jacobian(ind)
        if np.isscalar(jm) and np.isscalar(js):
            # a constant Jacobian chains in the kernel: one accumulator, and no
            # stiffness density at all where its factor is zero
            g_mass, g_stiff = run.prop.accumulator(), None
            mode = dict(g_mass=g_mass, combine=(jm, js), skip_layer=True)
        else:
            g_mass, g_stiff = run.prop.accumulator(), run.prop.accumulator()
            mode = dict(g_mass=g_mass, g_stiff=g_stiff, skip_layer=True)
        sweep = run.sweep(mode)
        record = next(sweep)
        out = torch.stack([sensors.traces(record[b]) for b in range(run.B)])
        ctx.state = (sim, ind, run, sweep, g_mass, g_stiff, sensors, indicator.device,
                     indicator.dtype)
        return out.to(indicator.device)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad: torch.Tensor):
        sim, ind, run, sweep, g_mass, g_stiff, sensors, device, dtype = ctx.state
        ctx.state = None
        grad = grad.to(run.dev, sim.dtype)
        dphi = torch.stack([sensors.scatter(grad[b]) for b in range(run.B)])
        try:
            sweep.send(dphi)
        except StopIteration:
            pass
        if g_stiff is None:
            grads = _zero_outside(_finalize(sim, {"g": g_mass}, run.dev), ~run.valid, run.dev)
            return grads["g"].to(device, dtype), None, None, None, None
        grads = _finalize(sim, {"mass": g_mass, "stiff": g_stiff}, run.dev)
        grads = _zero_outside(grads, ~run.valid, run.dev)
        gi = _chain(sim, ind, grads["mass"], grads["stiff"])
        return gi.to(device, dtype), None, None, None, None


def traces(indicator: torch.Tensor, sim: Simulation, sources: Sequence[Source],
           sensors: Sensors, checkpoints: int | None = None,
           batch: int | None = None) -> torch.Tensor:
    """The (B, N, count) traces of every shot, differentiable in `indicator`.

    `indicator` has the simulation's indicator shape (`sim.indicator_shape()`).
    Shots run `batch` at a time, at most 64 (all of them by default).
    `checkpoints` bounds the memory of the reverse march, as in
    `reconstruction_sensitivity`.
    """
    sources = list(sources)
    if not sources:
        raise ValueError("no shots")
    if not isinstance(sensors, Sensors):
        raise TypeError("sensors: a cuwave_amd.utils.Sensors")
    step = min(batch or MAX_BATCH, MAX_BATCH)
    if step <= 0:
        raise ValueError(f"batch must be positive, got {batch}")
    parts = [Traces.apply(indicator, sim, sources[i:i + step], sensors, checkpoints)
             for i in range(0, len(sources), step)]
    return parts[0] if len(parts) == 1 else torch.cat(parts)
