"""Application glue: points off the grid, receiver arrays, and the FWI misfit.

`distribute`, `point_source`, `shots` and `Sensors` place sources and receivers
at physical coordinates by multilinear interpolation (cuwave/utils.py:148-365);
`misfit_gradient` sums the misfit and its gradient over a shot list
(cuwave/utils.py:410-459). Here every shot of the list advances in one batched
launch a step, and the kernel sums the shots' gradients as it goes, already
chained onto the indicator by the parametrization's Jacobian
(cuwave/utils.py:19-24).
"""

from __future__ import annotations

import itertools
from collections.abc import Callable, Sequence

import numpy as np
import torch

from .sensitivity import (
    _Reconstruction,
    _finalize,
    _zero_outside,
    l2_misfit,
    reconstruction_sensitivity,
    require_full_domain,
    sensitivity,
)
from .wave import MAX_BATCH, Simulation, Source, as_numpy, simulate_shots

TOL = 1e-6  # for `distribute`, so exact-boundary coordinates survive rounding


def interior_slice(sim: Simulation) -> tuple[slice, ...]:
    """Index tuple selecting the interior nodes, dropping the ghosts and the padding."""
    return tuple(slice(1, n - 1) for n in sim.Nx)


def _corner_weights(sim: Simulation, coords: np.ndarray):
    """Cell corners and multilinear weights of `coords`, (ndim, num, 2**ndim) and (num, 2**ndim)."""
    index = np.empty_like(coords)
    for d in range(sim.ndim):
        nodal = coords[:, d] / sim.dx[d] + 1.0
        if nodal.min() < 1.0 - TOL or nodal.max() > sim.Nx[d] - 2.0 + TOL:
            length = (sim.Nx[d] - 3) * sim.dx[d]
            raise ValueError(
                f"axis {d} coordinate outside the domain [0, {length:g}]: "
                f"[{float(coords[:, d].min()):g}, {float(coords[:, d].max()):g}]"
            )
        index[:, d] = np.clip(nodal, 1.0, sim.Nx[d] - 2.0)
    base = np.stack([np.clip(np.floor(index[:, d]), 1, sim.Nx[d] - 3).astype(np.int32)
                     for d in range(sim.ndim)])
    offset = index.T - base
    corners = np.array(list(itertools.product((0, 1), repeat=sim.ndim)))
    nodes = base[:, :, None] + corners.T[:, None, :]
    weights = np.ones((len(coords), len(corners)))
    for d in range(sim.ndim):
        w = offset[d][:, None]
        weights *= np.where(corners[None, :, d] == 1, w, 1.0 - w)
    return nodes, weights


def distribute(sim: Simulation, coords) -> tuple[np.ndarray, np.ndarray]:
    """Multilinear interpolation of `coords` (num, ndim) onto the grid.

    Returns the (ndim, num * 2**ndim) corner nodes and the (num, 2**ndim) weights
    in the simulation's precision (cuwave/utils.py:185-247, nodal unknowns).
    """
    coords = np.atleast_2d(np.asarray(coords, dtype=float))
    if coords.shape[1] != sim.ndim:
        raise ValueError(f"coordinates need {sim.ndim} components, not {coords.shape[1]}")
    nodes, weights = _corner_weights(sim, coords)
    return nodes.reshape(sim.ndim, -1).astype(np.int32), weights.astype(sim.np_dtype)


def point_source(sim: Simulation, coords, signal) -> Source:
    """A `Source` injecting `signal` at `coords`, divided by the cell volume (cuwave/utils.py:250-277)."""
    nodes, weights = distribute(sim, coords)
    count = weights.shape[0]
    signal = np.asarray(as_numpy(signal), dtype=sim.np_dtype)
    if signal.ndim == 1:
        signal = np.broadcast_to(signal[:, None], (signal.shape[0], count))
    if signal.shape[1] != count:
        raise ValueError(f"{signal.shape[1]} signal columns for {count} sources")
    columns = (signal[:, :, None] * weights).reshape(signal.shape[0], -1)
    return Source(nodes, np.ascontiguousarray(columns / np.prod(sim.dx), dtype=sim.np_dtype))


def shots(sim: Simulation, coords, signal) -> list[Source]:
    """One single-coordinate `Source` per coordinate: the shot list of an inversion."""
    return [point_source(sim, c, signal) for c in np.atleast_2d(coords)]


def stack(sources: Sequence[Source]) -> Source:
    """Fire several sources in one simulation, their columns concatenated."""
    return Source(
        np.concatenate([np.asarray(as_numpy(s.position)) for s in sources], axis=1),
        np.ascontiguousarray(np.concatenate([as_numpy(s.signal) for s in sources], axis=1)),
    )


class Sensors:
    """Receivers at arbitrary coordinates, multilinearly interpolated onto the grid."""

    def __init__(self, sim: Simulation, coords) -> None:
        self.sim = sim
        self.coordinates = np.atleast_2d(np.asarray(coords, dtype=float))
        self.count = len(self.coordinates)
        self.nodes, self.weights = distribute(sim, self.coordinates)

    def _w(self, like: torch.Tensor) -> torch.Tensor:
        return torch.as_tensor(self.weights, device=like.device, dtype=like.dtype)

    def traces(self, record: torch.Tensor) -> torch.Tensor:
        """(N, count * 2**ndim) node record -> (N, count) receiver traces."""
        return (record.reshape(record.shape[0], self.count, -1) * self._w(record)).sum(2)

    def scatter(self, dphi: torch.Tensor) -> torch.Tensor:
        """The transpose of `traces`: (N, count) -> (N, count * 2**ndim)."""
        return (dphi[:, :, None] * self._w(dphi)).reshape(dphi.shape[0], -1)

    def objective(self, objective: Callable) -> Callable:
        """Lift a receiver-space objective into the node space the adjoint wants."""

        def wrapped(record):
            cost, dphi = objective(self.traces(record))
            return cost, self.scatter(dphi)

        return wrapped


def _batches(items: list, batch: int | None):
    """Slices of at most `batch` (default: all) and at most MAX_BATCH items."""
    step = min(len(items) if not batch else int(batch), MAX_BATCH)
    for i in range(0, len(items), max(step, 1)):
        yield i, items[i : i + step]


def intensity(sim: Simulation, frequencies, weights=None) -> Callable:
    """Objective factory: J = sum_f sum_s w |u_hat_s(f)|^2, the spectral intensity at
    the sensors, u_hat(f) = dt sum_t u(t) e^(-i 2 pi f t dt) (cuwave/utils.py:518-553).

    `weights` broadcasts to (frequencies, sensors) (default 1). The transform is
    linear in the record, so dJ/dtraces is the same tables read back: a sum of
    sinusoids at `frequencies`, what `spectral_sensitivity` needs.
    """
    f = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
    # the phase reaches 1e5 rad over a long run: built in double
    phase = np.outer(np.arange(sim.N, dtype=np.float64) * sim.dt, 2.0 * np.pi * f)
    cos_h, sin_h = np.cos(phase), np.sin(phase)
    scale = float(sim.np_dtype(sim.dt))

    def objective(traces):
        traces = torch.as_tensor(traces)
        dev, dtype = traces.device, traces.dtype
        cos = torch.as_tensor(cos_h, dtype=dtype, device=dev)
        sin = torch.as_tensor(sin_h, dtype=dtype, device=dev)
        w = 1.0 if weights is None else torch.as_tensor(np.asarray(as_numpy(weights)),
                                                       dtype=dtype, device=dev)
        re = scale * (cos.T @ traces)
        im = -scale * (sin.T @ traces)
        cost = float(torch.sum(w * (re * re + im * im)))
        return cost, 2.0 * scale * (cos @ (w * re) - sin @ (w * im))

    return objective


def measure(sim: Simulation, sources: Sequence[Source], indicator, sensors: Sensors,
            devices=None, batch: int | None = None) -> list[torch.Tensor]:
    """Receiver traces per shot, all shots of a batch in one launch a step."""
    if devices is not None:
        raise NotImplementedError("devices= shot dealing is not part of this build (one GPU)")
    out = []
    for _, chunk in _batches(list(sources), batch):
        _, record = simulate_shots(sim, chunk, indicator, sensors.nodes)
        out += [sensors.traces(r) for r in record]
    return out


def misfit(sim: Simulation, sources, indicator, sensors: Sensors, observed,
           objective: Callable = l2_misfit, devices=None, batch: int | None = None) -> float:
    """The cost `misfit_gradient` returns, forward passes only."""
    traces = measure(sim, sources, indicator, sensors, devices, batch)
    return sum(objective(data)(t)[0] for t, data in zip(traces, observed))


def misfit_gradient(sim: Simulation, sources: Sequence[Source], indicator, sensors: Sensors,
                    observed, objective: Callable = l2_misfit, adjoint: Callable = sensitivity,
                    devices=None, batch: int | None = None, checkpoints: int | None = None,
                    illum: torch.Tensor | None = None, below: float | None = None):
    """Misfit summed over the shots and its derivative with respect to `indicator`.

    The shots advance together, `batch` of them a launch (all by default): one
    forward sweep records every shot's traces and boundary strip, then one
    adjoint sweep rebuilds every forward field backwards and sums the gradients
    of all shots in the kernel, chained onto the indicator by
    `parametrization_jacobian`. `adjoint` keeps the reference's meaning for the
    memory a variant is allowed: both `sensitivity` and
    `reconstruction_sensitivity` give the exact transpose, which this sweep is
    for either (the stored history is never needed); any other callable runs
    shot by shot as the reference does.

    Two additions for iterative inversion, neither in the reference: `illum`, a
    per-node accumulator over the padded grid, gathers the illumination (the
    sum over steps and shots of u^2, as `wave.illumination`) in the forward
    sweep; and with `below` given, the adjoint half runs only if the cost comes
    out below it (a line-search trial that fails costs one forward sweep, and
    one that succeeds has the next gradient), every batch of shots then held
    until the cost is known.

    Returns (cost, gradient), the gradient a tensor over the padded grid (None
    when `below` stopped it).
    """
    if devices is not None:
        raise NotImplementedError("devices= shot dealing is not part of this build (one GPU)")
    require_full_domain(sim)
    sources = list(sources)
    from .wave import family_drivers

    own = family_drivers(sim)
    if own is not None:
        if adjoint not in (sensitivity, reconstruction_sensitivity):
            raise NotImplementedError(f"{type(sim).__name__} sums its shots through the reverse "
                                      f"march only")
        if illum is not None or below is not None:
            raise NotImplementedError("illum= and below= are built for the scalar equations")
        return own.misfit_gradient(sim, sources, indicator, sensors, observed, objective,
                                   batch, checkpoints)
    if type(sim).__name__ == "ElasticWave" and adjoint in (sensitivity, reconstruction_sensitivity):
        from . import elastic

        if illum is not None or below is not None:
            raise NotImplementedError("illum= and below= are built for the scalar equations")
        return elastic.misfit_gradient(sim, sources, indicator, sensors, observed, objective,
                                       batch, checkpoints)
    if adjoint not in (sensitivity, reconstruction_sensitivity):
        if illum is not None or below is not None:
            raise NotImplementedError("illum= and below= are for the batched adjoint sweep")
        cost, gradient = 0.0, None
        d_mass, d_stiff = sim.parametrization_jacobian(indicator)
        for source, data in zip(sources, observed):
            c, grads, _, _ = adjoint(sim, source, indicator, sensors.nodes,
                                     sensors.objective(objective(data)))
            g = d_mass * grads["mass"] + d_stiff * grads["stiff"]
            cost += c
            gradient = g if gradient is None else gradient + g
        return cost, gradient
    for s in sources:
        sim.require_interior(s.position, "source")
    sim.require_interior(sensors.nodes, "sensor")
    jm, js = sim.parametrization_jacobian(indicator)
    constant = np.isscalar(jm) and np.isscalar(js)
    if not constant and tuple(sim.indicator_shape()) != tuple(sim.Nx_padded):
        raise NotImplementedError("a stacked indicator goes through autograd (cuwave_amd.autograd)")
    mat = sim.build_materials(indicator)
    cost = 0.0
    gradient = None
    held = []

    def adjoint_half(sweep, dphi):
        try:
            sweep.send(dphi)
        except StopIteration:
            pass

    for start, chunk in _batches(sources, batch):
        run = _Reconstruction(sim, mat, chunk, sensors.nodes, checkpoints=checkpoints,
                              illum=illum, indicator=indicator)
        valid = run.valid
        if constant:  # the kernel chains the Jacobian as it sums: one accumulator
            acc = run.prop.accumulator()
            mode = dict(g_mass=acc, combine=(jm, js), skip_layer=True)
        else:  # a nodal Jacobian: both densities, chained after the sweep
            acc = torch.stack([run.prop.accumulator(), run.prop.accumulator()])
            mode = dict(g_mass=acc[0], g_stiff=acc[1], skip_layer=True)
        sweep = run.sweep(mode)
        record = next(sweep)
        dphis = []
        for b in range(len(chunk)):
            c, dphi = sensors.objective(objective(observed[start + b]))(record[b])
            cost += c
            dphis.append(torch.as_tensor(dphi, dtype=sim.dtype, device=record.device))
        if below is None:
            adjoint_half(sweep, torch.stack(dphis))
            gradient = acc if gradient is None else gradient + acc
        else:
            held.append((sweep, torch.stack(dphis), acc))
    if below is not None:
        if not cost < below:
            return cost, None
        for sweep, dphi, acc in held:
            adjoint_half(sweep, dphi)
            gradient = acc if gradient is None else gradient + acc
    if not constant:
        w = torch.as_tensor(sim.cell_weights(), device=gradient.device)

        def field(d):
            return torch.as_tensor(d, dtype=gradient.dtype, device=gradient.device)

        gradient = field(jm) * (gradient[0] * w) + field(js) * (gradient[1] * w)
        gradient[torch.as_tensor(~valid, device=gradient.device)] = 0.0
        return cost, gradient
    grads = _finalize(sim, {"g": gradient}, gradient.device)
    grads = _zero_outside(grads, ~valid, gradient.device)
    return cost, grads["g"]
