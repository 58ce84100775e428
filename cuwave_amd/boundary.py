"""Boundary conditions per face, and the convolutional PML behind them.

A `BoundaryCondition` is a declarative marker (cuwave/boundary.py:30-78): Neumann
refills a face ghost from its mirror, Dirichlet refills it negated and holds the
wall node at zero. Both are folded into the step kernel, which writes the ghost
from the mirror node's thread, so a condition names no kernel of its own here.

`Pml` is the layer of Komatitsch & Martin 2007 (cuwave/boundary.py:272-376): a
memory variable per derivative inside `width` nodes behind each lined face, its
profile graded from zero at the front to the wall and mirrored past it.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass
from typing import TYPE_CHECKING

import numpy as np

if TYPE_CHECKING:
    from .wave import Simulation


@dataclass(frozen=True, repr=False)
class BoundaryCondition:
    """A face condition; frozen, so it keys the face grouping."""

    name: str

    def __repr__(self) -> str:
        return self.name


Neumann = BoundaryCondition("neumann")
Dirichlet = BoundaryCondition("dirichlet")


def canonical_boundary(boundary, ndim: int, default: BoundaryCondition = Neumann) -> tuple:
    """((low, high),) * ndim from a condition, a per-axis one, or None for `default`."""
    if boundary is None:
        boundary = default
    if isinstance(boundary, BoundaryCondition):
        boundary = (boundary,) * ndim
    if len(boundary) != ndim:
        raise ValueError(f"boundary needs one (low, high) pair per axis: {ndim}")
    pairs = []
    for pair in boundary:
        if isinstance(pair, BoundaryCondition):
            pair = (pair, pair)
        if len(pair) != 2:
            raise ValueError("a boundary axis is a (low, high) pair of conditions")
        for c in pair:
            if c not in (Neumann, Dirichlet):
                raise ValueError(f"a pressure face is Neumann or Dirichlet, not {c!r}")
        pairs.append(tuple(pair))
    return tuple(pairs)


def faces_with(sim: Simulation, condition: BoundaryCondition) -> list[int]:
    """The `2 * axis + side` codes of the faces marked with `condition`."""
    return [2 * d + side for d, pair in enumerate(sim.boundary)
            for side, marker in enumerate(pair) if marker == condition]


def face_mask(sim: Simulation, condition: BoundaryCondition) -> int:
    """The faces carrying `condition`, as the bitmask the kernels take."""
    return sum(1 << f for f in faces_with(sim, condition))


def require_off_held_walls(sim: Simulation, rows: np.ndarray) -> None:
    """Raise if a source sits on the wall node of a Dirichlet face, held at zero."""
    for face in faces_with(sim, Dirichlet):
        d, side = divmod(face, 2)
        wall = sim.Nx[d] - 2 if side else 1
        if np.any(rows[d] == wall):
            raise ValueError(f"a source on the wall node {wall} of Dirichlet face {face}, held at zero")


@dataclass(frozen=True)
class Pml:
    """A convolutional perfectly matched layer behind `faces`, for `Simulation.pml`."""

    width: int  # layer depth in nodes, the wall node included
    wavespeed: float  # fastest wave speed in the layer, which sizes the damping
    reflection: float = 1e-6  # target reflection at normal incidence
    frequency: float = 0.0  # CFS shift pi * frequency through the layer, 0 for none
    faces: tuple[int, ...] | None = None  # the `2 * axis + side` codes, None for all


def _canonical_faces(ndim: int, faces: Sequence[int] | None) -> tuple[int, ...]:
    faces = tuple(range(2 * ndim) if faces is None else faces)
    if any(f not in range(2 * ndim) for f in faces):
        raise ValueError(f"face codes run to {2 * ndim - 1} in {ndim}D: {faces}")
    return faces


def validate_pml(sim: Simulation) -> None:
    """Raise if `sim.pml` cannot be stepped: too thin, too wide, or ill-posed."""
    layer = sim.pml
    if layer.width < 2 or layer.wavespeed <= 0.0 or not 0.0 < layer.reflection < 1:
        raise ValueError(f"a pml needs width >= 2, a speed and 0 < reflection < 1: {layer}")
    faces = _canonical_faces(sim.ndim, layer.faces)
    width = layer.width + sim.reach  # the transposed taps reach past the layer
    for d in range(sim.ndim):
        sides = sum(2 * d + side in faces for side in (0, 1))
        if sides * width >= sim.Nx[d] - 2:
            raise ValueError(
                f"{sides} layer(s) of {width} nodes leave no interior on axis {d}: Nx={sim.Nx[d]}"
            )


def pml_widths(sim: Simulation) -> list[tuple[int, int]]:
    """Per axis, the (low, high) layer depths in nodes, 0 for a face left unlined."""
    faces = _canonical_faces(sim.ndim, sim.pml.faces)
    return [tuple(sim.pml.width if 2 * d + side in faces else 0 for side in (0, 1))
            for d in range(sim.ndim)]


def pml_profiles(sim: Simulation) -> list[np.ndarray | None]:
    """Per lined axis the (4, Nx[d]) recursion coefficients a, b at nodes and half points.

    In float64 then rounded to the simulation's precision; the depth below the
    front is (L - distance) / L from the wall node, L = width - 1, so the front
    node is undamped; d0 = -3 c ln(R) / (2 L dx) depth^2, b = exp(-(d0 + alpha) dt),
    a = d0 / (d0 + alpha) (b - 1).
    """
    layer = sim.pml
    L = layer.width - 1
    out = []
    for d, (lo, hi) in enumerate(pml_widths(sim)):
        if not lo + hi:
            out.append(None)
            continue
        n = sim.Nx[d]
        scale = -3.0 * layer.wavespeed * np.log(layer.reflection) / (2.0 * L * sim.dx[d])
        rows = []
        for position in (np.arange(n) + shift for shift in (0.0, 0.5)):
            distance = np.full(n, np.inf)
            if lo:
                distance = np.minimum(distance, np.abs(position - 1))
            if hi:
                distance = np.minimum(distance, np.abs(n - 2 - position))
            depth = np.clip((L - distance) / L, 0.0, 1.0)
            d0 = scale * depth**2
            alpha = np.pi * layer.frequency
            inside = depth > 0.0
            b = np.where(inside, np.exp(-(d0 + alpha) * sim.dt), 1.0)
            a = np.where(inside, d0 / np.maximum(d0 + alpha, 1e-300) * (b - 1.0), 0.0)
            rows += [a, b]
        out.append(np.ascontiguousarray(np.stack(rows), dtype=sim.np_dtype))
    return out


def pml_nodes(sim: Simulation) -> np.ndarray:
    """Interior nodes inside the layer, where the step carries memory terms."""
    inside = np.zeros(sim.Nx_padded, dtype=bool)
    if sim.pml is None:
        return inside
    for d, (lo, hi) in enumerate(pml_widths(sim)):
        index = [slice(None)] * sim.ndim
        for rows in (slice(1, 1 + lo), slice(sim.Nx[d] - 1 - hi, sim.Nx[d] - 1)):
            index[d] = rows
            inside[tuple(index)] = True
    interior = np.zeros_like(inside)
    interior[tuple(slice(1, n - 1) for n in sim.Nx)] = True
    return inside & interior
