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
# the faces of the elastic family (cuwave/boundary.py:30-78): a traction-free
# face zeroes the normal stress on its wall node, a clamped one holds the
# displacement there at zero
Traction = BoundaryCondition("traction")
Clamped = BoundaryCondition("clamped")
# the faces of the vector Maxwell family (cuwave/boundary.py:30-78): a
# conductor holds the tangential electric field at zero on its wall node, a
# magnetic wall holds the tangential magnetic field (the curl) there
Conductor = BoundaryCondition("conductor")
Magnetic = BoundaryCondition("magnetic")


def canonical_boundary(boundary, ndim: int, default: BoundaryCondition = Neumann,
                       allowed: tuple = (Neumann, Dirichlet)) -> tuple:
    """((low, high),) * ndim from a condition, a per-axis one, or None for `default`;
    every condition one of `allowed`."""
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
            if c not in allowed:
                names = " or ".join(map(repr, allowed))
                raise ValueError(f"a face of this equation is {names}, not {c!r}")
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
    """A convolutional perfectly matched layer behind `faces`, for `Simulation.pml`.

    `width` counts the layer's nodes from its front to the wall node, both
    included. `wavespeed` is the fastest speed that reaches the layer and sets
    the damping that gives a normal-incidence echo of `reflection`. A non-zero
    `frequency` adds the complex-frequency shift alpha = pi * frequency
    throughout the layer. `faces` names the lined faces by `2 * axis + side`
    (None lines every face).
    """

    width: int
    wavespeed: float
    reflection: float = 1e-6
    frequency: float = 0.0
    faces: tuple[int, ...] | None = None


def _canonical_faces(ndim: int, faces: Sequence[int] | None) -> tuple[int, ...]:
    faces = tuple(range(2 * ndim) if faces is None else faces)
    if any(f not in range(2 * ndim) for f in faces):
        raise ValueError(f"face codes run to {2 * ndim - 1} in {ndim}D: {faces}")
    return faces


def validate_pml(sim: Simulation) -> None:
    """Raise if `sim.pml` cannot be stepped: too thin, too wide, or ill-posed."""
    layer = sim.pml
    least = 2 + sim.pml_margin
    if layer.width < least or layer.wavespeed <= 0.0 or not 0.0 < layer.reflection < 1:
        raise ValueError(
            f"a pml needs width >= {least}, a speed and 0 < reflection < 1: {layer}")
    if sim.pml_shifted and layer.frequency <= 0.0:
        raise ValueError(
            f"{type(sim).__name__} keeps a static mode in an unshifted pml: set frequency, "
            f"near the low end of the band")
    faces = _canonical_faces(sim.ndim, layer.faces)
    if sim.pml_walls is not None:
        for face in faces:
            wall = sim.boundary[face // 2][face % 2]
            if wall not in sim.pml_walls:
                raise ValueError(
                    f"face {face} ends the pml on {wall!r}, which traps a growing wave: make it "
                    f"{' or '.join(map(repr, sim.pml_walls))}")
    width = layer.width + sim.reach  # the transposed taps reach past the layer
    for d in range(sim.ndim):
        sides = sum(2 * d + side in faces for side in (0, 1))
        if sides * width >= sim.Nx[d] - 2:
            raise ValueError(
                f"{sides} layer(s) of {width} nodes leave no interior on axis {d}: Nx={sim.Nx[d]}"
            )


def pml_widths(sim: Simulation) -> list[tuple[int, int]]:
    """Layer depth behind the low and the high face of every axis (0 where unlined)."""
    lined = set(_canonical_faces(sim.ndim, sim.pml.faces))
    depth = sim.pml.width
    return [(depth * (2 * d in lined), depth * (2 * d + 1 in lined)) for d in range(sim.ndim)]


def pml_profiles(sim: Simulation) -> list[np.ndarray | None]:
    """Per lined axis the (4, Nx[d]) recursion coefficients a, b at nodes and half points.

    In float64 then rounded to the simulation's precision; the depth below the
    front is (L - distance) / L from the wall node, L = width - 1 - margin, so
    the front node and the `sim.pml_margin` nodes before it are undamped;
    d0 = -3 c ln(R) / (2 L dx) depth^2, b = exp(-(d0 + alpha) dt),
    a = d0 / (d0 + alpha) (b - 1).
    """
    layer = sim.pml
    L = layer.width - 1 - sim.pml_margin
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
    """Mask of the interior nodes that carry layer memory (on some lined axis within the
    layer depth of its wall node), over the padded grid."""
    mask = np.zeros(sim.Nx_padded, dtype=bool)
    if sim.pml is None:
        return mask
    inner = np.ones(sim.Nx_padded, dtype=bool)
    for d, (lo, hi) in enumerate(pml_widths(sim)):
        n = sim.Nx[d]
        a = np.arange(sim.Nx_padded[d])
        shape = [1] * sim.ndim
        shape[d] = -1
        # wall node 1 and the lo - 1 rows past it; wall node n - 2 and hi - 1 before it
        layer_rows = ((a >= 1) & (a <= lo)) | ((a >= n - 1 - hi) & (a <= n - 2))
        mask |= layer_rows.reshape(shape)
        inner &= ((a >= 1) & (a <= n - 2)).reshape(shape)
    return mask & inner


# ------------------------------------------------------------------ sponge
def _layer_faces(sim: Simulation, width: int, faces: Sequence[int] | None) -> tuple[int, ...]:
    """The faces a layer of `width` nodes lines, checked to leave an interior
    (cuwave/boundary.py:162-174)."""
    if width < 1:
        raise ValueError(f"width must be at least one node: {width}")
    faces = _canonical_faces(sim.ndim, faces)
    for d in range(sim.ndim):
        sides = sum(2 * d + side in faces for side in (0, 1))
        if sides * width >= sim.Nx[d] - 2:
            raise ValueError(f"{sides} layer(s) of {width} nodes leave no interior on axis {d}: "
                             f"Nx={sim.Nx[d]}")
    return faces


def _layer_ramp(sim: Simulation, width: int, faces: tuple[int, ...]) -> np.ndarray:
    """1 on a lined wall node falling linearly to 0 `width` nodes in, the larger of
    two faces' ramps in a corner, zero off the interior (cuwave/boundary.py:177-193)."""
    ramp = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)  # in the simulation's precision
    interior = np.ones(sim.Nx_padded, dtype=bool)
    for d in range(sim.ndim):
        shape = [1] * sim.ndim
        shape[d] = sim.Nx_padded[d]
        index = np.arange(sim.Nx_padded[d], dtype=sim.np_dtype).reshape(shape)
        interior &= (index >= 1) & (index <= sim.Nx[d] - 2)
        for side in (0, 1):
            if 2 * d + side in faces:
                wall = 1 if side == 0 else sim.Nx[d] - 2
                ramp = np.maximum(ramp, np.clip(1.0 - np.abs(index - wall) / width, 0.0, 1.0))
    return ramp * interior


def sponge(sim: Simulation, indicator, width: int, beta: float,
           faces: Sequence[int] | None = None) -> np.ndarray:
    """A damping field ramping quadratically to `beta` (the decay per step,
    d dt / 2m) at the lined walls over `width` nodes, zero inside
    (cuwave/boundary.py:238-269): d = 2 beta ramp^2 / (dt minv). Set it as
    `Simulation.damping`. It costs losslessness: `superposition_sensitivity`
    refuses it and `reconstruction_sensitivity` records the nodes it cannot run
    back."""
    from .wave import as_numpy

    if beta < 0.0:
        raise ValueError(f"beta must be non-negative: {beta}")
    faces = _layer_faces(sim, width, faces)
    minv = sim.inverse_inertia(np.asarray(as_numpy(indicator), dtype=sim.np_dtype))
    ramp = _layer_ramp(sim, width, faces)
    return np.ascontiguousarray((2.0 * beta * ramp**2 / (sim.dt * minv)).astype(sim.np_dtype))


def pad_for_sponge(Nx, dx, thickness: float, faces: Sequence[int] | None = None):
    """Grow a region of interest `Nx` (ghosts included) by a sponge of physical
    `thickness` behind `faces` (cuwave/boundary.py:196-235). Returns the grown
    extent, the layer width in nodes (off the finest axis, so no face is thinner
    than asked), the region's physical origin per axis, and the index tuple of its
    interior nodes in the grown grid."""
    if thickness <= 0.0:
        raise ValueError(f"thickness must be positive: {thickness}")
    faces = _canonical_faces(len(Nx), faces)
    width = round(thickness / min(dx))
    if width < 1:
        raise ValueError(f"thickness {thickness} is under one node at dx {min(dx)}")
    pads = [tuple(width if 2 * d + side in faces else 0 for side in (0, 1))
            for d in range(len(Nx))]
    grown = tuple(n + lo + hi for n, (lo, hi) in zip(Nx, pads))
    origin = tuple(lo * h for (lo, _), h in zip(pads, dx))
    region = tuple(slice(lo + 1, n - 1 - hi) for (lo, hi), n in zip(pads, grown))
    return grown, width, origin, region
