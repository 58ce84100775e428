"""The out-of-plane reductions of Maxwell's equations on the scalar kernels.

In 1D and 2D one field component is out of plane and the curl-curl form reduces
to a flux divergence, so both polarisations are `PressureWave`s with the two
coefficient fields assigned differently (cuwave/maxwell.py:72-160):

- `ElectricWave` carries E_z: the permittivity is its inertia and 1/mu its
  stiffness, so the design (which interpolates the permittivity) enters the
  inertia, with a constant Jacobian.
- `MagneticWave` carries H_z: mu is its inertia and 1/eps its stiffness, so the
  design enters the stiffness as 1/eps(gamma), with a nodal Jacobian.

Both step with the factors 2 dt^2 / dx^2 and an unscaled source.

`MaxwellWave` is the vector form 3D needs (cuwave/maxwell.py:162-708): the Yee
lattice written as -C^T nu C with a diagonal permittivity, electric component
c half a node up axis c and the curl of the pair (k, l) half a node up both.
A step is the pair of launches `cw_maxwell_step` makes (the curls, then the
update); `DielectricWave` interpolates the permittivity between two phases.
The drivers play the part of the reference's closures over a batch of shots,
the shots innermost in every per-shot array as in the other families:

- `simulate_shots` / `simulate`: the forward march (cuwave/wave.py:1071-1163);
- `sensitivity`: the stored-history adjoint, the adjoint step being the
  forward step itself (no pml breaks its symmetry) with the inertia density
  fused in (cuwave/maxwell.py:278-282, 400-445), chained through the harmonic
  point means onto the nodal permittivity (:601-613);
- `reconstruction_sensitivity`: the same adjoint with the forward field
  rebuilt level by level by the reverse march instead of stored.

The host-side parametrization (the point inertia, the excitation weights,
the chain rule) runs in numpy in the simulation's precision, as the reference
computes it. A nodal `damping` field (a sponge) damps the update as the
reference's USE_DAMPING build does; the stored-history `sensitivity` takes it,
the reverse march does not. A subclass with `magnetic = True` designs the
permeability too
(the reference's `USE_MAGNETIC` build): the curl then reads nu W per pair
point and the adjoint also accumulates the permeability density. Not built
here: a pml, damping, the fused and blocked launch tiers.
"""

from __future__ import annotations

import ctypes
import itertools
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from .boundary import Conductor, Magnetic, face_mask, faces_with
from .elastic import ShotFields, _require, adjoint_signal
from .scalar import PressureWave
from .stencils import staggered_weights
from .wave import MAX_BATCH, Simulation, Source, as_numpy, shot_lanes


@dataclass
class PolarizedWave(PressureWave):
    """Two-phase 2D Maxwell base: one out-of-plane component, gamma interpolating eps."""

    permittivity1: float = None  # gamma = 0
    permittivity2: float = None  # gamma = 1
    permeability: float = 1.0  # non-magnetic media: not designed

    def __post_init__(self) -> None:
        super().__post_init__()
        if None in (self.permittivity1, self.permittivity2):
            raise ValueError(f"{type(self).__name__} requires permittivity1 and permittivity2")
        if min(self.permittivity1, self.permittivity2, self.permeability) <= 0.0:
            raise ValueError(
                f"permittivity and permeability must be positive: {self.permittivity1}, "
                f"{self.permittivity2}, {self.permeability}")
        if self.ndim > 2:
            raise ValueError(f"an out-of-plane reduction is 1D or 2D, not {self.ndim}D")

    @property
    def light_speed(self) -> float:
        """The fastest speed on the grid, 1 / sqrt(min(eps) mu), which bounds dt."""
        return 1.0 / math.sqrt(min(self.permittivity1, self.permittivity2) * self.permeability)

    def permittivity(self, indicator):
        """eps1 + gamma (eps2 - eps1): linear in the design."""
        return self.permittivity1 + indicator * (self.permittivity2 - self.permittivity1)

    def _constant(self, value: float) -> np.ndarray:
        return np.full(self.Nx_padded, value, dtype=self.np_dtype)

    def step_factors(self) -> list:
        return [self.np_dtype(2.0 * self.dt**2 / d**2) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0


@dataclass
class ElectricWave(PolarizedWave):
    """E_z: the permittivity is the inertia, 1/mu the stiffness."""

    def parametrization(self, indicator):
        return self._constant(1.0 / self.permeability), 1.0 / self.permittivity(indicator)

    def parametrization_jacobian(self, indicator=None):
        return self.permittivity2 - self.permittivity1, 0.0


@dataclass
class MagneticWave(PolarizedWave):
    """H_z: mu is the inertia, 1/eps the stiffness (a nodal Jacobian)."""

    def parametrization(self, indicator):
        return 1.0 / self.permittivity(indicator), self._constant(1.0 / self.permeability)

    def parametrization_jacobian(self, indicator):
        ind = np.asarray(as_numpy(indicator), dtype=self.np_dtype)
        contrast = self.permittivity2 - self.permittivity1
        return 0.0, -contrast / self.permittivity(ind) ** 2


# ------------------------------------------------------------ vector Maxwell
@dataclass
class MaxwellWave(Simulation):
    """Curl-curl Maxwell on the Yee lattice (cuwave/maxwell.py:162-260).

    Component c of the electric field lives half a node up axis c and the curl
    of the pair (k, l) half a node up both of its axes. The operator is
    -C^T nu C, symmetric by construction and its own transpose at every order;
    the walls read images (odd across a conductor for a tangential component,
    the curl odd across a magnetic wall), and a conductor holds the tangential
    field on its wall node at zero through a zeroed inertia.
    """

    permeability: float = 1.0  # background mu, designed only where `magnetic` is set

    default_boundary = Conductor
    face_conditions = (Conductor, Magnetic)
    magnetic = False  # whether the permeability carries the design too
    accepts_damping = True  # a nodal conductivity-like field, as a sponge builds it

    def __post_init__(self) -> None:
        super().__post_init__()
        if self.ndim < 2:
            raise ValueError("a single in-plane component has no curl; use ElectricWave")
        if self.permeability <= 0.0:
            raise ValueError(f"permeability must be positive: {self.permeability}")
        if self.pml is not None:
            raise NotImplementedError("a pml on the vector Maxwell family is not part of this build")
        self.field_shape = (self.ncomp, *self.Nx_padded)

    @property
    def ncomp(self) -> int:
        return self.ndim

    @property
    def npairs(self) -> int:
        """Curl components, one per axis pair: 1 in 2D and 3 in 3D."""
        return self.ndim * (self.ndim - 1) // 2

    @property
    def radius(self) -> int:
        return self.space_order // 2

    @property
    def interior(self) -> tuple[slice, ...]:
        return (slice(None), *(slice(0, n) for n in self.Nx))

    def flatten_indices(self, position) -> np.ndarray:
        """(1 + ndim, num) indices, the component row first, as int32 flat indices
        (cuwave/wave.py:364-382)."""
        position = np.asarray(as_numpy(position), dtype=np.int64)
        if position.ndim != 2 or position.shape[0] != self.node_rows:
            raise ValueError(f"position needs {self.node_rows} rows (the component, then the "
                             f"axes), not {position.shape}")
        lin = position[0] * self.comp_stride
        if np.any(position[0] < 0) or np.any(position[0] >= self.ncomp):
            raise ValueError(f"a component row outside 0 .. {self.ncomp - 1}")
        for d in range(self.ndim):
            rows = position[1 + d]
            if np.any(rows < 0) or np.any(rows >= self.Nx[d]):
                raise ValueError(f"a position on axis {d} lies outside 0 .. {self.Nx[d] - 1}")
            lin = lin + rows * self.strides[d]
        return lin.astype(np.int32)

    def parametrization(self, indicator):
        """(nodal permittivity, nodal 1 / mu or None) of the design (a subclass defines it)."""
        raise NotImplementedError(f"{type(self).__name__} defines no parametrization")

    @property
    def pair_axes(self) -> list[tuple[int, int]]:
        """The axes of each curl row: (1, 2), (0, 2), (0, 1) in 3D, (0, 1) in 2D."""
        return [(0, 1)] if self.ndim == 2 else [(1, 2), (0, 2), (0, 1)]

    # ----------------------------------------------------------- materials
    def _wall_weights(self, axes) -> np.ndarray:
        """Cell weights of a point family staggered along `axes`: halved on the other
        walls (cuwave/wave.py:386-394)."""
        w = np.ones(self.Nx_padded, dtype=self.np_dtype)
        for d in set(range(self.ndim)) - set(axes):
            for index in (1, self.Nx[d] - 2):
                wall = [slice(None)] * self.ndim
                wall[d] = index
                w[tuple(wall)] *= 0.5
        return w

    def _point_average(self, field: np.ndarray, c: int) -> np.ndarray:
        out = np.zeros(self.Nx_padded, dtype=self.np_dtype)
        lo = tuple(slice(0, n - 1) if d == c else slice(0, n) for d, n in enumerate(self.Nx))
        hi = tuple(slice(1, n) if d == c else slice(0, n) for d, n in enumerate(self.Nx))
        out[lo] = 0.5 * (field[lo] + field[hi])
        return out

    def _point_harmonic(self, field: np.ndarray, c: int) -> np.ndarray:
        """The harmonic mean of the two nodes component c sits between
        (cuwave/wave.py:416-425)."""
        tiny = np.finfo(self.np_dtype).tiny
        mean = self._point_average(1.0 / np.maximum(field, tiny), c)
        return np.where(mean > 0.0, 1.0 / np.maximum(mean, tiny), 0.0).astype(self.np_dtype)

    def _pair_average(self, field: np.ndarray, axes) -> np.ndarray:
        """The harmonic mean of the four nodes a pair point straddles (cuwave/wave.py:436-457)."""
        out = np.zeros(self.Nx_padded, dtype=self.np_dtype)
        safe = np.maximum(field, np.finfo(self.np_dtype).tiny)
        inner = tuple(slice(0, n - 1) if d in axes else slice(0, n) for d, n in enumerate(self.Nx))
        for bits in itertools.product((0, 1), repeat=2):
            shifted = tuple(slice(bits[axes.index(d)], n - 1 + bits[axes.index(d)]) if d in axes
                            else slice(0, n) for d, n in enumerate(self.Nx))
            out[inner] += 1.0 / safe[shifted]
        out[inner] = self.np_dtype(4.0) / out[inner]
        return out

    def point_inverse_inertia(self, eps: np.ndarray) -> np.ndarray:
        """(ncomp, *Nx_padded): 1 / (W eps) on the component points, zero on a
        conductor's wall node for the tangential components (cuwave/maxwell.py:247-263)."""
        tiny = np.finfo(self.np_dtype).tiny
        minv = np.zeros((self.ncomp, *self.Nx_padded), dtype=self.np_dtype)
        for c in range(self.ncomp):
            mass = self._wall_weights((c,)) * self._point_harmonic(eps, c)
            minv[c] = 1.0 / np.maximum(mass, tiny)
        for face in faces_with(self, Conductor):
            wall = [slice(None)] * self.ndim
            wall[face // 2] = 1 if face % 2 == 0 else self.Nx[face // 2] - 2
            for c in range(self.ncomp):
                if c != face // 2:
                    minv[(c, *wall)] = 0.0
        return minv

    def inverse_inertia(self, indicator) -> np.ndarray:
        """Nodal 1 / (eps W), what a sponge scales its damping by (cuwave/maxwell.py:239-245)."""
        eps = np.asarray(self.parametrization(np.asarray(indicator, dtype=self.np_dtype))[0],
                         dtype=self.np_dtype)
        mass = eps * self.cell_weights()
        return (1.0 / np.maximum(mass, np.finfo(self.np_dtype).tiny)).astype(self.np_dtype)

    def indicator_shape(self) -> tuple[int, ...]:
        return self.Nx_padded

    def build_materials(self, indicator) -> dict:
        """The nodal permittivity and the point inverse inertia, on the host."""
        self.check_masks()
        gamma = np.ascontiguousarray(as_numpy(indicator), dtype=self.np_dtype)
        if gamma.shape != self.indicator_shape():
            raise ValueError(f"indicator of shape {gamma.shape}, expected {self.indicator_shape()}")
        eps, nu = self.parametrization(gamma)
        eps = np.ascontiguousarray(eps, dtype=self.np_dtype)
        mat = {"eps": eps, "minv": self.point_inverse_inertia(eps)}
        if self.damping is not None:
            mat["damping"] = self.damping
        if self.magnetic:
            # nu W of every pair point, the weight folded in (cuwave/maxwell.py:264-270)
            nu = np.ascontiguousarray(nu, dtype=self.np_dtype)
            mat["nu"] = nu
            mat["nu_pair"] = np.stack([self._wall_weights(axes) * self._pair_average(nu, axes)
                                       for axes in self.pair_axes]).astype(self.np_dtype)
        return mat

    def excitation_weights(self, mat: dict, lin) -> np.ndarray:
        """Source weights dt^2 / (eps V), over 1 + beta where damped (cuwave/maxwell.py:
        561-574)."""
        volume = self.np_dtype(1.0 / float(np.prod(self.dx)))
        lin = np.asarray(lin, dtype=np.int64)
        minv = mat["minv"].ravel()[lin]
        weight = minv * volume
        if self.damping is not None:
            beta = 0.5 * minv * mat["damping"].ravel()[lin % self.comp_stride] * self.dt
            weight = weight / (1.0 + beta)
        return (self.np_dtype(self.dt**2 * self.source_factor()) * weight).astype(self.np_dtype)

    def adjoint_weights(self, sensors) -> np.ndarray:
        """1 / V: dJ/du is nodal, so it undoes the volume the source weights divide by."""
        volume = self.np_dtype(1.0 / float(np.prod(self.dx)))
        return np.full(np.asarray(as_numpy(sensors)).shape[1], volume, dtype=self.np_dtype)

    def step_factors(self) -> list:
        """Per-axis update factors dt^2 / dx, the curl carrying the other 1 / dx."""
        return [self.np_dtype(self.dt**2 / d) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0

    def adjoint_mass_factor(self) -> float:
        return 1.0 / self.dt**2

    def abi_grid(self, nbatch: int) -> _abi.MaxwellGrid:
        g = _abi.MaxwellGrid()
        g.ndim, g.order, g.nbatch = self.ndim, self.space_order, int(nbatch)
        g.lanes = shot_lanes(nbatch)
        g.dtype = 0 if self.precision == "float32" else 1
        for d in range(self.ndim):
            g.n[d], g.np[d] = self.Nx[d], self.Nx_padded[d]
        g.conductor = face_mask(self, Conductor)
        for k, w in enumerate(staggered_weights(self.radius, self.stencil)):
            g.stag[k] = float(self.np_dtype(w))
        g.nu = float(self.np_dtype(1.0 / self.permeability))
        for d in range(self.ndim):
            g.inv_dx[d] = float(self.np_dtype(1.0 / self.dx[d]))
            g.step[d] = float(self.step_factors()[d])
        return g


@dataclass
class DielectricWave(MaxwellWave):
    """Two-phase dielectric, gamma interpolating the permittivity (cuwave/maxwell.py:676-708)."""

    permittivity1: float = None  # background, gamma = 0
    permittivity2: float = None  # design material, gamma = 1

    def __post_init__(self) -> None:
        super().__post_init__()
        if None in (self.permittivity1, self.permittivity2):
            raise ValueError("DielectricWave requires permittivity1 and permittivity2")
        if min(self.permittivity1, self.permittivity2) <= 0.0:
            raise ValueError(f"permittivity must be positive: {self.permittivity1}, "
                             f"{self.permittivity2}")

    @property
    def light_speed(self) -> float:
        """The fastest speed on the grid, 1 / sqrt(min(eps) mu), which bounds dt."""
        return 1.0 / math.sqrt(min(self.permittivity1, self.permittivity2) * self.permeability)

    def parametrization(self, indicator):
        """eps1 + gamma (eps2 - eps1) (cuwave/maxwell.py:65-67, 700-704); the permeability
        left uniform."""
        eps = self.permittivity1 + indicator * (self.permittivity2 - self.permittivity1)
        return eps.astype(self.np_dtype), None

    def parametrization_jacobian(self, indicator=None):
        """The permittivity is affine in gamma; the permeability does not depend on it."""
        return self.permittivity2 - self.permittivity1, 0.0


class MaxwellEngine(ShotFields):
    """The pair of one Maxwell simulation bound to its material, for `nbatch` shots.

    Per-shot arrays end in the lanes axis: a field level is (ncomp, *Nx_padded,
    lanes), the curl scratch (npairs, *Nx_padded, lanes). The point lists are
    the elastic family's (`cw_maxwell_points` runs the same kernel).
    """

    points_entry = "cw_maxwell_points"

    def __init__(self, sim: MaxwellWave, mat: dict, nbatch: int):
        super().__init__(sim, nbatch)
        self.minv = torch.as_tensor(mat["minv"], device=self.device)
        self.curl = torch.zeros((sim.npairs, *sim.Nx_padded, self.lanes), dtype=sim.dtype,
                                device=self.device)
        self.nu_pair = (torch.as_tensor(mat["nu_pair"], device=self.device) if sim.magnetic
                        else None)
        self.damping = (torch.as_tensor(mat["damping"], device=self.device)
                        if sim.damping is not None else None)
        self.args = _abi.MaxwellStepArgs()
        self.args.grid = self.grid
        self.args.minv, self.args.curl = _abi.ptr(self.minv), _abi.ptr(self.curl)
        self.args.nu_pair = _abi.ptr(self.nu_pair)
        self.args.damping, self.args.dt = _abi.ptr(self.damping), float(sim.dt)

    def step(self, u0, u1, u2, g_mass=None, v1=None, mf=0.0, g_nu=None):
        """u2 of the pair (u2 may be u0); with g_mass the adjoint step's inertia density of
        u1 against the forward v1, with g_nu too the permeability density (cw_maxwell_step)."""
        a = self.args
        a.u0, a.u1, a.u2 = _abi.ptr(u0), _abi.ptr(u1), _abi.ptr(u2)
        a.g_mass, a.v1, a.mass_factor = _abi.ptr(g_mass), _abi.ptr(v1), float(mf)
        a.g_nu = _abi.ptr(g_nu)
        _abi.call("cw_maxwell_step", ctypes.byref(a), _abi.stream())
        return u2

    def pair(self, adjoint: bool = False, with_pml: bool = True) -> "MaxwellEngine":
        """The step's launches as the elastic engine's pair offers them (no pml here)."""
        return self

    def gradient_fields(self) -> dict:
        sim = self.sim
        g = {"mass": torch.zeros((sim.ncomp, *sim.Nx_padded), dtype=sim.dtype,
                                 device=self.device)}
        if sim.magnetic:
            g["nu"] = torch.zeros((sim.npairs, *sim.Nx_padded), dtype=sim.dtype,
                                  device=self.device)
        return g

    def adjoint_step(self, l0, l1, grads, v1, mf):
        """The adjoint step over l0 with every density this family accumulates."""
        return self.step(l0, l1, l0, grads["mass"], v1, mf, grads.get("nu"))


def _point_average_adjoint(density: np.ndarray, c: int) -> np.ndarray:
    """Transpose of the point mean: half of `density` back onto each node it spans
    (cuwave/wave.py:405-413)."""
    out = 0.5 * density
    to = [slice(None)] * density.ndim
    fro = [slice(None)] * density.ndim
    to[c], fro[c] = slice(1, None), slice(0, -1)
    out[tuple(to)] += 0.5 * density[tuple(fro)]
    return out


def _pair_average_adjoint(sim: MaxwellWave, density, field, axes) -> np.ndarray:
    """Transpose of the pair mean: d(harmonic mean)/d(node) = (mean / node)^2 / 4
    (cuwave/wave.py:460-475)."""
    safe = np.maximum(field, np.finfo(sim.np_dtype).tiny)
    mean = sim._pair_average(field, axes)
    scattered = density * mean * mean * 0.25
    out = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)
    for bits in itertools.product((0, 1), repeat=2):
        to = [slice(None)] * sim.ndim
        fro = [slice(None)] * sim.ndim
        for d, b in zip(axes, bits):
            to[d], fro[d] = slice(b, None), slice(0, -b if b else None)
        out[tuple(to)] += scattered[tuple(fro)] / safe[tuple(to)] ** 2
    return out


def finalize_gradients(sim: MaxwellWave, grads: dict, mat: dict) -> dict:
    """The point densities chained through the harmonic means onto the nodal
    permittivity, and with a designed permeability the pair densities through the
    pair means onto the nodal 1 / mu (cuwave/maxwell.py:601-613, cuwave/wave.py:
    427-475); zero stiffness otherwise."""
    eps = mat["eps"]
    g_mass = grads["mass"].cpu().numpy()
    tiny = np.finfo(sim.np_dtype).tiny
    safe = np.maximum(eps, tiny)
    out = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)
    for c in range(sim.ncomp):
        density = g_mass[c] * sim._wall_weights((c,))
        mean = sim._point_harmonic(eps, c)
        out += _point_average_adjoint(density * mean * mean, c) / (safe * safe)
    stiff = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)
    if sim.magnetic:
        g_nu = grads["nu"].cpu().numpy()
        for p, axes in enumerate(sim.pair_axes):
            stiff += _pair_average_adjoint(sim, g_nu[p] * sim._wall_weights(axes), mat["nu"], axes)
    return {"mass": out, "stiff": stiff}


def simulate_shots(sim: MaxwellWave, sources, indicator, sensors=None):
    """Every shot of `sources` forward in one batch: the (B, ncomp, *Nx) final fields
    and, with `sensors`, the (B, N, count) records."""
    sources = list(sources)
    _require(sim, sources, sensors)
    mat = sim.build_materials(indicator)
    eng = MaxwellEngine(sim, mat, len(sources))
    src, signal = eng.sources(mat, sources)
    U = eng.levels(2)
    rec = record = None
    if sensors is not None:
        rec = eng.sensors(sensors)
        record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    for t in range(sim.N):
        # slot t % 2 takes u^t over u^(t - 2)
        u = eng.step(U[t % 2], U[1 - t % 2], U[t % 2])
        src.inject(u, signal, t)
        if rec is not None:
            rec.record(u, record, t)
    field = eng.shots(U[(sim.N - 1) % 2])[(slice(None), *sim.interior)].clone()
    return (field, eng.shots(record)) if sensors is not None else field


def simulate(sim: MaxwellWave, source: Source, indicator, sensors=None):
    """The final interior field and, with `sensors`, the (N, count) record."""
    out = simulate_shots(sim, [source], indicator, sensors)
    if isinstance(out, tuple):
        return tuple(o[0] for o in out)
    return out[0]


def sensitivity(sim: MaxwellWave, source: Source, indicator, sensors, objective):
    """Cost and d(cost)/d(mass, stiff) by the adjoint over a stored history
    (cuwave/sensitivity.py:384-497): N + 2 forward levels, then the adjoint steps
    accumulating the inertia density, chained onto the nodal permittivity."""
    _require(sim, [source], sensors)
    mat = sim.build_materials(indicator)
    eng = MaxwellEngine(sim, mat, 1)
    src, signal = eng.sources(mat, [source])
    rec = eng.sensors(sensors)
    record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    V = eng.levels(sim.N + 2)
    for t in range(sim.N):
        u = eng.step(V[t], V[t + 1], V[t + 2])
        src.inject(u, signal, t)
        rec.record(u, record, t)
    um = eng.shots(record)[0]
    cost, dphi = objective(um)
    dphi = torch.as_tensor(dphi, dtype=sim.dtype, device=eng.device)
    adjoint = eng.lane_table(adjoint_signal(sim, dphi[None], sensors))
    lin = sim.flatten_indices(sensors)
    adj = eng.sensors(sensors, weight=sim.excitation_weights(mat, lin))
    grads = eng.gradient_fields()
    mf = sim.np_dtype(sim.adjoint_mass_factor())
    P = eng.levels(2)
    for m in range(sim.N):
        v1 = V[sim.N + 1 - m]
        l0, l1 = P[m % 2], P[1 - m % 2]
        eng.adjoint_step(l0, l1, grads, v1, mf)
        adj.adjoint_inject(l0, adjoint, m, grads["mass"], v1, mf)
    out = finalize_gradients(sim, grads, mat)
    return cost, {k: torch.as_tensor(v, device=eng.device) for k, v in out.items()}, um, {}


def _reverse_sweep(sim: MaxwellWave, mat: dict, sources, sensors, objectives):
    """Every shot of `sources` forward in one batch, then one adjoint sweep with the
    forward field rebuilt level by level by the reverse march: with neither a pml nor
    damping every point is reversible, u^(n-2) = -u^n + 2 u^(n-1) + minv force(u^(n-1))
    + the source of step n (the step with u0 and u2 swapped), so no strip is kept.
    Every shot's densities are summed over the lanes in the kernels.
    `objectives(records)` maps the (B, N, count) records to (costs, (B, N, count) dJ/du).
    Returns (costs, grads, records, drift), `drift` the rebuilt initial state against
    the last level."""
    if sim.damping is not None:
        raise NotImplementedError("a damped field does not run backwards: use sensitivity, "
                                  "which stores the history")
    eng = MaxwellEngine(sim, mat, len(sources))
    src, signal = eng.sources(mat, sources)
    rec = eng.sensors(sensors)
    record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    U = eng.levels(2)
    N = sim.N
    for t in range(N):
        u = eng.step(U[t % 2], U[1 - t % 2], U[t % 2])
        src.inject(u, signal, t)
        rec.record(u, record, t)
    records = eng.shots(record)
    costs, dphi = objectives(records)
    adjoint = eng.lane_table(adjoint_signal(sim, dphi, sensors))
    lin = sim.flatten_indices(sensors)
    adj = eng.sensors(sensors, weight=sim.excitation_weights(mat, lin))
    grads = eng.gradient_fields()
    mf = sim.np_dtype(sim.adjoint_mass_factor())
    P = eng.levels(2)
    cur, prev = U[(N - 1) % 2], U[N % 2]  # u^(N-1), u^(N-2)
    scale = float(torch.max(torch.abs(cur)))
    for m in range(N):
        n = N - 1 - m
        l0, l1 = P[m % 2], P[1 - m % 2]
        eng.adjoint_step(l0, l1, grads, cur, mf)
        adj.adjoint_inject(l0, adjoint, m, grads["mass"], cur, mf)
        # u^(n-2) over u^n's slot, then the pair moves down a level
        src.inject(eng.step(cur, prev, cur), signal, n)
        cur, prev = prev, cur
    drift = float(torch.max(torch.abs(U))) / scale if scale > 0 else 0.0
    return costs, grads, records, drift


def _stored_sweep(sim: MaxwellWave, mat: dict, sources, sensors, objectives):
    """As `_reverse_sweep`, the forward field stored for all N + 2 levels instead of
    rebuilt: what a damped field, which does not run backwards, needs."""
    eng = MaxwellEngine(sim, mat, len(sources))
    src, signal = eng.sources(mat, sources)
    rec = eng.sensors(sensors)
    record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    N = sim.N
    V = eng.levels(N + 2)
    for t in range(N):
        u = eng.step(V[t], V[t + 1], V[t + 2])
        src.inject(u, signal, t)
        rec.record(u, record, t)
    records = eng.shots(record)
    costs, dphi = objectives(records)
    adjoint = eng.lane_table(adjoint_signal(sim, dphi, sensors))
    adj = eng.sensors(sensors, weight=sim.excitation_weights(mat, sim.flatten_indices(sensors)))
    grads = eng.gradient_fields()
    mf = sim.np_dtype(sim.adjoint_mass_factor())
    P = eng.levels(2)
    for m in range(N):
        l0, l1 = P[m % 2], P[1 - m % 2]
        eng.adjoint_step(l0, l1, grads, V[N + 1 - m], mf)
        adj.adjoint_inject(l0, adjoint, m, grads["mass"], V[N + 1 - m], mf)
    return costs, grads, records, 0.0


def reconstruction_sensitivity(sim: MaxwellWave, source: Source, indicator, sensors, objective,
                               strip_dtype: str | None = None, checkpoints: int | None = None):
    """Cost and gradients as `sensitivity`, the forward field rebuilt backwards instead of
    stored (cuwave/sensitivity.py:500-750) by the reverse march, which needs no strip here.
    `info` carries `strip` (0) and the `drift` of the rebuilt initial state."""
    if strip_dtype is not None or checkpoints is not None:
        raise NotImplementedError("nothing is stored on this family's reverse march: no strip "
                                  "dtype or checkpoints to choose")
    _require(sim, [source], sensors)
    mat = sim.build_materials(indicator)

    def objectives(records):
        cost, dphi = objective(records[0])
        return [cost], torch.as_tensor(dphi, dtype=sim.dtype, device=records.device)[None]

    costs, grads, records, drift = _reverse_sweep(sim, mat, [source], sensors, objectives)
    out = finalize_gradients(sim, grads, mat)
    device = records.device
    return (costs[0], {k: torch.as_tensor(v, device=device) for k, v in out.items()}, records[0],
            {"strip": 0, "drift": drift})


def misfit_gradient(sim: MaxwellWave, sources, indicator, sensors, observed, objective,
                    batch: int | None = None, checkpoints: int | None = None):
    """The misfit summed over the shots and its derivative with respect to the indicator
    (cuwave/utils.py:410-459), `batch` shots a sweep (all by default, at most 64): one
    forward sweep records every shot, one reverse sweep sums every shot's densities in
    the kernels, one chain rule finishes the sum. A damped field, which does not run
    backwards, keeps the forward history instead."""
    if checkpoints is not None:
        raise NotImplementedError("nothing is stored on this family's reverse march: no "
                                  "checkpoints to choose")
    sources = list(sources)
    nodes = sensors.nodes
    _require(sim, sources, nodes)
    step = min(batch or MAX_BATCH, MAX_BATCH)
    mat = sim.build_materials(indicator)
    cost, total, device = 0.0, None, None
    for start in range(0, len(sources), step):
        chunk = sources[start:start + step]

        def objectives(records, start=start, chunk=chunk):
            costs, dphis = [], []
            for b in range(len(chunk)):
                c, dphi = sensors.objective(objective(observed[start + b]))(records[b])
                costs.append(c)
                dphis.append(torch.as_tensor(dphi, dtype=sim.dtype, device=records.device))
            return costs, torch.stack(dphis)

        sweep = _stored_sweep if sim.damping is not None else _reverse_sweep
        costs, grads, records, _ = sweep(sim, mat, chunk, nodes, objectives)
        cost += sum(costs)
        device = records.device
        total = grads if total is None else {k: total[k] + grads[k] for k in total}
    out = finalize_gradients(sim, total, mat)
    jm, js = sim.parametrization_jacobian(as_numpy(indicator))
    return cost, torch.as_tensor(jm * out["mass"] + js * out["stiff"], device=device)
