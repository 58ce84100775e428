"""Cell-assembled elasticity: `AnisotropicElasticWave`, its device engine and drivers.

The collocated sibling of the staggered `ElasticWave` (cuwave/anisotropic.py:
1-12): every displacement component on the nodes, the operator gathered cell
by cell from one cell's stencil -B^T C B, with C any symmetric Voigt matrix
scaled on the cell by the harmonic mean of the design field over its corners.
That makes the operator exactly symmetric for a varying material and leaves a
traction-free face the natural condition of the interior-cell sum; a clamped
face holds its wall node through a zeroed inertia. Order 2 only.

A step is one launch (`cw_aniso_step`); the drivers play the part of the
reference's closures over a batch of shots, the shots innermost:

- `simulate_shots` / `simulate`: the forward march (cuwave/wave.py:1071-1163);
- `sensitivity`: the stored-history adjoint of cuwave/sensitivity.py:482-497,
  the adjoint stepped by the step itself (the operator is symmetric) and the
  densities of each forward triplet against it from `cw_aniso_gradient`, then
  chained onto the nodes (cuwave/anisotropic.py:457-506);
- `reconstruction_sensitivity`: the same with the forward field rebuilt by the
  reverse march instead of stored.

The host-side parametrization (the one-cell stencil, the lumped inertia, the
harmonic cell means, the chain rule) runs in numpy in the simulation's
precision, as the reference computes it. A nodal `damping` field (a sponge)
damps the update as the reference's USE_DAMPING build does; the stored-history
`sensitivity` takes it, the reverse march does not. Not built here: a pml, the
blocked launch tier.
"""

from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from .boundary import Clamped, Traction, faces_with
from .elastic import PAIRS, ShotFields, _require, adjoint_signal
from .wave import MAX_BATCH, Simulation, Source, as_numpy, shot_lanes


def corner_bits(ndim: int) -> list[tuple[int, ...]]:
    """The 2^ndim corners of a cell, corner c with bit d set on the plus side of axis d
    (cuwave/anisotropic.py:58-60)."""
    return [tuple((c >> d) & 1 for d in range(ndim)) for c in range(2**ndim)]


def voigt(ndim: int, lame: float, shear: float, plane: str = "strain") -> np.ndarray:
    """The isotropic Voigt stiffness, plane strain or stress in 2D (cuwave/elastic.py:57-75)."""
    if ndim == 1:
        return np.array([[lame + 2.0 * shear]])
    if ndim == 2:
        lam = lame if plane == "strain" else 2.0 * lame * shear / (lame + 2.0 * shear)
        return np.array([[lam + 2.0 * shear, lam, 0.0], [lam, lam + 2.0 * shear, 0.0],
                         [0.0, 0.0, shear]])
    C = np.zeros((6, 6))
    C[:3, :3] = lame
    for i in range(3):
        C[i, i] = lame + 2.0 * shear
        C[3 + i, 3 + i] = shear
    return C


def cell_stencil(ndim: int, dx, C) -> np.ndarray:
    """One cell's stiffness sum_q w_q B_q^T C B_q over the 2^ndim Gauss points, ordered
    (corner, component) (cuwave/anisotropic.py:63-119).

    B is the symmetric gradient of the bilinear (trilinear) corner functions at a
    quadrature point, the shear rows engineering strains; the two-point Gauss rule
    per axis is what keeps the checkerboard out of the null space.
    """
    dx = np.asarray(dx, dtype=float)
    corners = corner_bits(ndim)
    n = len(corners) * ndim
    x, w = np.polynomial.legendre.leggauss(2)
    x = 0.5 * (x + 1.0)  # on [0, 1]
    w = 0.5 * w * float(np.prod(dx)) ** (1.0 / ndim)
    rows = PAIRS[ndim]
    K = np.zeros((n, n))
    for point in itertools.product(range(2), repeat=ndim):
        weight = float(np.prod([w[q] for q in point]))
        shape = [(1.0 - x[q], x[q]) for q in point]  # the 1D functions at the point
        grad = np.zeros((len(corners), ndim))
        for m, corner in enumerate(corners):
            for k in range(ndim):
                g = (1.0 if corner[k] else -1.0) / dx[k]
                for a in range(ndim):
                    if a != k:
                        g *= shape[a][corner[a]]
                grad[m, k] = g
        B = np.zeros((len(rows), n))
        for r, (k, l) in enumerate(rows):
            for m in range(len(corners)):
                if k == l:
                    B[r, m * ndim + k] += grad[m, k]
                else:
                    B[r, m * ndim + l] += grad[m, k]
                    B[r, m * ndim + k] += grad[m, l]
        K += weight * (B.T @ C @ B)
    return K


@dataclass
class AnisotropicElasticWave(Simulation):
    """Cell-assembled elasticity, gamma scaling density and stiffness alike
    (cuwave/anisotropic.py:157-240). `C` overrides the isotropic Voigt matrix at
    gamma = 1."""

    density: float = None
    wavespeed_p: float = None
    wavespeed_s: float = None
    plane: str = "strain"
    C: np.ndarray | None = None

    default_boundary = Traction
    face_conditions = (Traction, Clamped)
    accepts_damping = True  # a nodal damping field, as a sponge builds it

    def __post_init__(self) -> None:
        super().__post_init__()
        if None in (self.density, self.wavespeed_p, self.wavespeed_s):
            raise ValueError("AnisotropicElasticWave requires density, wavespeed_p, wavespeed_s")
        if self.plane not in ("strain", "stress"):
            raise ValueError(f"plane must be strain or stress: {self.plane}")
        if self.ndim == 3 and self.plane != "strain":
            raise ValueError("plane stress is a 2D reduction, not a 3D one")
        if self.space_order != 2:
            raise ValueError(f"the cell gather costs (2r)**(2 ndim) per node, so order "
                             f"{self.space_order} belongs to the staggered ElasticWave")
        if self.wavespeed_s >= self.wavespeed_p:
            raise ValueError(f"wavespeed_s must be below wavespeed_p: "
                             f"{self.wavespeed_s} >= {self.wavespeed_p}")
        if self.pml is not None:
            raise NotImplementedError("a pml on the cell-assembled elastic family is not part "
                                      "of this build")
        nvoigt = len(PAIRS[self.ndim])
        if self.C is not None:
            self.C = np.asarray(self.C, dtype=float)
            if self.C.shape != (nvoigt, nvoigt):
                raise ValueError(f"C must be ({nvoigt}, {nvoigt}): {self.C.shape}")
            if not np.allclose(self.C, self.C.T):
                raise ValueError("C must be symmetric")
        self.field_shape = (self.ncomp, *self.Nx_padded)

    @property
    def ncomp(self) -> int:
        return self.ndim

    @property
    def interior(self) -> tuple[slice, ...]:
        return (slice(None), *(slice(0, n) for n in self.Nx))

    @property
    def lame(self) -> float:
        return self.density * (self.wavespeed_p**2 - 2.0 * self.wavespeed_s**2)

    @property
    def shear(self) -> float:
        return self.density * self.wavespeed_s**2

    def stiffness_matrix(self) -> np.ndarray:
        """The Voigt stiffness at gamma = 1: `C`, or the isotropic one."""
        if self.C is not None:
            return self.C
        return voigt(self.ndim, self.lame, self.shear, self.plane)

    def stencil_table(self) -> np.ndarray:
        """One cell's B^T C B (the step subtracts it) in the simulation's precision."""
        return cell_stencil(self.ndim, self.dx, self.stiffness_matrix()).astype(self.np_dtype)

    def flatten_indices(self, position) -> np.ndarray:
        """(1 + ndim, num) indices, the component row first, as int32 flat indices."""
        position = np.asarray(as_numpy(position), dtype=np.int64)
        if position.ndim != 2 or position.shape[0] != self.node_rows:
            raise ValueError(f"position needs {self.node_rows} rows, not {position.shape}")
        if self.ncomp > 1 and (np.any(position[0] < 0) or np.any(position[0] >= self.ncomp)):
            raise ValueError(f"a component row outside 0 .. {self.ncomp - 1}")
        rows = position[-self.ndim:]
        lin = position[0] * self.comp_stride if self.ncomp > 1 else np.zeros(position.shape[1],
                                                                              dtype=np.int64)
        for d in range(self.ndim):
            if np.any(rows[d] < 0) or np.any(rows[d] >= self.Nx[d]):
                raise ValueError(f"a position on axis {d} lies outside 0 .. {self.Nx[d] - 1}")
            lin = lin + rows[d] * self.strides[d]
        return lin.astype(np.int32)

    # ----------------------------------------------------------- materials
    def cell_weights(self) -> np.ndarray:
        """W, halved once per wall a node sits on, the axes in sequence (cuwave/wave.py:336-344)."""
        w = np.ones(self.Nx_padded, dtype=self.np_dtype)
        for d in range(self.ndim):
            for index in (1, self.Nx[d] - 2):
                face = [slice(None)] * self.ndim
                face[d] = index
                w[tuple(face)] *= 0.5
        return w

    def inverse_inertia(self, gamma) -> np.ndarray:
        """The lumped 1 / (gamma rho0 V W) (cuwave/anisotropic.py:260-264), what a sponge
        scales its damping by."""
        volume = float(np.prod(self.dx))
        mass = (np.asarray(gamma, dtype=self.np_dtype) * self.np_dtype(self.density * volume)
                * self.cell_weights())
        return (1.0 / np.maximum(mass, np.finfo(self.np_dtype).tiny)).astype(self.np_dtype)

    def held_inverse_inertia(self, gamma: np.ndarray) -> np.ndarray:
        """The inverse inertia the step reads: zero on a clamped wall node, which holds it
        (cuwave/anisotropic.py:266-272)."""
        minv = self.inverse_inertia(gamma)
        for face in faces_with(self, Clamped):
            wall = [slice(None)] * self.ndim
            wall[face // 2] = 1 if face % 2 == 0 else self.Nx[face // 2] - 2
            minv[tuple(wall)] = 0.0
        return minv

    def cell_average(self, gamma: np.ndarray) -> np.ndarray:
        """The harmonic mean of gamma over a cell's corners, at its low corner
        (cuwave/anisotropic.py:137-154)."""
        out = np.zeros(self.Nx_padded, dtype=self.np_dtype)
        safe = np.maximum(gamma, np.finfo(self.np_dtype).tiny)
        inner = tuple(slice(0, n - 1) for n in self.Nx)
        for corner in corner_bits(self.ndim):
            shifted = tuple(slice(b, n - 1 + b) for b, n in zip(corner, self.Nx))
            out[inner] += 1.0 / safe[shifted]
        out[inner] = self.np_dtype(2.0**self.ndim) / out[inner]
        return out

    def indicator_shape(self) -> tuple[int, ...]:
        return self.Nx_padded

    def build_materials(self, indicator) -> dict:
        self.check_masks()
        gamma = np.ascontiguousarray(as_numpy(indicator), dtype=self.np_dtype)
        if gamma.shape != self.indicator_shape():
            raise ValueError(f"indicator of shape {gamma.shape}, expected {self.indicator_shape()}")
        mat = {"gamma": gamma, "minv": self.held_inverse_inertia(gamma),
               "cell": self.cell_average(gamma), "table": self.stencil_table()}
        if self.damping is not None:
            mat["damping"] = self.damping
        return mat

    def excitation_weights(self, mat: dict, lin) -> np.ndarray:
        """dt^2 / inertia at the node of each point, over 1 + beta where damped
        (cuwave/anisotropic.py:290-301)."""
        node = np.asarray(lin, dtype=np.int64) % self.comp_stride
        weight = mat["minv"].ravel()[node]
        if self.damping is not None:
            beta = 0.5 * weight * mat["damping"].ravel()[node] * self.dt
            weight = weight / (1.0 + beta)
        return (self.np_dtype(self.dt**2 * self.source_factor()) * weight).astype(self.np_dtype)

    def adjoint_weights(self, sensors) -> np.ndarray:
        """Ones: the lumped inertia already carries W (cuwave/anisotropic.py:529-531)."""
        return np.ones(np.asarray(as_numpy(sensors)).shape[1], dtype=self.np_dtype)

    def parametrization_jacobian(self, indicator=None):
        return 1.0, 1.0

    def step_factors(self) -> list:
        """dt^2; the spacing sits in the stencil table."""
        return [self.np_dtype(self.dt**2)] * self.ndim

    def source_factor(self) -> float:
        return 1.0

    def mass_factor(self) -> float:
        """rho0 V / dt^2: d(mass)/d(gamma) before the W the finalization applies."""
        return self.density * float(np.prod(self.dx)) / self.dt**2

    def abi_grid(self, nbatch: int) -> _abi.AnisoGrid:
        g = _abi.AnisoGrid()
        g.ndim, g.nbatch = self.ndim, int(nbatch)
        g.lanes = shot_lanes(nbatch)
        g.dtype = 0 if self.precision == "float32" else 1
        for d in range(self.ndim):
            g.n[d], g.np[d] = self.Nx[d], self.Nx_padded[d]
        g.step = float(self.np_dtype(self.dt**2))
        return g


class AnisoEngine(ShotFields):
    """The step of one cell-assembled simulation bound to its material, for `nbatch`
    shots: fields (ncomp, *Nx_padded, lanes). The point lists are the elastic
    family's (`cw_aniso_points` runs the same kernel)."""

    points_entry = "cw_aniso_points"

    def __init__(self, sim: AnisotropicElasticWave, mat: dict, nbatch: int):
        super().__init__(sim, nbatch)
        self.keep = {k: torch.as_tensor(mat[k], device=self.device)
                     for k in ("table", "minv", "cell", "damping") if k in mat}
        a = self.args = _abi.AnisoStepArgs()
        a.grid = self.grid
        a.table, a.minv, a.cell = (_abi.ptr(self.keep[k]) for k in ("table", "minv", "cell"))
        a.damping, a.dt = _abi.ptr(self.keep.get("damping")), float(sim.dt)

    def step(self, u0, u1, u2):
        """u2 of the step (u2 may be u0) (cw_aniso_step)."""
        a = self.args
        a.u0, a.u1, a.u2 = _abi.ptr(u0), _abi.ptr(u1), _abi.ptr(u2)
        _abi.call("cw_aniso_step", ctypes.byref(a), _abi.stream())
        return u2

    def pair(self, adjoint: bool = False, with_pml: bool = True) -> "AnisoEngine":
        """The step as the elastic engine's pair offers it (no pml here)."""
        return self

    def gradient_fields(self) -> dict:
        sim = self.sim
        return {k: torch.zeros(sim.Nx_padded, dtype=sim.dtype, device=self.device)
                for k in ("mass", "cell")}

    def gradient(self, grads, u0, u1, u2, l1) -> None:
        """Both densities of the forward triplet against the adjoint l1 (cw_aniso_gradient)."""
        a = _abi.AnisoGradientArgs()
        a.grid, a.table = self.grid, _abi.ptr(self.keep["table"])
        a.g_mass, a.g_cell = _abi.ptr(grads["mass"]), _abi.ptr(grads["cell"])
        a.u0, a.u1, a.u2, a.l1 = _abi.ptr(u0), _abi.ptr(u1), _abi.ptr(u2), _abi.ptr(l1)
        a.mass_factor = float(self.sim.np_dtype(self.sim.mass_factor()))
        _abi.call("cw_aniso_gradient", ctypes.byref(a), _abi.stream())


def finalize_gradients(sim: AnisotropicElasticWave, g_mass: np.ndarray, g_cell: np.ndarray,
                       mat: dict) -> dict:
    """The mass density weighted by W; the cell density spread onto the corners through
    d(harmonic mean)/d(corner) = (mean / corner)^2 / 2^ndim (cuwave/anisotropic.py:
    468-485, cuwave/kernels/anisotropic_sensitivity.cu:204-224)."""
    mass = g_mass * sim.cell_weights()
    gamma, cell = mat["gamma"], mat["cell"]
    stiff = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)
    inner = tuple(slice(1, n - 1) for n in sim.Nx)
    acc = np.zeros(tuple(n - 2 for n in sim.Nx), dtype=sim.np_dtype)
    for c in range(2**sim.ndim):
        # the cells in the kernel's block order, the one below the node first: cell c
        # has its low corner at the node - corner, corner[d] = 1 - bit d of c
        corner = [1 - ((c >> d) & 1) for d in range(sim.ndim)]
        low = tuple(slice(1 - b, n - 1 - b) for b, n in zip(corner, sim.Nx))
        ok = np.ones(acc.shape, dtype=bool)
        for d, (b, n) in enumerate(zip(corner, sim.Nx)):
            a = np.arange(1, n - 1) - b
            shape = [1] * sim.ndim
            shape[d] = n - 2
            ok = ok & ((a >= 1) & (a <= n - 3)).reshape(shape)
        ratio = cell[low] / gamma[inner]
        acc = np.where(ok, acc + g_cell[low] * ratio * ratio, acc)
    stiff[inner] += sim.np_dtype(1.0 / 2.0**sim.ndim) * acc
    return {"mass": mass, "stiff": stiff}


def simulate_shots(sim: AnisotropicElasticWave, sources, indicator, sensors=None):
    """Every shot of `sources` forward in one batch: the (B, ncomp, *Nx) final fields
    and, with `sensors`, the (B, N, count) records."""
    sources = list(sources)
    _require(sim, sources, sensors)
    mat = sim.build_materials(indicator)
    eng = AnisoEngine(sim, mat, len(sources))
    src, signal = eng.sources(mat, sources)
    U = eng.levels(2)
    rec = record = None
    if sensors is not None:
        rec = eng.sensors(sensors)
        record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    for t in range(sim.N):
        u = eng.step(U[t % 2], U[1 - t % 2], U[t % 2])
        src.inject(u, signal, t)
        if rec is not None:
            rec.record(u, record, t)
    field = eng.shots(U[(sim.N - 1) % 2])[(slice(None), *sim.interior)].clone()
    return (field, eng.shots(record)) if sensors is not None else field


def simulate(sim: AnisotropicElasticWave, source: Source, indicator, sensors=None):
    out = simulate_shots(sim, [source], indicator, sensors)
    if isinstance(out, tuple):
        return tuple(o[0] for o in out)
    return out[0]


def sensitivity(sim: AnisotropicElasticWave, source: Source, indicator, sensors, objective):
    """Cost and d(cost)/d(mass, stiff) over a stored history (cuwave/sensitivity.py:
    430-497): N + 2 forward levels; then per level the adjoint step, its load, and the
    densities of the forward triplet around it against lambda."""
    _require(sim, [source], sensors)
    mat = sim.build_materials(indicator)
    eng = AnisoEngine(sim, mat, 1)
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
    P = eng.levels(2)
    N = sim.N
    for m in range(N):
        p = eng.step(P[m % 2], P[1 - m % 2], P[m % 2])  # lambda^n, n = N - 1 - m
        adj.inject(p, adjoint, m)
        eng.gradient(grads, V[N - 1 - m], V[N - m], V[N + 1 - m], p)
    out = finalize_gradients(sim, grads["mass"].cpu().numpy(), grads["cell"].cpu().numpy(), mat)
    return cost, {k: torch.as_tensor(v, device=eng.device) for k, v in out.items()}, um, {}


def _reverse_sweep(sim: AnisotropicElasticWave, mat: dict, sources, sensors, objectives):
    """Every shot of `sources` forward in one batch, then one adjoint sweep with the
    forward field rebuilt by the reverse march: with neither a pml nor damping every
    node is reversible, u^(t-2) = -u^t + 2 u^(t-1) + dt^2 minv force(u^(t-1)) + the
    source of step t (the step with u0 and u2 swapped), one level ahead of the triplet
    the gradient pairs with lambda, so no strip is kept. Every shot's densities are
    summed over the lanes in the kernels. `objectives(records)` maps the (B, N, count)
    records to (costs, (B, N, count) dJ/du). Returns (costs, grads, records, drift)."""
    if sim.damping is not None:
        raise NotImplementedError("a damped field does not run backwards: use sensitivity, "
                                  "which stores the history")
    eng = AnisoEngine(sim, mat, len(sources))
    src, signal = eng.sources(mat, sources)
    rec = eng.sensors(sensors)
    record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    U = eng.levels(3)
    N = sim.N
    for t in range(N):
        u = eng.step(U[t % 3], U[(t + 1) % 3], U[(t + 2) % 3])
        src.inject(u, signal, t)
        rec.record(u, record, t)
    records = eng.shots(record)
    costs, dphi = objectives(records)
    adjoint = eng.lane_table(adjoint_signal(sim, dphi, sensors))
    lin = sim.flatten_indices(sensors)
    adj = eng.sensors(sensors, weight=sim.excitation_weights(mat, lin))
    grads = eng.gradient_fields()
    P = eng.levels(2)
    # u^t sits in slot (t + 2) % 3: the march leaves hi = u^(N-1), mid = u^(N-2)
    # and lo = u^(N-3)
    hi, mid, lo = U[(N + 1) % 3], U[N % 3], U[(N + 2) % 3]
    scale = float(torch.max(torch.abs(hi)))
    for m in range(N):
        p = eng.step(P[m % 2], P[1 - m % 2], P[m % 2])  # lambda^n, n = N - 1 - m
        adj.inject(p, adjoint, m)
        eng.gradient(grads, lo, mid, hi, p)  # u^(N-3-m), u^(N-2-m), u^(N-1-m)
        if m < N - 1:
            # u^(N-4-m) over the top level's slot, from step N - 2 - m
            src.inject(eng.step(mid, lo, hi), signal, N - 2 - m)
            hi, mid, lo = mid, lo, hi
    drift = float(torch.max(torch.abs(torch.stack([lo, mid])))) / scale if scale > 0 else 0.0
    return costs, grads, records, drift


def _stored_sweep(sim: AnisotropicElasticWave, mat: dict, sources, sensors, objectives):
    """As `_reverse_sweep`, the forward field stored for all N + 2 levels instead of
    rebuilt: what a damped field, which does not run backwards, needs."""
    eng = AnisoEngine(sim, mat, len(sources))
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
    P = eng.levels(2)
    for m in range(N):
        p = eng.step(P[m % 2], P[1 - m % 2], P[m % 2])
        adj.inject(p, adjoint, m)
        eng.gradient(grads, V[N - 1 - m], V[N - m], V[N + 1 - m], p)
    return costs, grads, records, 0.0


def reconstruction_sensitivity(sim: AnisotropicElasticWave, source: Source, indicator, sensors,
                               objective, strip_dtype: str | None = None,
                               checkpoints: int | None = None):
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
    out = finalize_gradients(sim, grads["mass"].cpu().numpy(), grads["cell"].cpu().numpy(), mat)
    device = records.device
    return (costs[0], {k: torch.as_tensor(v, device=device) for k, v in out.items()}, records[0],
            {"strip": 0, "drift": drift})


def misfit_gradient(sim: AnisotropicElasticWave, sources, indicator, sensors, observed,
                    objective, batch: int | None = None, checkpoints: int | None = None):
    """The misfit summed over the shots and its derivative with respect to gamma
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
    out = finalize_gradients(sim, total["mass"].cpu().numpy(), total["cell"].cpu().numpy(), mat)
    jm, js = sim.parametrization_jacobian(indicator)
    return cost, torch.as_tensor(jm * out["mass"] + js * out["stiff"], device=device)
