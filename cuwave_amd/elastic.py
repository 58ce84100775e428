"""Isotropic elasticity on a staggered grid: `ElasticWave`, its device engine and drivers.

Displacement component `c` sits half a node up axis `c`, so the normal strains
land on the nodes and the shear strain (k, l) half a node up both axes
(cuwave/elastic.py:1-10). The design field gamma scales density and stiffness
alike at fixed wave speeds. The walls close with the staggered
summation-by-parts rows of `sbp.py`, whose norms weigh the component and
stress points, and a traction-free or clamped condition per face.

A step is the pair of launches `cw_elastic_step` makes (stresses, then the
update), a convolutional PML folded into both; the drivers here play the part
of the reference's closures (`_define_pair`, `define_adjoint_gradient`,
`define_gradient`, cuwave/elastic.py:345-762) over a batch of shots, the shots
innermost in every per-shot array as in the pressure path:

- `simulate_shots` / `simulate`: the forward march (cuwave/wave.py:1071-1163);
- `sensitivity`: the stored-history adjoint, the adjoint pair accumulating
  both densities (cuwave/sensitivity.py:384-497, cuwave/elastic.py:436-488);
- `reconstruction_sensitivity`: the forward field rebuilt backwards by the
  PML-free pair from a strip of the irreversible layer points, with treeverse
  checkpoints (cuwave/sensitivity.py:500-750), the densities by the gradient
  kernel (cuwave/elastic.py:731-762);
- `misfit_gradient`: every shot of a batch in one sweep, summed.

The host-side parametrization (point inertia, excitation weights, the norms,
the chain rule of `finalize_gradients`) runs in numpy in the simulation's
precision, as the reference computes it, so what the kernels read is the
reference's to the bit.
"""

from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from .boundary import Clamped, Traction, face_mask, pml_nodes, pml_profiles, pml_widths
from .sbp import closure_band, closure_norms, closure_tables
from .stencils import staggered_weights
from .wave import MAX_BATCH, Simulation, Source, as_numpy, shot_lanes

# the Voigt rows: the normal ones, then one per axis pair (cuwave/wave.py PAIRS)
PAIRS = {1: ((0, 0),), 2: ((0, 0), (1, 1), (0, 1)),
         3: ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))}


@dataclass
class ElasticWave(Simulation):
    """Staggered isotropic elasticity parametrized by a density-scaling indicator gamma.

    Both wave speeds are held fixed and gamma scales the density, so
    C = gamma rho0 C0 scales inertia and stiffness alike
    (cuwave/elastic.py:154-345). `plane` picks plane strain or plane stress in
    2D. A `relaxation` (the reference's generalized Zener body) is not part of
    this build and raises.
    """

    density: float = None
    wavespeed_p: float = None
    wavespeed_s: float = None
    plane: str = "strain"
    relaxation: object | None = None

    default_boundary = Traction
    face_conditions = (Traction, Clamped)
    gradient_names = ("mass", "stiff")
    pml_shifted = True
    pml_walls = (Clamped,)

    def __post_init__(self) -> None:
        super().__post_init__()
        if None in (self.density, self.wavespeed_p, self.wavespeed_s):
            raise ValueError("ElasticWave requires density, wavespeed_p, wavespeed_s")
        if self.plane not in ("strain", "stress"):
            raise ValueError(f"plane must be strain or stress: {self.plane}")
        if self.ndim == 3 and self.plane != "strain":
            raise ValueError("plane stress is a 2D reduction, not a 3D one")
        if self.wavespeed_s >= self.wavespeed_p:
            raise ValueError(f"wavespeed_s must be below wavespeed_p: "
                             f"{self.wavespeed_s} >= {self.wavespeed_p}")
        if self.relaxation is not None:
            raise NotImplementedError("a relaxation (attenuating material) is not part of this build")
        least = 2 * closure_band(self.space_order, self.stencil) + 3
        if min(self.Nx) < least:
            raise ValueError(f"Nx {self.Nx} at order {self.space_order} needs {least} nodes an "
                             f"axis for both closures")
        self.field_shape = (self.ncomp, *self.Nx_padded)

    # ------------------------------------------------------------ geometry
    @property
    def ncomp(self) -> int:
        return self.ndim

    @property
    def nvoigt(self) -> int:
        return len(PAIRS[self.ndim])

    @property
    def npairs(self) -> int:
        return self.nvoigt - self.ndim

    @property
    def radius(self) -> int:
        return self.space_order // 2

    @property
    def reach(self) -> int:
        """The strain reach compounds with the divergence one: 2 radius - 1 nodes."""
        return 2 * self.radius - 1

    @property
    def pml_margin(self) -> int:
        """Undamped layer nodes before the front: a stress the step corrects is read this far."""
        return self.radius - 1

    @property
    def interior(self) -> tuple[slice, ...]:
        return (slice(None), *(slice(0, n) for n in self.Nx))

    @property
    def shear(self) -> float:
        """Second Lame parameter rho0 c_s^2."""
        return self.density * self.wavespeed_s**2

    @property
    def lame(self) -> float:
        """First Lame parameter rho0 (c_p^2 - 2 c_s^2), reduced under plane stress."""
        lam = self.density * (self.wavespeed_p**2 - 2.0 * self.wavespeed_s**2)
        if self.plane == "stress":
            return 2.0 * lam * self.shear / (lam + 2.0 * self.shear)
        return lam

    def flatten_indices(self, position) -> np.ndarray:
        """(node_rows, num) indices as int32 flat indices of the field, the component
        row folded in as `component * comp_stride` (cuwave/wave.py:364-382)."""
        position = np.asarray(as_numpy(position), dtype=np.int64)
        if position.ndim != 2 or position.shape[0] != self.node_rows:
            raise ValueError(f"position needs {self.node_rows} rows for ncomp={self.ncomp} in "
                             f"{self.ndim}D, not {position.shape}")
        rows = position[-self.ndim:]
        lin = np.zeros(position.shape[1], dtype=np.int64)
        for d in range(self.ndim):
            if np.any(rows[d] < 0) or np.any(rows[d] >= self.Nx[d]):
                raise ValueError(f"a position on axis {d} lies outside 0 .. {self.Nx[d] - 1}")
            lin += rows[d] * self.strides[d]
        if self.ncomp > 1:
            if np.any(position[0] < 0) or np.any(position[0] >= self.ncomp):
                raise ValueError(f"a component row outside 0 .. {self.ncomp - 1}")
            lin += position[0] * self.comp_stride
        return lin.astype(np.int32)

    def with_components(self, nodes: np.ndarray) -> np.ndarray:
        """Spatial `nodes` once per component, the component row prepended."""
        if self.ncomp == 1:
            return nodes
        tiled = np.tile(nodes, (1, self.ncomp))
        row = np.repeat(np.arange(self.ncomp, dtype=np.int32), nodes.shape[1])
        return np.ascontiguousarray(np.concatenate((row[None, :], tiled), axis=0))

    # --------------------------------------------------------------- norms
    def _norm_lines(self, half: tuple[int, ...]) -> list[np.ndarray]:
        """Per axis the closure norms along its padded extent, 1 on the ghosts and
        past the walls; the half points' along the axes `half` (cuwave/elastic.py:290-303)."""
        lines = []
        for d, n in enumerate(self.Nx):
            node, half_points = closure_norms(self.space_order, n - 3)
            line = np.ones(self.Nx_padded[d])
            if d in half:
                line[1:n - 2] = half_points
            else:
                line[1:n - 1] = node
            lines.append(line)
        return lines

    def point_weights(self, half: tuple[int, ...]) -> np.ndarray:
        """The norm W of the points on the half grid of the axes `half`, nodes elsewhere."""
        w = np.ones(self.Nx_padded, dtype=self.np_dtype)
        for d, line in enumerate(self._norm_lines(half)):
            shape = [1] * self.ndim
            shape[d] = -1
            w *= np.asarray(line.reshape(shape), dtype=self.np_dtype)
        return w

    def _point_norms(self, lin: np.ndarray) -> np.ndarray:
        """W on the component points `lin`: the half points' along its own axis."""
        component, node = np.divmod(np.asarray(lin, dtype=np.int64), self.comp_stride)
        nodal = self._norm_lines(())
        half = self._norm_lines(tuple(range(self.ndim)))
        weight = np.ones(len(node))
        for d in range(self.ndim):
            a = node // self.strides[d] % self.Nx_padded[d]
            weight *= np.where(component == d, half[d][a], nodal[d][a])
        return np.asarray(weight, dtype=self.np_dtype)

    # ------------------------------------------------------------ material
    def indicator_shape(self) -> tuple[int, ...]:
        return self.Nx_padded

    def build_materials(self, indicator) -> dict:
        """The design field on the host and on the GPU: the kernels average it onto the
        component and stress points (cuwave/elastic.py:305-311)."""
        self.check_masks()  # masks set after construction raise here
        gamma = np.ascontiguousarray(as_numpy(indicator), dtype=self.np_dtype)
        if gamma.shape != self.indicator_shape():
            raise ValueError(f"indicator of shape {gamma.shape}, expected {self.indicator_shape()}")
        device = _abi.require_device()
        return {"gamma_host": gamma, "gamma": torch.as_tensor(gamma, device=device)}

    def point_inverse_inertia(self, gamma: np.ndarray, lin) -> np.ndarray:
        """1 / (rho0 W mean gamma) on the component points `lin`, zero on a clamped wall
        across the component's axis (cuwave/elastic.py:313-332)."""
        lin = np.asarray(lin, dtype=np.int64)
        component, node = lin // self.comp_stride, lin % self.comp_stride
        rows = [node // s % n for s, n in zip(self.strides, self.Nx_padded)]
        step = np.asarray(self.strides, dtype=np.int64)[component]
        flat = gamma.ravel()
        mean = 0.5 * (flat[node] + flat[node + step])
        mass = self.density * self._point_norms(lin) * mean
        minv = 1.0 / np.maximum(mass, np.finfo(self.np_dtype).tiny)
        for d, pair in enumerate(self.boundary):
            for side, condition in enumerate(pair):
                if condition == Clamped:
                    wall = 1 if side == 0 else self.Nx[d] - 2
                    minv = np.where((rows[d] == wall) & (component != d), 0.0, minv)
        return minv.astype(self.np_dtype)

    def excitation_weights(self, mat: dict, lin) -> np.ndarray:
        """Source weights dt^2 / (inertia V) (cuwave/elastic.py:672-686)."""
        volume = self.np_dtype(1.0 / float(np.prod(self.dx)))
        weight = self.point_inverse_inertia(mat["gamma_host"], lin) * volume
        return (self.np_dtype(self.dt**2 * self.source_factor()) * weight).astype(self.np_dtype)

    def adjoint_weights(self, sensors) -> np.ndarray:
        """1 / V: dJ/du is nodal, so it undoes the volume the source weights divide by."""
        volume = self.np_dtype(1.0 / float(np.prod(self.dx)))
        return np.full(np.asarray(as_numpy(sensors)).shape[1], volume, dtype=self.np_dtype)

    def parametrization_jacobian(self, indicator=None):
        """Gamma scales inertia and stiffness alike, so both derivatives are 1."""
        return 1.0, 1.0

    def step_factors(self) -> list:
        """Per-axis update factors dt^2 / dx, the strain carrying the other 1 / dx."""
        return [self.np_dtype(self.dt**2 / d) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0

    def adjoint_mass_factor(self) -> float:
        return 1.0 / self.dt**2

    # --------------------------------------------------------------- ABI
    def abi_grid(self, nbatch: int) -> _abi.ElasticGrid:
        g = _abi.ElasticGrid()
        g.ndim, g.order, g.nbatch = self.ndim, self.space_order, int(nbatch)
        g.lanes = shot_lanes(nbatch)
        g.dtype = 0 if self.precision == "float32" else 1
        for d in range(self.ndim):
            g.n[d], g.np[d] = self.Nx[d], self.Nx_padded[d]
        g.clamped = face_mask(self, Clamped)
        tables = closure_tables(self.space_order, self.stencil)
        for k, c in enumerate(tables["counts"]):
            g.closure[k] = int(c)
        for k, w in enumerate(staggered_weights(self.radius, self.stencil)):
            g.stag[k] = float(self.np_dtype(w))
        g.lame = float(self.np_dtype(self.lame))
        g.shear = float(self.np_dtype(self.shear))
        g.density = float(self.np_dtype(self.density))
        for d in range(self.ndim):
            g.inv_dx[d] = float(self.np_dtype(1.0 / self.dx[d]))
            g.step[d] = float(self.step_factors()[d])
        return g


# ----------------------------------------------------------------- the engine
class _Points:
    """Points of a field on the device, entries on one node grouped into a run in their
    given order (cw_elastic_point_args)."""

    def __init__(self, eng, lin, shot, col, strides, weight=None):
        lin = np.asarray(lin, dtype=np.int64)
        count = len(lin)
        order = np.argsort(lin, kind="stable")
        starts = np.flatnonzero(np.diff(lin[order], prepend=-1)) if count else np.zeros(0, int)
        starts = np.append(starts, count).astype(np.int32)
        dev = eng.device

        def up(a, dtype):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)

        shot = np.broadcast_to(np.asarray(shot, dtype=np.int32), (count,))
        self.keep = [up(starts, torch.int32), up(lin[order], torch.int32),
                     up(shot[order], torch.int32), up(np.asarray(col, np.int32)[order], torch.int32)]
        a = self.args = _abi.ElasticPointArgs()
        a.starts, a.lin, a.shot, a.col = (_abi.ptr(t) for t in self.keep)
        a.runs = len(starts) - 1
        a.row_stride, a.col_stride, a.shot_stride = (int(s) for s in strides)
        if weight is not None:
            w = up(np.asarray(weight, dtype=eng.sim.np_dtype)[order], eng.sim.dtype)
            self.keep.append(w)
            a.weight = _abi.ptr(w)
        self.eng = eng
        self.count = count

    def _call(self, kind, field, table, row, g_mass=None, u1=None, mf=0.0):
        if not self.count:
            return
        a = self.args
        a.kind, a.row = kind, int(row)
        a.field, a.table = _abi.ptr(field), _abi.ptr(table)
        a.g_mass, a.u1, a.mass_factor = _abi.ptr(g_mass), _abi.ptr(u1), float(mf)
        _abi.call(self.eng.points_entry, ctypes.byref(self.eng.grid), ctypes.byref(a),
                  _abi.stream())

    def inject(self, field, table, row):
        self._call(0, field, table, row)

    def record(self, field, table, row):
        self._call(1, field, table, row)

    def replay(self, field, table, row):
        self._call(2, field, table, row)

    def adjoint_inject(self, field, table, row, g_mass, u1, mf):
        self._call(3, field, table, row, g_mass, u1, mf)


class _Pair:
    """The stress and update launches of a step, a PML folded into both (cw_elastic_step).

    The adjoint pair (`adjoint`) keeps its strain memory twice and alternates the
    block it reads each step (cuwave/elastic.py:355-384, 597-612).
    """

    def __init__(self, eng, adjoint: bool = False, with_pml: bool = True):
        self.eng = eng
        sim = eng.sim
        a = self.args = _abi.ElasticStepArgs()
        a.grid, a.mat = eng.grid, eng.mat
        a.sigma = _abi.ptr(eng.sigma)
        self.memory = []
        self.adjoint = adjoint and with_pml and sim.pml is not None
        if with_pml and sim.pml is not None:
            rows = (3 if self.adjoint else 2) * sim.ndim
            for d, lo, hi, profile, plane in eng.slabs:
                aux = torch.zeros((rows, (lo + hi) * plane, eng.lanes), dtype=sim.dtype,
                                  device=eng.device)
                self.memory.append(aux)
                a.pml.lo[d], a.pml.hi[d] = lo, hi
                a.pml.profile[d] = _abi.ptr(profile)
                a.pml.memory[d] = _abi.ptr(aux)
        a.pml.phase = 1 if self.adjoint else 0

    def step(self, u0, u1, u2, grads=None, v1=None, mf=0.0):
        a = self.args
        a.u0, a.u1, a.u2 = _abi.ptr(u0), _abi.ptr(u1), _abi.ptr(u2)
        if grads is not None:
            a.g_mass, a.g_normal = _abi.ptr(grads["mass"]), _abi.ptr(grads["normal"])
            a.g_shear = _abi.ptr(grads.get("shear"))
            a.v1, a.mass_factor = _abi.ptr(v1), float(mf)
        else:
            a.g_mass = a.g_normal = a.g_shear = a.v1 = None
        _abi.call("cw_elastic_step", ctypes.byref(a), _abi.stream())
        if self.adjoint:  # the strain memory swaps its two row blocks
            a.pml.phase = 3 - a.pml.phase
        return u2


class ShotFields:
    """What the engines of the vector families share: the device, the batch of shots
    and its lanes, the ABI grid, the time levels and the point lists. Per-shot arrays
    end in the lanes axis: a field level is (ncomp, *Nx_padded, lanes). A family names
    the entry point its point lists run through (`points_entry`)."""

    points_entry = "cw_elastic_points"

    def __init__(self, sim, nbatch: int):
        self.device = _abi.require_device()
        self.sim, self.nbatch = sim, int(nbatch)
        self.grid = sim.abi_grid(self.nbatch)
        self.lanes = self.grid.lanes

    def levels(self, slots: int) -> torch.Tensor:
        return torch.zeros((slots, *self.sim.field_shape, self.lanes), dtype=self.sim.dtype,
                           device=self.device)

    def shots(self, t: torch.Tensor) -> torch.Tensor:
        """The (nbatch, ...) view of a per-shot array whose last axis is the lanes axis."""
        return torch.movedim(t, -1, 0)[: self.nbatch]

    def lane_table(self, per_shot: torch.Tensor) -> torch.Tensor:
        out = torch.zeros((*per_shot.shape[1:], self.lanes), dtype=self.sim.dtype,
                          device=self.device)
        out[..., : self.nbatch] = torch.movedim(per_shot.to(self.device, self.sim.dtype), 0, -1)
        return out

    # --------------------------------------------------------------- points
    def sources(self, mat: dict, sources) -> tuple[_Points, torch.Tensor]:
        """Every shot's source points and one (N, columns) signal table, shot b's columns its own."""
        sim = self.sim
        lins, shots, cols, weights, signals = [], [], [], [], []
        column = 0
        for b, source in enumerate(sources):
            lin = sim.flatten_indices(source.position)
            sig = as_numpy(source.signal)
            if sig.ndim != 2 or sig.shape[0] < sim.N or sig.shape[1] != len(lin):
                raise ValueError(f"shot {b}: signal needs ({sim.N}, {len(lin)}), not {sig.shape}")
            lins.append(lin)
            shots.append(np.full(len(lin), b, dtype=np.int32))
            cols.append(column + np.arange(len(lin), dtype=np.int32))
            weights.append(sim.excitation_weights(mat, lin))
            signals.append(np.asarray(sig[: sim.N], dtype=sim.np_dtype))
            column += len(lin)
        table = np.ascontiguousarray(np.concatenate(signals, axis=1), dtype=sim.np_dtype)
        pts = _Points(self, np.concatenate(lins), np.concatenate(shots), np.concatenate(cols),
                      (column, 1, 0), np.concatenate(weights))
        return pts, torch.as_tensor(table, device=self.device)

    def sensors(self, sensors, weight=None) -> _Points:
        """Points every shot shares, column k the k-th, over an (N, count, lanes) table."""
        lin = self.sim.flatten_indices(sensors)
        count = len(lin)
        return _Points(self, lin, -1, np.arange(count, dtype=np.int32),
                       (count * self.lanes, self.lanes, 1), weight)


class ElasticEngine(ShotFields):
    """The kernels of one elastic simulation bound to its material, for `nbatch` shots:
    the stress scratch is (nvoigt, *Nx_padded, lanes)."""

    def __init__(self, sim: ElasticWave, mat: dict, nbatch: int):
        super().__init__(sim, nbatch)
        tables = closure_tables(sim.space_order, sim.stencil)
        rows = np.concatenate((tables["node"].ravel(), tables["half"].ravel()))
        self.rows = torch.as_tensor(rows.astype(sim.np_dtype), device=self.device)
        self.gamma = mat["gamma"]
        self.prep = torch.zeros((sim.nvoigt, *sim.Nx_padded), dtype=sim.dtype, device=self.device)
        self.mat = _abi.ElasticMaterial(_abi.ptr(self.gamma), _abi.ptr(self.prep),
                                        _abi.ptr(self.rows))
        _abi.call("cw_elastic_prepare", ctypes.byref(self.grid), ctypes.byref(self.mat),
                  _abi.stream())
        self.sigma = torch.zeros((sim.nvoigt, *sim.Nx_padded, self.lanes), dtype=sim.dtype,
                                 device=self.device)
        self.slabs = []  # (axis, lo, hi, profile, plane)
        if sim.pml is not None:
            total = int(np.prod(sim.Nx_padded))
            for d, ((lo, hi), prof) in enumerate(zip(pml_widths(sim), pml_profiles(sim))):
                if prof is not None:
                    self.slabs.append((d, lo, hi, torch.as_tensor(prof, device=self.device),
                                       total // sim.Nx_padded[d]))

    def pair(self, adjoint: bool = False, with_pml: bool = True) -> _Pair:
        return _Pair(self, adjoint, with_pml)

    def gradient_fields(self) -> dict:
        sim = self.sim
        g = {"mass": torch.zeros((sim.ncomp, *sim.Nx_padded), dtype=sim.dtype, device=self.device),
             "normal": torch.zeros(sim.Nx_padded, dtype=sim.dtype, device=self.device)}
        if sim.ndim > 1:
            g["shear"] = torch.zeros((sim.npairs, *sim.Nx_padded), dtype=sim.dtype,
                                     device=self.device)
        return g

    def gradient(self, grads, u0, u1, u2, l1, mf) -> None:
        """Both densities of the triplet against the adjoint l1 (cw_elastic_gradient)."""
        a = _abi.ElasticGradientArgs()
        a.grid, a.rows = self.grid, _abi.ptr(self.rows)
        a.g_mass, a.g_normal = _abi.ptr(grads["mass"]), _abi.ptr(grads["normal"])
        a.g_shear = _abi.ptr(grads.get("shear"))
        a.u0, a.u1, a.u2, a.l1 = _abi.ptr(u0), _abi.ptr(u1), _abi.ptr(u2), _abi.ptr(l1)
        a.mass_factor = float(mf)
        _abi.call("cw_elastic_gradient", ctypes.byref(a), _abi.stream())

# ------------------------------------------------------------------ gradients
def _point_average_adjoint(sim: ElasticWave, density: np.ndarray, c: int) -> np.ndarray:
    """The transpose of the mean over the two nodes component c sits between."""
    out = 0.5 * density
    to = [slice(None)] * sim.ndim
    fro = [slice(None)] * sim.ndim
    to[c], fro[c] = slice(1, None), slice(0, -1)
    out[tuple(to)] += 0.5 * density[tuple(fro)]
    return out


def _pair_average(sim: ElasticWave, field: np.ndarray, axes) -> np.ndarray:
    """The harmonic mean of `field` over the four nodes around a pair point of `axes`."""
    out = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)
    safe = np.maximum(field, np.finfo(sim.np_dtype).tiny)
    inner = tuple(slice(0, n - 1) if d in axes else slice(0, n) for d, n in enumerate(sim.Nx))
    for bits in itertools.product((0, 1), repeat=2):
        shifted = tuple(slice(bits[axes.index(d)], n - 1 + bits[axes.index(d)])
                        if d in axes else slice(0, n) for d, n in enumerate(sim.Nx))
        out[inner] += 1.0 / safe[shifted]
    out[inner] = 4.0 / out[inner]
    return out


def _pair_average_adjoint(sim: ElasticWave, density, field, axes) -> np.ndarray:
    """The transpose of the harmonic pair mean's derivative, scattered onto the four nodes."""
    safe = np.maximum(field, np.finfo(sim.np_dtype).tiny)
    mean = _pair_average(sim, field, axes)
    scattered = density * mean * mean * 0.25
    out = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)
    for bits in itertools.product((0, 1), repeat=2):
        to = [slice(None)] * sim.ndim
        fro = [slice(None)] * sim.ndim
        for d, b in zip(axes, bits):
            to[d], fro[d] = slice(b, None), slice(0, -b if b else None)
        out[tuple(to)] += scattered[tuple(fro)] / safe[tuple(to)] ** 2
    return out


def finalize_gradients(sim: ElasticWave, grads: dict, gamma: np.ndarray) -> dict:
    """The point densities chained through the averages onto the nodal design field
    (cuwave/elastic.py:716-729, cuwave/wave.py:386-475), on the host."""
    host = {k: v.cpu().numpy() for k, v in grads.items()}
    g_mass = np.zeros(sim.Nx_padded, dtype=sim.np_dtype)
    for c in range(sim.ncomp):
        density = host["mass"][c] * sim.point_weights((c,)) * sim.density
        g_mass += _point_average_adjoint(sim, density, c)
    g_stiff = host["normal"] * sim.point_weights(())
    for p, axes in enumerate(PAIRS[sim.ndim][sim.ndim:]):
        density = host["shear"][p] * sim.point_weights(axes)
        g_stiff += _pair_average_adjoint(sim, density, gamma, axes)
    return {"mass": g_mass, "stiff": g_stiff}


def _on_device(grads: dict, device) -> dict:
    return {k: torch.as_tensor(np.ascontiguousarray(v), device=device) for k, v in grads.items()}


def reconstruction_nodes(sim: ElasticWave):
    """The strip points a reverse march replays, one per component, and the reversible
    nodes (cuwave/sensitivity.py:160-216)."""
    from .sensitivity import reconstruction_nodes as nodes

    strip, valid = nodes(sim)
    return sim.with_components(strip), valid


def adjoint_signal(sim: ElasticWave, dphi: torch.Tensor, sensors) -> torch.Tensor:
    """dphi reversed in time over the adjoint weights (cuwave/sensitivity.py:272-292),
    for a (B, N, count) batch."""
    w = torch.as_tensor(sim.adjoint_weights(sensors), device=dphi.device)
    return (torch.flip(dphi, (1,)) / w).contiguous()


def _require(sim: ElasticWave, sources, sensors) -> None:
    for s in sources:
        sim.require_interior(s.position, "source")
    if sensors is not None:
        sim.require_interior(sensors, "sensor")


# -------------------------------------------------------------------- drivers
def simulate_shots(sim: ElasticWave, sources, indicator, sensors=None):
    """Every shot of `sources` forward in one batch: the (B, ncomp, *Nx) final fields
    (ncomp 1: (B, *Nx)) and, with `sensors`, the (B, N, count) records."""
    sources = list(sources)
    _require(sim, sources, sensors)
    mat = sim.build_materials(indicator)
    eng = ElasticEngine(sim, mat, len(sources))
    pair = eng.pair()
    src, signal = eng.sources(mat, sources)
    U = eng.levels(2)
    rec = record = None
    if sensors is not None:
        rec = eng.sensors(sensors)
        record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    for t in range(sim.N):
        # slot t % 2 takes u^t over u^(t - 2)
        u = pair.step(U[t % 2], U[1 - t % 2], U[t % 2])
        src.inject(u, signal, t)
        if rec is not None:
            rec.record(u, record, t)
    field = eng.shots(U[(sim.N - 1) % 2])[(slice(None), *sim.interior)].clone()
    if sim.ncomp == 1:
        field = field[:, 0]
    return (field, eng.shots(record)) if sensors is not None else field


def simulate(sim: ElasticWave, source: Source, indicator, sensors=None):
    """The final interior field and, with `sensors`, the (N, count) record."""
    out = simulate_shots(sim, [source], indicator, sensors)
    if isinstance(out, tuple):
        return tuple(o[0] for o in out)
    return out[0]


def sensitivity(sim: ElasticWave, source: Source, indicator, sensors, objective):
    """Cost and d(cost)/d(mass, stiff) by the adjoint pair over a stored history
    (cuwave/sensitivity.py:384-497): N + 2 forward levels, then the adjoint pair
    accumulating both densities, the layer then frozen out of the gradient."""
    _require(sim, [source], sensors)
    mat = sim.build_materials(indicator)
    eng = ElasticEngine(sim, mat, 1)
    pair = eng.pair()
    src, signal = eng.sources(mat, [source])
    rec = eng.sensors(sensors)
    record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    V = eng.levels(sim.N + 2)
    for t in range(sim.N):
        u = pair.step(V[t], V[t + 1], V[t + 2])
        src.inject(u, signal, t)
        rec.record(u, record, t)
    um = eng.shots(record)[0]
    cost, dphi = objective(um)
    dphi = torch.as_tensor(dphi, dtype=sim.dtype, device=eng.device)
    adjoint = eng.lane_table(adjoint_signal(sim, dphi[None], sensors))
    lin = sim.flatten_indices(sensors)
    adj = eng.sensors(sensors, weight=sim.excitation_weights(mat, lin))
    grads = eng.gradient_fields()
    back = eng.pair(adjoint=True)
    mf = sim.np_dtype(sim.adjoint_mass_factor())
    P = eng.levels(2)
    for m in range(sim.N):
        v1 = V[sim.N + 1 - m]
        l0, l1 = P[m % 2], P[1 - m % 2]
        back.step(l0, l1, l0, grads, v1, mf)
        adj.adjoint_inject(l0, adjoint, m, grads["mass"], v1, mf)
    out = finalize_gradients(sim, grads, mat["gamma_host"])
    if sim.pml is not None:
        layer = pml_nodes(sim)
        for f in out.values():
            f[..., layer] = 0.0
    return cost, _on_device(out, eng.device), um, {}


class _Reconstruction:
    """The strip-replayed reverse march over a batch of shots, in treeverse order.

    Three forward slots hold the triplet the gradient reads; a forward step t
    writes u^t into slot (t + 2) % 3 (cuwave/sensitivity.py:500-750).

    Set `timings` to a list to have every forward segment and every segment's
    reverse levels bracketed by HIP events on the launch stream: entries
    (kind, steps, start, end), kind "forward" or "backward".
    """

    timings: list | None = None

    def _timed(self, kind: str, steps: int):
        import contextlib

        timings = _Reconstruction.timings
        if timings is None or steps <= 0:
            return contextlib.nullcontext()

        @contextlib.contextmanager
        def bracket():
            stream = torch.cuda.current_stream()
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record(stream)
            yield
            end.record(stream)
            timings.append((kind, steps, start, end))

        return bracket()

    def __init__(self, sim: ElasticWave, mat: dict, sources, sensors, checkpoints=None):
        self.sim, self.mat, self.sensors = sim, mat, sensors
        eng = self.eng = ElasticEngine(sim, mat, len(sources))
        self.fwd = eng.pair()
        self.back = eng.pair(with_pml=False)
        self.adjp = eng.pair(adjoint=True)
        self.grads = eng.gradient_fields()
        self.src, self.signal = eng.sources(mat, sources)
        self.rec = eng.sensors(sensors)
        lin = sim.flatten_indices(sensors)
        self.adj = eng.sensors(sensors, weight=sim.excitation_weights(mat, lin))
        self.U = eng.levels(3)
        snapshot = 2 * self.U[0].nbytes + sum(a.nbytes for a in self.fwd.memory)
        strip, self.valid = reconstruction_nodes(sim)
        self.count = strip.shape[1]
        self.strip = eng.sensors(strip) if self.count else None
        segment = sim.N
        if checkpoints is not None and self.count:
            row = 2 * self.count * np.dtype(sim.np_dtype).itemsize * eng.lanes
            segment = min(sim.N, max(2, snapshot // row))
        self.segment = segment
        self.checkpoints = checkpoints or 0
        self.starts = list(range(0, sim.N, segment)) + [sim.N]
        self.record = torch.zeros((sim.N, self.rec.count, eng.lanes), dtype=sim.dtype,
                                  device=eng.device)
        self.P = eng.levels(2)
        self.snapshots = {}
        self.recorded = 0
        self.steps = 0
        self.mf = sim.np_dtype(sim.adjoint_mass_factor())
        self.adjoint = None
        self.seed = None

    def slots(self, t):
        return self.U[t % 3], self.U[(t + 1) % 3]

    def triplet(self, m):
        N = self.sim.N
        return tuple(self.U[(N + i - m) % 3] for i in (2, 0, 1))

    def forward(self, t, record=False, buffer=None, row=0):
        u = self.fwd.step(self.U[t % 3], self.U[(t + 1) % 3], self.U[(t + 2) % 3])
        self.src.inject(u, self.signal, t)
        if record:
            self.rec.record(u, self.record, t)
        if buffer is not None:
            self.strip.record(u, buffer, row)

    def advance(self, a, b, buffer=None):
        first = a == self.recorded
        if first:
            self.recorded = b
        with self._timed("forward", b - a):
            for k in range(b - a):
                self.forward(a + k, first, buffer, 2 + k)
        self.steps += b - a

    def store(self, k):
        self.snapshots[k] = [f.clone() for f in (*self.slots(self.starts[k]), *self.fwd.memory)]

    def restore(self, k):
        fields = (*self.slots(self.starts[k]), *self.fwd.memory)
        if k == 0:
            for f in fields:
                f.zero_()
            return
        for f, saved in zip(fields, self.snapshots[k]):
            f.copy_(saved)

    def adjoint_level(self, m, row):
        l0, l1 = self.P[m % 2], self.P[1 - m % 2]
        self.adjp.step(l0, l1, l0)
        self.adj.inject(l0, self.adjoint, row)
        return l0

    def reverse(self, m, src_row, strip_row, buffer):
        a, b, c = self.triplet(m)
        c = self.back.step(b, a, c)
        self.src.inject(c, self.signal, src_row)
        if buffer is not None:
            self.strip.replay(c, buffer, strip_row)

    def _valid_norm(self, level: torch.Tensor) -> float:
        valid = torch.as_tensor(self.valid, device=self.eng.device)
        return float(torch.linalg.vector_norm(self.eng.shots(level) * valid))

    def leaf(self, k):
        sim, eng = self.sim, self.eng
        s, e = self.starts[k], self.starts[k + 1]
        buffer = None
        if self.count:
            buffer = torch.zeros((e - s + 2, self.count, eng.lanes), dtype=sim.dtype,
                                 device=eng.device)
            if s:
                for row, f in enumerate(self.slots(s)):
                    self.strip.record(f, buffer, row)
        self.advance(s, e, buffer)
        if self.adjoint is None:
            dphi = yield eng.shots(self.record)
            dphi = torch.as_tensor(dphi, dtype=sim.dtype, device=eng.device)
            self.adjoint = eng.lane_table(adjoint_signal(sim, dphi, self.sensors))
        m0 = sim.N - e
        if not s:
            self.seed = self._valid_norm(self.triplet(m0)[2])
        with self._timed("backward", e - s):
            eng.gradient(self.grads, *self.triplet(m0), self.adjoint_level(m0, m0), self.mf)
            for j in range(e - s - 1):
                m = m0 + j
                self.reverse(m, e - 2 - j, e - s - 2 - j, buffer)
                p = self.adjoint_level(m + 1, m0 + 1 + j)
                eng.gradient(self.grads, *self.triplet(m + 1), p, self.mf)

    def sweep(self):
        from .sensitivity import checkpoint_schedule

        for action, *segs in checkpoint_schedule(len(self.starts) - 1, self.checkpoints):
            if action == "advance":
                self.advance(self.starts[segs[0]], self.starts[segs[1]])
            elif action == "store":
                self.store(segs[0])
            elif action == "restore":
                self.restore(segs[0])
            elif action == "drop":
                del self.snapshots[segs[0]]
            else:
                yield from self.leaf(segs[0])

    def run(self, objectives):
        """Forward, objectives, adjoint; returns the per-shot costs. `objectives(record)`
        takes the (B, N, count) record and returns the costs and the derivative."""
        sweep = self.sweep()
        costs, dphi = objectives(next(sweep))
        try:
            sweep.send(dphi)
        except StopIteration:
            pass
        return costs

    def drift(self) -> float:
        """Norm of the rebuilt u^(-1), which is zero, over that of the top state."""
        rebuilt = self._valid_norm(self.triplet(self.sim.N - 1)[0])
        return rebuilt / self.seed if self.seed else float("inf")

    def gradients(self) -> dict:
        out = finalize_gradients(self.sim, self.grads, self.mat["gamma_host"])
        for f in out.values():
            f[~self.valid] = 0.0
        return out


def reconstruction_sensitivity(sim: ElasticWave, source: Source, indicator, sensors, objective,
                               strip_dtype: str | None = None, checkpoints: int | None = None):
    """Cost and gradients as `sensitivity`, storing the strip of irreversible layer
    points instead of the field (cuwave/sensitivity.py:500-750). `info` carries the
    `strip` point count, the `drift` of the rebuilt initial state, the steps per
    `segment`, the forward `sweeps` and the `valid` mask."""
    if strip_dtype is not None:
        raise NotImplementedError("the elastic strip is kept in the simulation's precision")
    if checkpoints is not None and checkpoints < 0:
        raise ValueError(f"checkpoints must be non-negative: {checkpoints}")
    _require(sim, [source], sensors)
    mat = sim.build_materials(indicator)
    run = _Reconstruction(sim, mat, [source], sensors, checkpoints)

    def objectives(record):
        cost, dphi = objective(record[0])
        return [cost], torch.as_tensor(dphi, dtype=sim.dtype, device=record.device)[None]

    costs = run.run(objectives)
    info = {"strip": run.count, "drift": run.drift(), "segment": run.segment,
            "sweeps": run.steps / sim.N, "valid": run.valid}
    return costs[0], _on_device(run.gradients(), run.eng.device), run.eng.shots(run.record)[0], info


def misfit_gradient(sim: ElasticWave, sources, indicator, sensors, observed, objective,
                    batch: int | None = None, checkpoints: int | None = None):
    """The misfit summed over the shots and its derivative with respect to gamma
    (cuwave/utils.py:410-459), `batch` shots a sweep (all by default, at most 64):
    one forward sweep records every shot, one reverse sweep sums every shot's
    densities in the kernels, and one chain rule finishes the sum."""
    sources = list(sources)
    nodes = sensors.nodes
    _require(sim, sources, nodes)
    step = min(batch or MAX_BATCH, MAX_BATCH)
    mat = sim.build_materials(indicator)
    cost, total, valid, device = 0.0, None, None, None
    for start in range(0, len(sources), step):
        chunk = sources[start:start + step]
        run = _Reconstruction(sim, mat, chunk, nodes, checkpoints)

        def objectives(record, start=start, chunk=chunk):
            costs, dphis = [], []
            for b in range(len(chunk)):
                c, dphi = sensors.objective(objective(observed[start + b]))(record[b])
                costs.append(c)
                dphis.append(torch.as_tensor(dphi, dtype=sim.dtype, device=record.device))
            return costs, torch.stack(dphis)

        cost += sum(run.run(objectives))
        if total is None:
            total, valid, device = run.grads, run.valid, run.eng.device
        else:
            for k in total:
                total[k] += run.grads[k]
    out = finalize_gradients(sim, total, mat["gamma_host"])
    jm, js = sim.parametrization_jacobian(indicator)
    gradient = jm * out["mass"] + js * out["stiff"]
    gradient[~valid] = 0.0
    return cost, torch.as_tensor(gradient, device=device)
