"""The grid contract, the simulation description and the forward driver.

`Simulation` carries the reference's constructor arguments and grid contract
(cuwave/wave.py:479-631): `Nx` counts the ghost nodes, the interior is
`1 .. Nx[d] - 2`, and every field is allocated over `Nx_padded`, the fastest axis
rounded up to a multiple of 32 so each row starts on a 128-byte boundary. The
physics lives in subclasses (`scalar.py`) that supply the parametrization hooks.

Fields are torch tensors on the GPU. A batch of shots is a leading axis: one
launch steps every shot of the batch, the material read once for all of them.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import TYPE_CHECKING

import numpy as np
import torch

from . import _abi
from .boundary import Dirichlet, Neumann, canonical_boundary, face_mask, validate_pml
from .stencils import STENCILS, flux_row, stable_dt  # noqa: F401  (re-exported)

MAX_BATCH = 64  # shots one launch advances: one wave's lanes per node

if TYPE_CHECKING:
    from .boundary import Pml


def padded_shape(Nx: tuple[int, ...]) -> tuple[int, ...]:
    """Pad the fastest (last) axis of `Nx` up to a multiple of 32."""
    return (*Nx[:-1], ((Nx[-1] + 31) // 32) * 32)


def grid_coords(Nx: tuple[int, ...], dx: tuple[float, ...], dtype=np.float64) -> list[np.ndarray]:
    """Padded grid coordinates for shape `Nx` at spacing `dx`, node 1 at the origin."""
    axes = [(np.arange(n, dtype=dtype) - 1) * d for n, d in zip(padded_shape(Nx), dx)]
    return np.meshgrid(*axes, indexing="ij")


@dataclass
class Source:
    position: np.ndarray  # (ndim, num_sources) grid indices
    signal: np.ndarray | torch.Tensor  # (N, num_sources) time series


def as_numpy(a) -> np.ndarray:
    """A host copy of an array or tensor."""
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


@dataclass
class Simulation:
    """Grid, timestep and stencil shared by the wave equations.

    `threads` is accepted for the reference's signature (its CUDA block shape);
    the gfx950 kernels choose their own tiles. `damping` is a nodal field d over
    the padded grid (the pressure family; `boundary.sponge` builds one).
    `domain` and `dirichlet` are boolean nodal masks over the padded grid (the
    pressure family, forward runs only): the step updates the `domain` nodes
    (every node by default) less the `dirichlet` ones, which stay at zero, the
    cells at their walls graded (`domain_cells`). `devices` is not part of this
    build and raises if set.
    """

    Nx: tuple[int, ...]
    dx: tuple[float, ...]
    N: int
    dt: float
    threads: tuple[int, ...] = (4, 64)
    precision: str = "float32"
    space_order: int = 2
    stencil: str = "optimized"
    boundary: tuple = None
    damping: object | None = None
    domain: object | None = None
    dirichlet: object | None = None
    pml: Pml | None = None
    devices: tuple[int, ...] | None = None

    default_boundary = Neumann
    face_conditions = (Neumann, Dirichlet)  # the conditions a face may carry
    pml_shifted = False  # whether a pml needs a frequency shift (a static mode)
    pml_walls = None  # the conditions a lined face may end on, None for any
    accepts_damping = False  # whether the step takes a damping field
    accepts_domain = False  # whether the step takes `domain` / `dirichlet` masks

    @property
    def pml_margin(self) -> int:
        """Undamped layer nodes before the front (cuwave/boundary.py:335-336)."""
        return 0

    @property
    def ncomp(self) -> int:
        """Field components: one for a pressure field."""
        return 1

    @property
    def node_rows(self) -> int:
        """Rows of a position: the axes, plus a leading component row for a vector field."""
        return self.ndim + (self.ncomp > 1)

    def grid_rows(self, position) -> np.ndarray:
        """The `ndim` spatial rows of `position`, a leading component row dropped."""
        return np.asarray(as_numpy(position))[-self.ndim:]

    def __post_init__(self) -> None:
        self.Nx = tuple(int(n) for n in self.Nx)
        self.dx = tuple(float(d) for d in self.dx)
        self.ndim = len(self.Nx)
        if not 1 <= self.ndim <= 3:
            raise ValueError(f"1 to 3 axes, not {self.ndim}")
        if len(self.dx) != self.ndim:
            raise ValueError("dx needs one spacing per axis")
        self.Nx_padded = padded_shape(self.Nx)
        strides = [1] * self.ndim
        for d in range(self.ndim - 2, -1, -1):
            strides[d] = strides[d + 1] * self.Nx_padded[d + 1]
        self.strides = tuple(strides)
        if self.precision not in ("float32", "float64"):
            raise ValueError(f"precision is float32 or float64, not {self.precision!r}")
        self.np_dtype = np.float32 if self.precision == "float32" else np.float64
        self.dtype = torch.float32 if self.precision == "float32" else torch.float64
        self.boundary = canonical_boundary(self.boundary, self.ndim, self.default_boundary,
                                           self.face_conditions)
        if self.space_order % 2 != 0 or not 2 <= self.space_order <= 8:
            raise ValueError("space_order must be 2, 4, 6 or 8")
        if self.stencil not in STENCILS:
            raise ValueError(f"unknown stencil {self.stencil!r}, expected one of {STENCILS}")
        if min(self.Nx) < self.space_order // 2 + 3:
            raise ValueError(
                f"Nx {self.Nx} cannot mirror the taps of space_order {self.space_order}: "
                f"at least {self.space_order // 2 + 3} nodes an axis"
            )
        if self.damping is not None:
            if not self.accepts_damping:
                raise NotImplementedError(
                    f"{type(self).__name__} takes no damping field in this build: use pml")
            damping = np.asarray(as_numpy(self.damping), dtype=self.np_dtype)
            if damping.shape != tuple(self.Nx_padded):
                raise ValueError(f"damping of shape {damping.shape}, expected {self.Nx_padded}")
            if not np.all(np.isfinite(damping)):
                raise ValueError("damping must be finite")
            self.damping = np.ascontiguousarray(damping)
        self.check_masks()
        if self.devices is not None:
            raise NotImplementedError("the devices= axis split is not part of this build (one GPU)")
        self.comp_stride = int(np.prod(self.Nx_padded))
        if self.comp_stride >= 2**31:
            raise ValueError(f"{self.comp_stride} nodes overflow the int32 flat index")
        self.field_shape = self.Nx_padded
        if self.pml is not None:
            validate_pml(self)

    @property
    def masked(self) -> bool:
        """Whether `domain` or `dirichlet` takes nodes out of the stepped set."""
        return self.domain is not None or self.dirichlet is not None

    def check_masks(self) -> None:
        """Raise unless the masks are boolean nodal fields this equation can step
        (cuwave/wave.py:505-511, cuwave/boundary.py:294-295)."""
        if not self.masked:
            return
        if not self.accepts_domain:
            raise ValueError(f"{type(self).__name__} steps every node: unset domain")
        for name in ("domain", "dirichlet"):
            mask = getattr(self, name)
            if mask is None:
                continue
            mask = np.asarray(as_numpy(mask))
            if mask.dtype != np.bool_:
                raise ValueError(f"{name} must be a boolean mask, not {mask.dtype}")
            if mask.shape != tuple(self.Nx_padded):
                raise ValueError(f"{name} of shape {mask.shape}, expected {self.Nx_padded}")
        if self.pml is not None:
            raise ValueError("a pml takes neither a damping field nor a domain mask")

    # ----------------------------------------------------------- geometry
    @property
    def reach(self) -> int:
        """Nodes one step reads past a point, which a reconstruction strip must cover."""
        return self.space_order // 2

    @property
    def interior(self) -> tuple[slice, ...]:
        return tuple(slice(0, n) for n in self.Nx)

    def flatten_indices(self, position) -> np.ndarray:
        """(ndim, num) grid indices as int32 flat indices of one field (cuwave/wave.py:364-382)."""
        position = np.asarray(as_numpy(position), dtype=np.int64)
        if position.ndim != 2 or position.shape[0] != self.ndim:
            raise ValueError(f"position needs {self.ndim} rows, not {position.shape}")
        lin = np.zeros(position.shape[1], dtype=np.int64)
        for d in range(self.ndim):
            if np.any(position[d] < 0) or np.any(position[d] >= self.Nx[d]):
                raise ValueError(f"a position on axis {d} lies outside 0 .. {self.Nx[d] - 1}")
            lin += position[d] * self.strides[d]
        return lin.astype(np.int32)

    def require_interior(self, position, what: str) -> None:
        """Raise if a node of `position` is a ghost (cuwave/sensitivity.py:103-118)."""
        rows = self.grid_rows(position)
        if rows.size == 0:
            return
        lo = int(rows.min())
        if lo < 1:
            raise ValueError(f"{what} on a ghost node: index {lo} < 1")
        for d in range(self.ndim):
            hi = int(rows[d].max())
            if hi > self.Nx[d] - 2:
                raise ValueError(
                    f"{what} on a ghost node: axis {d} index {hi} exceeds Nx[{d}] - 2 = {self.Nx[d] - 2}"
                )

    def cell_weights(self) -> np.ndarray:
        """The cell volume W a node owns, 1/2 per wall row it sits on (cuwave/wave.py:336-344)."""
        w = np.ones(self.Nx_padded, dtype=self.np_dtype)
        for d in range(self.ndim):
            for index in (1, self.Nx[d] - 2):
                face = [slice(None)] * self.ndim
                face[d] = index
                w[tuple(face)] *= 0.5
        return w

    def sensor_cell_weights(self, sensors) -> np.ndarray:
        """Cell weights at the sensor nodes only (cuwave/wave.py:347-354)."""
        rows = np.asarray(as_numpy(sensors))
        w = np.ones(rows.shape[1], dtype=self.np_dtype)
        for d in range(self.ndim):
            for index in (1, self.Nx[d] - 2):
                w = np.where(rows[d] == index, w * 0.5, w)
        return w

    def abi_grid(self, nbatch: int) -> _abi.Grid:
        """The grid struct of the C ABI for a batch of `nbatch` shots."""
        g = _abi.Grid()
        g.ndim, g.order, g.nbatch = self.ndim, self.space_order, int(nbatch)
        g.lanes = shot_lanes(nbatch)
        g.dtype = 0 if self.precision == "float32" else 1
        for d in range(self.ndim):
            g.n[d], g.np[d] = self.Nx[d], self.Nx_padded[d]
        g.dirichlet = face_mask(self, Dirichlet)
        g.neumann = face_mask(self, Neumann)
        for k, v in enumerate(flux_row(self.space_order, self.stencil)):
            g.op[k] = float(v)
        for d, f in enumerate(self.step_factors()):
            g.f[d] = float(f)
        return g

    # --------------------------------------------------- hooks (subclass)
    def step_factors(self) -> list:
        raise NotImplementedError

    def source_factor(self) -> float:
        raise NotImplementedError

    def adjoint_mass_factor(self) -> float:
        """Scale of the inertia gradient density: `1 / dt**2` (cuwave/wave.py:597-599)."""
        return 1.0 / self.dt**2


def family_drivers(sim: Simulation):
    """The module whose drivers step `sim` when it is of the vector Maxwell or the
    cell-assembled elastic family (each has its own kernels), else None."""
    from .anisotropic import AnisotropicElasticWave
    from .maxwell import MaxwellWave

    if isinstance(sim, MaxwellWave):
        from . import maxwell

        return maxwell
    if isinstance(sim, AnisotropicElasticWave):
        from . import anisotropic

        return anisotropic
    return None


def shot_lanes(nbatch: int) -> int:
    """The innermost extent of a batch's per-shot arrays: a power of two >= nbatch, at most 64."""
    nbatch = int(nbatch)
    if not 1 <= nbatch <= MAX_BATCH:
        raise ValueError(f"a launch batches 1 .. {MAX_BATCH} shots, not {nbatch}")
    return 1 << (nbatch - 1).bit_length()


# ------------------------------------------------------------------- forward
GRAPH_WINDOW = 24  # steps a captured window holds, the reference's chunk (cuwave/wave.py:841)
# node-lanes of a sweep under which it replays captured windows untuned: none.
# On a 298 x 298 one-shot sweep a step's kernel takes about 7 us, longer than
# launching it, and the graphed march measured 9.1 us a step against 7.7 us
# launched plainly; the tuner measures each sweep shape instead.
GRAPH_BELOW = 0


def graph_window(sim: Simulation, lanes: int) -> int:
    """The steps a captured graph window of a forward march holds, 0 for none:
    the tuned choice for the sweep's shape (`cuwave_amd.tuning`) where there is
    one, else GRAPH_WINDOW for sweeps under GRAPH_BELOW node-lanes."""
    from . import tuning

    tuned = tuning.choice(sim, lanes, "graph_window")
    if tuned is not None:
        return int(tuned)
    return GRAPH_WINDOW if int(np.prod(sim.Nx_padded)) * lanes < GRAPH_BELOW else 0


def simulate_shots(sim: Simulation, sources, indicator, sensors=None, record_every=None):
    """Run every shot of `sources` forward in one batch under material `indicator`.

    Returns the (B, *Nx) final interior fields, the (B, N, num_sensors) records when
    `sensors` is given, and the (B, S, *Nx) host snapshots every `record_every`
    steps when that is given.
    """
    from .engine import Propagator, sensor_points, source_points

    own = family_drivers(sim)
    if own is not None:
        if record_every is not None:
            raise NotImplementedError("record_every is built for the pressure family")
        return own.simulate_shots(sim, sources, indicator, sensors)
    if sim.ncomp > 1 or type(sim).__name__ == "ElasticWave":
        from . import elastic

        if record_every is not None:
            raise NotImplementedError("record_every is built for the pressure family")
        return elastic.simulate_shots(sim, sources, indicator, sensors)
    sources = list(sources)
    for s in sources:
        sim.require_interior(s.position, "source")
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, len(sources))
    src, signal = source_points(prop, mat, sources)
    U = prop.levels_memory(2)
    aux = prop.aux_memory()
    rec = record = None
    if sensors is not None:
        sim.require_interior(sensors, "sensor")
        rec = sensor_points(prop, sensors)
        record = prop.record_memory(rec)
    snapshots = []
    if record_every is None:
        from . import tuning

        with tuning.applied(sim, prop.lanes):
            prop.march(U, 0, sim.N, graph_window(sim, prop.lanes), aux=aux, sources=src,
                       signal=signal, sensors=rec, record=record)
    else:
        for t in range(sim.N):
            prop.forward(U, t, t + 1, aux=aux, sources=src, signal=signal, sensors=rec,
                         record=record)
            if t % record_every == 0:
                snapshots.append(prop.shots(U[t % 2])[(slice(None), *sim.interior)].cpu().numpy())
    out = (prop.shots(U[(sim.N - 1) % 2])[(slice(None), *sim.interior)].clone(),)
    if sensors is not None:
        out += (prop.shots(record),)
    if record_every is not None:
        out += (np.stack(snapshots, axis=1),)
    return out if len(out) > 1 else out[0]


def illumination(sim: Simulation, sources, indicator, batch: int | None = None) -> torch.Tensor:
    """The illumination map: per node the sum over steps and shots of u^2.

    Every shot of `sources` runs forward under `indicator`, `batch` at a time
    (at most 64, all by default), and each step adds the squares of its
    final field, summed over the shots, into one per-node map over the
    padded grid. The forward kernels accumulate it as they step. Dividing a
    gradient by it (plus a small floor) compensates for the geometric
    spreading of the source wavefield.
    """
    from .engine import Propagator, source_points

    sources = list(sources)
    if not sources:
        raise ValueError("no shots")
    for s in sources:
        sim.require_interior(s.position, "source")
    step = min(batch or MAX_BATCH, MAX_BATCH)
    if step <= 0:
        raise ValueError(f"batch must be positive, got {batch}")
    mat = sim.build_materials(indicator)
    illum = None
    for i in range(0, len(sources), step):
        prop = Propagator(sim, mat, len(sources[i:i + step]))
        if illum is None:
            illum = prop.accumulator()
        src, signal = source_points(prop, mat, sources[i:i + step])
        U = prop.levels_memory(2)
        from . import tuning

        with tuning.applied(sim, prop.lanes):
            prop.march(U, 0, sim.N, graph_window(sim, prop.lanes), aux=prop.aux_memory(),
                       sources=src, signal=signal, illum=illum)
    return illum


# ------------------------------------------------------------ domain masks
DEEP = 1 << 30  # the cell code of a node whose cells all keep the full radius
HELD = 7  # the radius code of a cell open onto a node held at zero


def stepped_nodes(sim: Simulation) -> np.ndarray:
    """The nodes the step updates: `domain` (every node by default) less `dirichlet`
    (cuwave/wave.py:665-672)."""
    stepped = np.ones(sim.Nx_padded, dtype=bool)
    if sim.domain is not None:
        stepped &= np.asarray(as_numpy(sim.domain), dtype=bool)
    if sim.dirichlet is not None:
        stepped &= ~np.asarray(as_numpy(sim.dirichlet), dtype=bool)
    return stepped


def _runs(inside: np.ndarray, d: int) -> tuple[np.ndarray, np.ndarray]:
    """Per node, how many nodes after it and before it along axis d are `inside`
    without a break: the distance to the nearest node that is not, less one
    (the array's ends count as far away)."""
    n = inside.shape[d]
    shape = [1] * inside.ndim
    shape[d] = n
    at = np.arange(n).reshape(shape)
    far = 2 * n
    # the first outside node at or after each position, then strictly after
    ahead = np.flip(np.minimum.accumulate(np.flip(np.where(inside, far, at), d), axis=d), d)
    ahead = np.concatenate([np.take(ahead, np.arange(1, n), axis=d),
                            np.full_like(np.take(ahead, [0], axis=d), far)], axis=d)
    behind = np.maximum.accumulate(np.where(inside, -far, at), axis=d)
    behind = np.concatenate([np.full_like(np.take(behind, [0], axis=d), -far),
                             np.take(behind, np.arange(0, n - 1), axis=d)], axis=d)
    return ahead - at - 1, at - behind - 1


def domain_cells(sim: Simulation) -> np.ndarray:
    """Per node the int32 code the masked step reads (cuwave/wave.py:675-726).

    A cell's stencil reaches only stepped nodes, so along each axis the cell to the
    next node has radius rp = min(wall, 1 + down, up) and the cell to the previous
    rm = min(wall, down, 1 + up), `up` / `down` the run of stepped nodes after and
    before the node (ghosts count as stepped, so the box walls close nothing) and
    `wall` = min(R, a, Nx - 1 - a), the radius the box leaves at node a. A cell of
    radius 0 leaving the stepped set is closed, unless the node across is held by
    `dirichlet`: then it is open onto the zero halfway across (`HELD`). The code is 0
    for a node not stepped, `DEEP` where every radius is the wall's own, and else
    sum over axes d of (rp | rm << 3) << 6 d.
    """
    sim.check_masks()
    R = sim.space_order // 2
    interior = tuple(slice(1, n - 1) for n in sim.Nx)
    inside = np.ones(sim.Nx_padded, dtype=bool)
    inside[interior] = stepped_nodes(sim)[interior]
    held = np.zeros(sim.Nx_padded, dtype=bool)
    if sim.dirichlet is not None:
        held[interior] = np.asarray(as_numpy(sim.dirichlet), dtype=bool)[interior]
    code = np.zeros(sim.Nx_padded, dtype=np.int64)
    deep = inside.copy()
    for d in range(sim.ndim):
        n = sim.Nx_padded[d]
        shape = [1] * sim.ndim
        shape[d] = n
        a = np.arange(n).reshape(shape)
        wall = np.minimum(R, np.minimum(a, sim.Nx[d] - 1 - a))
        ahead, behind = _runs(inside, d)
        up, down = np.minimum(ahead, wall), np.minimum(behind, wall)
        rp = np.minimum(wall, np.minimum(1 + down, up))
        rm = np.minimum(wall, np.minimum(down, 1 + up))
        # the node across each cell held?
        nxt = np.zeros_like(held)
        prv = np.zeros_like(held)
        lo = [slice(None)] * sim.ndim
        hi = [slice(None)] * sim.ndim
        lo[d], hi[d] = slice(0, n - 1), slice(1, n)
        nxt[tuple(lo)] = held[tuple(hi)]
        prv[tuple(hi)] = held[tuple(lo)]
        rp = np.where((rp == 0) & nxt, HELD, rp)
        rm = np.where((rm == 0) & prv, HELD, rm)
        deep &= (rp == wall) & (rm == wall)
        code |= (np.clip(rp, 0, 7) | np.clip(rm, 0, 7) << 3) << (6 * d)
    code = np.where(deep, DEEP, code)
    return np.where(inside, code, 0).astype(np.int32)


def require_stepped(sim: Simulation, position, what: str) -> None:
    """Raise if a point of `position` sits on a node the masked step never updates
    (cuwave/wave.py:964-967)."""
    if not sim.masked:
        return
    rows = sim.grid_rows(position).astype(np.int64)
    if rows.size and not bool(stepped_nodes(sim)[tuple(rows)].all()):
        raise ValueError(f"{what} reaches a node that is never stepped")


def stable_timestep(sim: Simulation, indicator, iterations: int = 60,
                    safety: float = 0.95) -> float:
    """The largest stable timestep for `sim` under `indicator`, measured on the step itself.

    The leapfrog is stable while the spectral radius of dt^2 minv L stays under 4,
    and L is symmetric with minv diagonal, so a power iteration through the step
    kernels (u0 = 0, u1 = v: 2 v - u2 = -dt^2 minv L v) converges to that radius
    for any order, material and wall layout (cuwave/wave.py:56-116). The PML is
    left out: its memory would make the iteration nonlinear. Returns `safety`
    times the bound; a wide stencil over a strong contrast can put it well below
    `stable_dt`, which knows only the speed and the spacing.
    """
    import dataclasses

    sim = dataclasses.replace(sim, pml=None, damping=None)
    field = np.random.default_rng(0).standard_normal((sim.ncomp, *sim.Nx_padded))
    mask = np.zeros(field.shape, dtype=bool)
    for c in range(sim.ncomp):
        # a staggered component has no unknown on its own axis' last node
        stag = getattr(sim, "radius", None) is not None and sim.ncomp > 1
        inner = tuple(slice(1, n - 1 - (stag and d == c)) for d, n in enumerate(sim.Nx))
        mask[(c, *inner)] = True
    if sim.ncomp > 1 or type(sim).__name__ == "ElasticWave" or family_drivers(sim) is not None:
        from .anisotropic import AnisoEngine, AnisotropicElasticWave
        from .elastic import ElasticEngine
        from .maxwell import MaxwellEngine, MaxwellWave

        mat = sim.build_materials(indicator)
        engine = (MaxwellEngine if isinstance(sim, MaxwellWave) else
                  AnisoEngine if isinstance(sim, AnisotropicElasticWave) else ElasticEngine)
        eng = engine(sim, mat, 1)
        pair = eng.pair()
        slots = eng.levels(3)

        def apply(v):
            slots[0].zero_()
            slots[1].copy_(v)
            pair.step(slots[0], slots[1], slots[2])
            return slots[2]

        shape = (*sim.field_shape, 1)
    else:
        from .engine import Propagator

        mat = sim.build_materials(indicator)
        prop = Propagator(sim, mat, 1)
        slots = prop.levels_memory(3)

        def apply(v):
            # level -2 (slot 0) zero, level -1 (slot 1) v, step 0 into slot 2
            slots[0].zero_()
            slots[1].copy_(v)
            prop.forward(slots, 0, 1, offset=2)
            return slots[2]

        shape = (*sim.Nx_padded, 1)
    keep = torch.as_tensor(mask.reshape(shape), device=slots.device)
    v = torch.as_tensor(field.reshape(shape), dtype=sim.dtype, device=slots.device) * keep
    value = 0.0
    for _ in range(iterations):
        v = v / torch.linalg.vector_norm(v)
        applied = (2.0 * v - apply(v)) * keep
        value = float(torch.sum(v * applied))
        v = applied
    if value <= 0.0:
        raise ValueError(f"the operator came back non-positive: {value}")
    return safety * sim.dt * float(np.sqrt(4.0 / value))


def simulate(sim: Simulation, source: Source, indicator, sensors=None, record_every=None):
    """Run `sim` forward under `source` and material `indicator` (cuwave/wave.py:1071-1163).

    Returns the final interior field, followed by the (N, num_sensors) record when
    `sensors` is given and the stacked host snapshots when `record_every` is.
    """
    out = simulate_shots(sim, [source], indicator, sensors, record_every)
    if not isinstance(out, tuple):
        return out[0]
    return tuple(o[0] for o in out)
