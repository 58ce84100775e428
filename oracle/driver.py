"""CPU oracle of the scalar FD path: host orchestration over `scalar.c`.

TEST INFRASTRUCTURE ONLY: `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` import this module, as the checker and as the
CPU baseline; the product (`cuwave_amd`) never does.

A numpy restatement of the reference's host code for the pressure family, run
on the C kernels of `oracle/scalar.c` (each restating one reference kernel):

- stencil rows: `cuwave/stencils.py:11-58, 186-213`
- materials, factors, weights: `cuwave/scalar.py:407-420, 739-745, 858-950`,
  `cuwave/wave.py:125-134, 336-382`
- the CPML profiles and slabs: `cuwave/boundary.py:327-376`, `cuwave/wave.py:620-631`
- `simulate`: `cuwave/wave.py:1071-1163` (the CPU path: plain steps, no graphs)
- `sensitivity`: `cuwave/sensitivity.py:384-497` (the fused adjoint branch)
- `reconstruction_sensitivity`: `cuwave/sensitivity.py:500-750`, strip nodes
  `:160-195`, treeverse `:229-269`, bfloat16 strips `cuwave/wave.py:844-860`
- `superposition_sensitivity`: `cuwave/sensitivity.py:753-893`
- `misfit_gradient`, `Sensors`, `distribute`: `cuwave/utils.py:19-24, 148-459`
- domain masks: `stepped_nodes`, `domain_cells` `cuwave/wave.py:661-726` (node by
  node), the graded step `cuwave/kernels/scalar.cu:85-158, 434-452`, the point
  check `cuwave/wave.py:964-967`; pinned by the order-2 symmetry of the graded
  operator, polynomial exactness of the graded cells, the method of images for a
  held row, a domain covering everything being the plain step, and no node
  outside the stepped set ever being read (tests/test_domain.py)

Parity pins (tests/test_oracle.py): every entry point is held to the golden
vectors the reference produced (`tests/golden/*.npz`), bit for bit where the
arithmetic is the same, and the adjoint additionally to the transpose identity
and to central finite differences in float64.
"""

from __future__ import annotations

import ctypes
import itertools
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import scipy.ndimage as ndi

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "liboracle.so"
DISPERSION_TOL = 1e-3
CELL_WEIGHT_FACES = (0, 1)


# ----------------------------------------------------------------- stencils
def weights(R: int) -> np.ndarray:
    """Central second-derivative weights w[-R..R] (cuwave/stencils.py:11-17)."""
    x = np.arange(-R, R + 1)
    A = x[None, :] ** np.arange(2 * R + 1)[:, None]
    rhs = np.zeros(2 * R + 1)
    rhs[2] = 2.0
    return np.linalg.solve(A, rhs)


def band_fit(R: int, staggered: bool) -> np.ndarray:
    """The widest-band least-squares rows (cuwave/stencils.py:40-58)."""
    k = np.arange(1, R + 1)
    for band in np.linspace(np.pi, 0.0, 400, endpoint=False):
        theta = np.linspace(band / 400, band, 400)
        G = 2.0 * np.sin(np.outer(theta, k - 0.5)) / theta[:, None]
        if not staggered:
            G *= 2.0 * np.sin(theta / 2.0)[:, None] / theta[:, None]
        consistency = (2 * k - 1.0)[None]
        K = np.block([[G.T @ G, consistency.T], [consistency, np.zeros((1, 1))]])
        c = np.linalg.solve(K, np.append(G.sum(axis=0), 1.0))[:R]
        ratio = G @ c
        error = ratio - 1.0 if staggered else np.sqrt(ratio) - 1.0
        if np.abs(error).max() <= DISPERSION_TOL:
            return c
    raise ValueError("no band fits")


def cell_coefficients(R: int, stencil: str = "taylor") -> np.ndarray:
    """Cell-flux coefficients at radius R (cuwave/stencils.py:20-25)."""
    if stencil != "taylor":
        return band_fit(R, False)
    w = weights(R)
    return np.array([w[R + 1 + k :].sum() for k in range(R)])


def staggered_weights(R: int, stencil: str = "taylor") -> np.ndarray:
    """First-derivative half-offset weights (cuwave/stencils.py:28-37)."""
    if stencil != "taylor":
        return band_fit(R, True)
    x = np.arange(1, R + 1) - 0.5
    A = x[None, :] ** (2 * np.arange(R)[:, None] + 1)
    rhs = np.zeros(R)
    rhs[0] = 0.5
    return np.linalg.solve(A, rhs)


def op_table(order: int, stencil: str) -> np.ndarray:
    """The (4, 4) table of cell rows per radius, the full one `stencil` (:186-200)."""
    R = order // 2
    table = np.zeros((4, 4))
    for r in range(1, R + 1):
        table[r - 1, :r] = cell_coefficients(r, stencil if r == R else "taylor")
    if R == 1:
        table[0, 0] = 1.0  # the radius-1 weight is the literal one (common.cuh:27-29)
    return table


def stable_dt(dx, wavespeed, space_order=2, stencil="optimized") -> float:
    """CFL bound (cuwave/wave.py:44-54)."""
    c = cell_coefficients(space_order // 2, stencil)
    lam = 4.0 * abs(float(c @ (-1.0) ** np.arange(len(c))))
    return 2.0 / (wavespeed * float(np.sqrt(lam * sum(1.0 / d**2 for d in dx))))


# ------------------------------------------------------------------ library
class Grid(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("R", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("np", ctypes.c_int * 3),
        ("s", ctypes.c_long * 3),
        ("op", ctypes.c_double * 16),
    ]


_LIB = None


def library() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not LIBRARY.exists():
            raise RuntimeError(f"{LIBRARY} is not built: run `make -C oracle`")
        _LIB = ctypes.CDLL(str(LIBRARY))
    return _LIB


def _ptr(a):
    if a is None:
        return None
    assert a.flags.c_contiguous, "oracle arrays are C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


def _dbl(values):
    return (ctypes.c_double * 3)(*[float(v) for v in values] + [0.0] * (3 - len(values)))


# --------------------------------------------------------------- simulation
NEUMANN, DIRICHLET = 0, 1


@dataclass
class Pml:
    width: int
    wavespeed: float
    reflection: float = 1e-6
    frequency: float = 0.0
    faces: tuple | None = None


@dataclass
class Sim:
    """The pressure simulation the reference's `ScalarWave` / `AcousticWave` describe."""

    kind: str  # "ScalarWave" or "AcousticWave"
    Nx: tuple
    dx: tuple
    N: int
    dt: float
    precision: str = "float32"
    space_order: int = 2
    stencil: str = "optimized"
    boundary: tuple = None  # ((low, high),) per axis, NEUMANN or DIRICHLET
    pml: Pml | None = None
    wavespeed: float = None
    density: float = None
    materials: tuple = None  # rho1, rho2, kappa1, kappa2
    damping: np.ndarray | None = None
    pml_margin: int = 0  # undamped layer nodes before the front (elastic: radius - 1)
    domain: np.ndarray | None = None  # nodal mask of the stepped domain, None for all
    dirichlet: np.ndarray | None = None  # nodal mask held at zero, its cells left open
    ndim: int = field(init=False)

    def __post_init__(self):
        self.Nx = tuple(int(n) for n in self.Nx)
        self.dx = tuple(float(d) for d in self.dx)
        self.ndim = len(self.Nx)
        self.Nx_padded = (*self.Nx[:-1], -(-self.Nx[-1] // 32) * 32)
        strides = [1] * self.ndim
        for d in range(self.ndim - 2, -1, -1):
            strides[d] = strides[d + 1] * self.Nx_padded[d + 1]
        self.strides = tuple(strides)
        self.dtype = np.float32 if self.precision == "float32" else np.float64
        if self.boundary is None:
            self.boundary = ((NEUMANN, NEUMANN),) * self.ndim
        self.boundary = tuple(tuple(int(c) for c in pair) for pair in self.boundary)
        self.R = self.space_order // 2
        self.derive_inertia = self.kind == "ScalarWave"

    @property
    def reach(self) -> int:
        return self.R

    @property
    def masked(self) -> bool:
        return self.domain is not None or self.dirichlet is not None

    def face_mask(self, condition) -> int:
        return sum(1 << (2 * d + s) for d, pair in enumerate(self.boundary)
                   for s, c in enumerate(pair) if c == condition)

    def grid(self) -> Grid:
        g = Grid()
        g.ndim, g.R = self.ndim, self.R
        for d in range(self.ndim):
            g.n[d], g.np[d], g.s[d] = self.Nx[d], self.Nx_padded[d], self.strides[d]
        table = op_table(self.space_order, self.stencil)
        for i in range(16):
            g.op[i] = table.flat[i]
        return g

    # hooks (cuwave/scalar.py:875-950); "BulkDensityWave" has no reference class:
    # the reference's PressureWave hooks (cuwave/scalar.py:140-160, 739-745) fed a
    # (kappa, rho) indicator stack, stiff = 1 / rho and minv = kappa
    def parametrization(self, indicator):
        if self.kind == "ScalarWave":
            return indicator, None
        if self.kind == "BulkDensityWave":
            kappa, rho = indicator[0], indicator[1]
            return 1 / rho, kappa
        rho1, rho2, kappa1, kappa2 = self.materials
        rho_inv = 1 / rho1 + indicator * (1 / rho2 - 1 / rho1)
        kappa_inv = 1 / kappa1 + indicator * (1 / kappa2 - 1 / kappa1)
        return rho_inv, 1 / kappa_inv

    def parametrization_jacobian(self, indicator=None):
        if self.kind == "ScalarWave":
            return 1.0, 1.0
        if self.kind == "BulkDensityWave":
            kappa, rho = np.asarray(indicator, dtype=self.dtype)
            return -1 / kappa**2, -1 / rho**2
        rho1, rho2, kappa1, kappa2 = self.materials
        return 1 / kappa2 - 1 / kappa1, 1 / rho2 - 1 / rho1

    def step_factors(self):
        if self.kind == "ScalarWave":
            return [self.dtype(2.0 * self.wavespeed**2 * self.dt**2 / d**2) for d in self.dx]
        return [self.dtype(2.0 * self.dt**2 / d**2) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0 / self.density if self.kind == "ScalarWave" else 1.0

    def mirror_ghosts(self, f):
        for d in range(self.ndim):
            for ghost, mirror in ((0, 2), (self.Nx[d] - 1, self.Nx[d] - 3)):
                dst = [slice(None)] * self.ndim
                dst[d] = ghost
                src = [slice(None)] * self.ndim
                src[d] = mirror
                f[tuple(dst)] = f[tuple(src)]
        return f

    def build_materials(self, indicator) -> dict:
        stiff, minv = self.parametrization(np.asarray(indicator, dtype=self.dtype))
        mat = {"stiff": self.mirror_ghosts(np.array(stiff, dtype=self.dtype))}
        mat["minv"] = None if minv is None else self.mirror_ghosts(
            np.array(minv, dtype=self.dtype))
        return mat

    def inverse_inertia(self, indicator) -> np.ndarray:
        """Nodal 1 / m from the parametrization (cuwave/scalar.py:417-420)."""
        stiff, minv = self.parametrization(np.asarray(indicator, dtype=self.dtype))
        return 1.0 / stiff if minv is None else minv

    def kernel_minv(self, mat):
        return mat["stiff"] if self.derive_inertia else mat["minv"]

    def flatten(self, position) -> np.ndarray:
        lin = np.zeros(position.shape[1], dtype=np.int64)
        for d in range(self.ndim):
            lin += np.asarray(position[d], dtype=np.int64) * self.strides[d]
        return lin.astype(np.int32)

    def excitation_weights(self, mat, lin, damping="sim") -> np.ndarray:
        """dt^2 source_factor / m at `lin`, over 1 + beta where the step is damped
        (cuwave/scalar.py:857-872); `damping` overrides the simulation's field."""
        damping = self.damping if isinstance(damping, str) else damping
        fld = mat["stiff"] if self.derive_inertia else mat["minv"]
        w = fld.ravel()[lin]
        if self.derive_inertia:
            w = 1.0 / w
        if damping is not None:
            beta = 0.5 * w * damping.ravel()[lin] * self.dt
            w = w / (1.0 + beta)
        return (self.dtype(self.dt**2 * self.source_factor()) * w).astype(self.dtype)

    def sensor_cell_weights(self, sensors) -> np.ndarray:
        w = np.ones(sensors.shape[1], dtype=self.dtype)
        for d in range(self.ndim):
            for index in (1, self.Nx[d] - 2):
                w = np.where(sensors[d] == index, w * 0.5, w)
        return w

    def apply_cell_weights(self, f):
        for d in range(self.ndim):
            for index in (1, self.Nx[d] - 2):
                face = [slice(None)] * self.ndim
                face[d] = index
                f[tuple(face)] *= 0.5
        return f

    # the pml (cuwave/boundary.py:318-376)
    def pml_widths(self):
        faces = range(2 * self.ndim) if self.pml.faces is None else self.pml.faces
        return [tuple(self.pml.width if 2 * d + s in faces else 0 for s in (0, 1))
                for d in range(self.ndim)]

    def pml_profiles(self):
        layer = self.pml
        L = layer.width - 1 - self.pml_margin  # cuwave/boundary.py:335-336
        out = []
        for d, (lo, hi) in enumerate(self.pml_widths()):
            if not lo + hi:
                out.append(None)
                continue
            n = self.Nx[d]
            damping = -3.0 * layer.wavespeed * np.log(layer.reflection) / (2.0 * L * self.dx[d])
            rows = []
            for position in (np.arange(n) + shift for shift in (0.0, 0.5)):
                distance = np.full(n, np.inf)
                if lo:
                    distance = np.minimum(distance, np.abs(position - 1))
                if hi:
                    distance = np.minimum(distance, np.abs(n - 2 - position))
                depth = np.clip((L - distance) / L, 0.0, 1.0)
                d0 = damping * depth**2
                alpha = np.pi * layer.frequency
                inside = depth > 0.0
                b = np.where(inside, np.exp(-(d0 + alpha) * self.dt), 1.0)
                a = np.where(inside, d0 / np.maximum(d0 + alpha, 1e-300) * (b - 1.0), 0.0)
                rows += [a, b]
            out.append(np.ascontiguousarray(np.stack(rows), dtype=self.dtype))
        return out

    def pml_slabs(self):
        if self.pml is None:
            return []
        total = int(np.prod(self.Nx_padded))
        return [(d, lo, hi, prof, total // self.Nx_padded[d])
                for d, ((lo, hi), prof) in enumerate(zip(self.pml_widths(), self.pml_profiles()))
                if prof is not None]

    def pml_nodes(self) -> np.ndarray:
        inside = np.zeros(self.Nx_padded, dtype=bool)
        if self.pml is None:
            return inside
        for d, (lo, hi) in enumerate(self.pml_widths()):
            index = [slice(None)] * self.ndim
            for rows in (slice(1, 1 + lo), slice(self.Nx[d] - 1 - hi, self.Nx[d] - 1)):
                index[d] = rows
                inside[tuple(index)] = True
        interior = np.zeros_like(inside)
        interior[tuple(slice(1, n - 1) for n in self.Nx)] = True
        return inside & interior

    def interior_view(self, f):
        return f[(Ellipsis, *(slice(0, n) for n in self.Nx))]


def sim_from_npz(z) -> Sim:
    """The simulation a golden case stores."""
    pml = None
    if "pml" in z:
        w, c, r, f = z["pml"]
        pml = Pml(int(w), float(c), float(r), float(f), tuple(int(x) for x in z["pml_faces"]))
    kw = {}
    if "wavespeed" in z:
        kw.update(wavespeed=float(z["wavespeed"]), density=float(z["density"]))
    else:
        kw.update(materials=tuple(float(v) for v in z["materials"]))
    return Sim(str(z["kind"]), tuple(z["Nx"]), tuple(z["dx"]), int(z["N"]),
               float(z["dt"]), str(z["precision"]), int(z["space_order"]),
               str(z["stencil"]), tuple(map(tuple, z["boundary"])), pml, **kw)


# ------------------------------------------------------------- domain masks
DEEP_CELLS, HELD_CELL = 1 << 30, 7  # cuwave/wave.py:661-662


def stepped_nodes(sim: Sim) -> np.ndarray:
    """`domain` (every node by default) less `dirichlet` (cuwave/wave.py:665-672)."""
    stepped = np.ones(sim.Nx_padded, dtype=bool)
    if sim.domain is not None:
        stepped &= np.asarray(sim.domain, dtype=bool)
    if sim.dirichlet is not None:
        stepped &= ~np.asarray(sim.dirichlet, dtype=bool)
    return stepped


def domain_cells(sim: Sim) -> np.ndarray:
    """The masked step's per-node cell codes (cuwave/wave.py:675-726), node by node.

    At every stepped interior node and along each axis: walk outwards while the
    nodes are inside (stepped, or a ghost or padding node, or past the array) and
    within the box's radius `wall` = min(R, a, Nx - 1 - a), giving the runs `up` and
    `down`; the cell to the next node takes radius min(wall, 1 + down, up), the cell
    to the previous min(wall, down, 1 + up), and a radius 0 cell whose node across is
    held becomes `HELD_CELL`. Every radius at `wall` is `DEEP_CELLS`, else the radii
    packed (rp | rm << 3) << 6 d. Only interior nodes are coded (the rest 0).
    """
    R = sim.R
    stepped = stepped_nodes(sim)
    held = (np.zeros(sim.Nx_padded, dtype=bool) if sim.dirichlet is None
            else np.asarray(sim.dirichlet, dtype=bool))

    def interior(b):
        return all(1 <= b[d] <= sim.Nx[d] - 2 for d in range(sim.ndim))

    def inside(b):
        if any(not 0 <= b[d] < sim.Nx_padded[d] for d in range(sim.ndim)):
            return True
        return bool(stepped[b]) if interior(b) else True

    def held_at(b):
        return interior(b) and bool(held[b])

    code = np.zeros(sim.Nx_padded, dtype=np.int64)
    for a in itertools.product(*(range(1, n - 1) for n in sim.Nx)):
        if not stepped[a]:
            continue
        packed, deep = 0, True
        for d in range(sim.ndim):
            wall = min(R, a[d], sim.Nx[d] - 1 - a[d])

            def at(k, d=d):
                b = list(a)
                b[d] += k
                return tuple(b)

            up = 0
            while up < wall and inside(at(up + 1)):
                up += 1
            down = 0
            while down < wall and inside(at(-down - 1)):
                down += 1
            rp, rm = min(wall, 1 + down, up), min(wall, down, 1 + up)
            if rp == 0 and held_at(at(1)):
                rp = HELD_CELL
            if rm == 0 and held_at(at(-1)):
                rm = HELD_CELL
            deep &= rp == wall and rm == wall
            packed |= (rp | rm << 3) << (6 * d)
        code[a] = DEEP_CELLS if deep else packed
    return code.astype(np.int32)


def require_full_domain(sim: Sim) -> None:
    """cuwave/sensitivity.py:121-126: the material gradients refuse a mask."""
    if sim.masked:
        raise NotImplementedError("the material gradients step every node: unset domain")


def require_stepped(sim: Sim, position, what: str) -> None:
    """cuwave/wave.py:964-967: a point on a node the masked step never updates raises."""
    if sim.masked and not stepped_nodes(sim)[tuple(np.asarray(position))].all():
        raise ValueError(f"{what} reaches a node that is never stepped")


# -------------------------------------------------------------------- steps
class Steps:
    """The kernels of one simulation bound to its materials, as the reference's closures."""

    def __init__(self, sim: Sim, mat: dict, with_pml: bool = True, damping="sim"):
        self.sim, self.mat = sim, mat
        # the damping field the step reads: the simulation's, or another (the
        # reverse march's, negated where it runs backwards)
        self.damping = sim.damping if isinstance(damping, str) else damping
        self.lib = library()
        self.suf = "_f32" if sim.dtype == np.float32 else "_f64"
        self.g = sim.grid()
        self.f = _dbl(sim.step_factors())
        self.dirichlet = sim.face_mask(DIRICHLET)
        self.neumann = sim.face_mask(NEUMANN)
        self.slabs = sim.pml_slabs() if with_pml else []
        self.memory = [np.zeros((3, (lo + hi) * plane), dtype=sim.dtype)
                       for _, lo, hi, _, plane in self.slabs]
        self.minv = sim.kernel_minv(mat)
        self.cells = domain_cells(sim) if sim.masked else None

    def fn(self, name):
        return getattr(self.lib, f"cwo_{name}{self.suf}")

    def fd(self, u0, u1, u2):
        s = self.sim
        self.fn("fd_step")(ctypes.byref(self.g), _ptr(u0), _ptr(u1), _ptr(u2),
                           _ptr(self.mat["stiff"]), _ptr(self.minv),
                           ctypes.c_int(s.derive_inertia), _ptr(self.damping),
                           ctypes.c_double(s.dt), self.f, ctypes.c_int(self.dirichlet),
                           _ptr(self.cells))
        for (d, lo, hi, prof, _), aux in zip(self.slabs, self.memory):
            self.fn("pml_step")(ctypes.byref(self.g), _ptr(u1), _ptr(u2),
                                _ptr(self.mat["stiff"]), _ptr(self.minv),
                                ctypes.c_int(s.derive_inertia), _ptr(aux), _ptr(prof),
                                d, lo, hi, ctypes.c_double(float(self.f[d])),
                                ctypes.c_int(self.dirichlet))
        return u2

    def refill(self, u):
        self.fn("refill")(ctypes.byref(self.g), _ptr(u), self.neumann, self.dirichlet)

    def excite(self, u, row, lin, weight):
        self.fn("excite")(_ptr(u), _ptr(np.ascontiguousarray(row)), _ptr(lin),
                          len(lin), _ptr(weight))

    def record(self, u, row, lin):
        self.fn("record")(_ptr(u), _ptr(row), _ptr(lin), len(lin))

    def replay(self, u, row, lin):
        self.fn("replay")(_ptr(u), _ptr(np.ascontiguousarray(row)), _ptr(lin), len(lin))


# ----------------------------------------------------------------- simulate
def simulate(sim: Sim, src_pos, src_sig, indicator, sensors=None):
    """Final interior field and (N, num_sensors) record (cuwave/wave.py:1071-1163)."""
    require_stepped(sim, src_pos, "a source")
    if sensors is not None:
        require_stepped(sim, sensors, "a sensor")
    mat = sim.build_materials(indicator)
    st = Steps(sim, mat)
    src_lin = sim.flatten(src_pos)
    weight = sim.excitation_weights(mat, src_lin)
    signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
    U = np.zeros((2, *sim.Nx_padded), dtype=sim.dtype)
    um = None
    if sensors is not None:
        rec_lin = sim.flatten(sensors)
        um = np.zeros((sim.N, sensors.shape[1]), dtype=sim.dtype)
    for t in range(sim.N):
        u = st.fd(U[t % 2], U[1 - t % 2], U[t % 2])
        st.excite(u, signal[t], src_lin, weight)
        st.refill(u)
        if um is not None:
            st.record(u, um[t], rec_lin)
    field_ = sim.interior_view(U[(sim.N - 1) % 2]).copy()
    return (field_, um) if um is not None else field_


def illumination(sim: Sim, shots, indicator):
    """Per node the sum over the shots and steps of u^2, u each step's final
    field as `simulate` marches it (no reference counterpart: the closed form
    sum_s sum_t u_s(t)^2), accumulated in float64."""
    mat = sim.build_materials(indicator)
    st = Steps(sim, mat)
    acc = np.zeros(sim.Nx_padded)
    for src_pos, src_sig in shots:
        st.memory = [np.zeros_like(m) for m in st.memory]
        src_lin = sim.flatten(src_pos)
        weight = sim.excitation_weights(mat, src_lin)
        signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
        U = np.zeros((2, *sim.Nx_padded), dtype=sim.dtype)
        for t in range(sim.N):
            u = st.fd(U[t % 2], U[1 - t % 2], U[t % 2])
            st.excite(u, signal[t], src_lin, weight)
            st.refill(u)
            acc += u.astype(np.float64) ** 2
    inner = np.zeros(sim.Nx_padded, dtype=bool)
    inner[tuple(slice(1, n - 1) for n in sim.Nx)] = True
    return np.where(inner, acc, 0.0)


def l2_misfit(observed):
    def objective(traces):
        residual = traces - observed
        return 0.5 * float(np.sum(residual**2)), residual
    return objective


def adjoint_signal(sim: Sim, dphi, sensors, scale=1.0, delay=0):
    """(cuwave/sensitivity.py:272-292)"""
    w = sim.sensor_cell_weights(sensors) * sim.source_factor()
    signal = np.asarray(dphi, dtype=sim.dtype)[::-1] / w
    if delay:
        signal = np.concatenate((signal[delay:], np.zeros_like(signal[:delay])))
    return np.ascontiguousarray(sim.dtype(scale) * signal, dtype=sim.dtype)


def _finalize(sim: Sim, grads: dict) -> dict:
    for f in grads.values():
        sim.apply_cell_weights(f)
    return grads


def _frozen(sim: Sim, grads: dict) -> dict:
    if sim.pml is not None:
        layer = sim.pml_nodes()
        for f in grads.values():
            f[layer] = 0.0
    return grads


class AdjointPml:
    """The transposed pml launches of one simulation (cuwave/scalar.py:369-405)."""

    def __init__(self, st: Steps, grads: dict | None):
        self.st, self.grads = st, grads
        self.memory = [np.zeros((3, (lo + hi) * plane), dtype=st.sim.dtype)
                       for _, lo, hi, _, plane in st.slabs]

    def __call__(self, l1, l2, u1=None):
        st, sim = self.st, self.st.sim
        fused = self.grads is not None
        for (d, lo, hi, prof, _), eta in zip(st.slabs, self.memory):
            st.fn("pml_aux")(ctypes.byref(st.g), _ptr(l1), _ptr(st.mat["stiff"]), _ptr(eta),
                             _ptr(prof), d, lo, hi, ctypes.c_double(float(st.f[d])),
                             ctypes.c_int(st.dirichlet))
            gm = self.grads["mass"] if fused else l2
            st.fn("pml_adjoint")(ctypes.byref(st.g), _ptr(l2), _ptr(gm),
                                 _ptr(l2 if u1 is None else u1),
                                 ctypes.c_double(float(sim.dtype(1.0 / sim.dt**2))),
                                 ctypes.c_int(int(fused)), _ptr(st.mat["stiff"]),
                                 _ptr(st.minv), ctypes.c_int(sim.derive_inertia), _ptr(eta),
                                 _ptr(prof), d, lo, hi, ctypes.c_double(float(st.f[d])),
                                 ctypes.c_int(st.dirichlet))
        return l2


def sensitivity(sim: Sim, src_pos, src_sig, indicator, sensors, objective):
    """Stored-history adjoint, the fused branch (cuwave/sensitivity.py:384-497)."""
    require_full_domain(sim)
    mat = sim.build_materials(indicator)
    st = Steps(sim, mat)
    src_lin = sim.flatten(src_pos)
    weight = sim.excitation_weights(mat, src_lin)
    signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
    grads = {k: np.zeros(sim.Nx_padded, dtype=sim.dtype) for k in ("mass", "stiff")}
    V = np.zeros((sim.N + 2, *sim.Nx_padded), dtype=sim.dtype)
    for t in range(sim.N):
        u = st.fd(V[t], V[t + 1], V[t + 2])
        st.excite(u, signal[t], src_lin, weight)
        st.refill(u)
    rec_lin = sim.flatten(sensors)
    um = np.ascontiguousarray(V[2:].reshape(sim.N, -1)[:, rec_lin])
    cost, dphi = objective(um)
    adj = adjoint_signal(sim, dphi, sensors)
    P = np.zeros((2, *sim.Nx_padded), dtype=sim.dtype)
    pml = AdjointPml(st, grads) if st.slabs else None
    adj_weight = sim.excitation_weights(mat, rec_lin)
    mf = float(sim.dtype(1.0 / sim.dt**2))
    for m in range(sim.N):
        u1 = V[sim.N + 1 - m]
        l0, l1 = P[m % 2], P[1 - m % 2]
        st.fn("adjoint_gradient")(ctypes.byref(st.g), _ptr(l0), _ptr(l1),
                                  _ptr(grads["mass"]), _ptr(grads["stiff"]), _ptr(u1),
                                  _ptr(mat["stiff"]), _ptr(st.minv),
                                  ctypes.c_int(sim.derive_inertia), _ptr(sim.damping),
                                  ctypes.c_double(sim.dt), ctypes.c_int(st.dirichlet),
                                  ctypes.c_double(mf), st.f)
        if pml is not None:
            pml(l1, l0, u1)
        st.fn("adjoint_excite")(_ptr(l0), _ptr(np.ascontiguousarray(adj[m])), _ptr(rec_lin),
                                len(rec_lin), _ptr(adj_weight), _ptr(grads["mass"]),
                                _ptr(u1), ctypes.c_double(mf))
    return cost, _frozen(sim, _finalize(sim, grads)), um, {}


# ------------------------------------------------------------ reconstruction
def narrow(values, to):
    """Stored as `to`, bfloat16 as round-to-nearest-even bits (cuwave/wave.py:848-853)."""
    if np.dtype(to) != np.uint16:
        return values.astype(to)
    bits = values.astype(np.float32).view(np.uint32)
    return ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16).astype(np.uint16)


def widen(values, dtype):
    if values.dtype != np.uint16:
        return values.astype(dtype)
    return (values.astype(np.uint32) << 16).view(np.float32).astype(dtype)


REVERSIBLE_GROWTH = {np.float32: 1e2, np.float64: 1e6}


def reversible_damping(sim: Sim, indicator, span: int) -> np.ndarray:
    """Damped nodes a reverse march of `span` steps can run backwards
    (cuwave/sensitivity.py:139-157): the damped update solved for its trailing
    slot is the update with the damping negated, whose growth (1 + beta) / (1 - beta)
    a step must stay under REVERSIBLE_GROWTH over the span; beta taken at the
    largest inverse inertia of the node's 3^ndim box."""
    if sim.damping is None:
        return np.zeros(sim.Nx_padded, dtype=bool)
    minv = ndi.maximum_filter(sim.inverse_inertia(indicator), size=3)
    beta = 0.5 * sim.dt * sim.damping * minv
    tiny = np.finfo(sim.dtype).tiny
    with np.errstate(over="ignore"):  # beta >= 1 overflows to inf: irreversible
        growth = span * np.log((1.0 + beta) / np.maximum(1.0 - beta, tiny))
    bound = math.log(REVERSIBLE_GROWTH[sim.dtype])
    return (sim.damping > 0) & (beta < 1.0) & (growth <= bound)


def reconstruction_nodes(sim: Sim, indicator=None, span=None):
    """The strip and the reversible mask (cuwave/sensitivity.py:160-195): layer
    nodes are irreversible, damped ones where `reversible_damping` says so."""
    interior = np.zeros(sim.Nx_padded, dtype=bool)
    interior[tuple(slice(1, n - 1) for n in sim.Nx)] = True
    if sim.pml is None and sim.damping is None:
        return np.zeros((sim.ndim, 0), dtype=np.int32), interior
    reversible = interior.copy()
    if sim.pml is not None:
        reversible &= ~sim.pml_nodes()
    if sim.damping is not None:
        span = sim.N if span is None else span
        reversible &= ~(sim.damping > 0) | reversible_damping(sim, indicator, span)
    square = ndi.generate_binary_structure(sim.ndim, sim.ndim)
    reach = ndi.binary_dilation(reversible, structure=square, iterations=sim.reach,
                                brute_force=True)
    strip = np.stack(np.nonzero(interior & ~reversible & reach)).astype(np.int32)
    return strip, reversible


def segment_split(lo, hi, free):
    r = 0
    while math.comb(free + r, free) < hi - lo:
        r += 1
    return max(lo + 1, hi - math.comb(free - 1 + r, free - 1))


def checkpoint_schedule(segments, slots):
    """Treeverse actions (cuwave/sensitivity.py:243-269)."""
    actions = []

    def reverse(lo, hi, free):
        while hi - lo > 1:
            if free == 0:
                for k in range(hi - 1, lo, -1):
                    actions.extend((("restore", lo), ("advance", lo, k), ("leaf", k)))
                break
            mid = segment_split(lo, hi, free)
            actions.extend((("restore", lo), ("advance", lo, mid), ("store", mid)))
            reverse(mid, hi, free - 1)
            actions.append(("drop", mid))
            hi = mid
        actions.extend((("restore", lo), ("leaf", lo)))

    reverse(0, segments, slots)
    return actions


def reconstruction_sensitivity(sim: Sim, src_pos, src_sig, indicator, sensors, objective,
                               strip_dtype=None, checkpoints=None):
    """Strip-replayed reverse march (cuwave/sensitivity.py:500-750), serial order."""
    require_full_domain(sim)
    stored = sim.dtype
    if strip_dtype == "bfloat16":
        stored = np.uint16
    elif strip_dtype == "float32" and sim.dtype == np.float64:
        stored = np.float32
    mat = sim.build_materials(indicator)
    st = Steps(sim, mat)
    src_lin = sim.flatten(src_pos)
    weight = sim.excitation_weights(mat, src_lin)
    signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
    rec_lin = sim.flatten(sensors)
    adj_weight = sim.excitation_weights(mat, rec_lin)
    grads = {k: np.zeros(sim.Nx_padded, dtype=sim.dtype) for k in ("mass", "stiff")}
    adj_pml = AdjointPml(st, None) if st.slabs else None
    U = np.zeros((3, *sim.Nx_padded), dtype=sim.dtype)
    snapshot_bytes = 2 * U[0].nbytes + sum(a.nbytes for a in st.memory)
    strip, valid = reconstruction_nodes(sim, indicator)
    segment = sim.N
    if checkpoints is not None and strip.shape[1]:
        row_bytes = 2 * strip.shape[1] * np.dtype(stored).itemsize
        segment = min(sim.N, max(2, snapshot_bytes // row_bytes))
    if segment < sim.N:
        strip, valid = reconstruction_nodes(sim, indicator, segment)
    if not valid.any():
        raise ValueError("damping too strong to reverse covers every interior node")
    # backwards the layer sits behind the strip and a reversible damping is negated
    back_damping = None
    if sim.damping is not None:
        back_damping = np.where(valid, -sim.damping, sim.damping).astype(sim.dtype)
    back = Steps(sim, mat, with_pml=False, damping=back_damping)
    back_weight = sim.excitation_weights(mat, src_lin, damping=back_damping)
    num_strip = strip.shape[1]
    lin_strip = sim.flatten(strip)
    starts = list(range(0, sim.N, segment)) + [sim.N]
    um = np.zeros((sim.N, sensors.shape[1]), dtype=sim.dtype)
    P = np.zeros((2, *sim.Nx_padded), dtype=sim.dtype)
    snapshots = {}
    state = {"recorded": 0, "steps": 0}
    Fg = _dbl([float(f) / sim.dt**2 for f in sim.step_factors()])
    inv_dt2 = float(sim.dtype(1.0 / sim.dt**2))

    def forward(t, row_sig, um_row=None, strip_row=None):
        u = st.fd(U[t % 3], U[(t + 1) % 3], U[(t + 2) % 3])
        st.excite(u, row_sig, src_lin, weight)
        st.refill(u)
        if um_row is not None:
            st.record(u, um_row, rec_lin)
        if strip_row is not None:
            st.record(u, strip_row, lin_strip)

    def advance(a, b, buffer=None):
        first = a == state["recorded"]
        if first:
            state["recorded"] = b
        for k in range(b - a):
            t = a + k
            srow = None
            if buffer is not None:
                srow = np.zeros(num_strip, dtype=sim.dtype)
            forward(t, signal[t], um[t] if first else None, srow)
            if buffer is not None:
                buffer[k] = narrow(srow, stored)
        state["steps"] += b - a

    def slots(t):
        return U[t % 3], U[(t + 1) % 3]

    def store(k):
        snapshots[k] = [f.copy() for f in (*slots(starts[k]), *st.memory)]

    def restore(k):
        for f, saved in zip((*slots(starts[k]), *st.memory), snapshots.get(k, ())):
            f[...] = saved
        if k == 0:
            for f in (*slots(0), *st.memory):
                f[...] = 0.0

    def triplet(m):
        return tuple(U[(sim.N + i - m) % 3] for i in (2, 0, 1))

    def adjoint(m, row):
        l0, l1 = P[m % 2], P[1 - m % 2]
        st.fn("adjoint_step")(ctypes.byref(st.g), _ptr(l0), _ptr(l1), _ptr(l0),
                              _ptr(mat["stiff"]), _ptr(st.minv),
                              ctypes.c_int(sim.derive_inertia), _ptr(sim.damping),
                              ctypes.c_double(sim.dt), st.f, ctypes.c_int(st.dirichlet))
        if adj_pml is not None:
            adj_pml(l1, l0)
        st.excite(l0, state["signal"][row], rec_lin, adj_weight)
        st.refill(l0)
        return l0

    def gradient_step(u0, u1, u2, l1):
        st.fn("gradient")(ctypes.byref(st.g), _ptr(grads["mass"]), _ptr(grads["stiff"]),
                          _ptr(u0), _ptr(u1), _ptr(u2), _ptr(l1), _ptr(mat["stiff"]),
                          ctypes.c_double(inv_dt2), Fg, ctypes.c_int(st.dirichlet))

    def reverse(m, src_row, strip_row=None):
        a, b, c = triplet(m)
        c = back.fd(b, a, c)
        back.excite(c, src_row, src_lin, back_weight)
        if strip_row is not None:
            back.replay(c, strip_row, lin_strip)
        back.refill(c)

    def leaf(k):
        s, e = starts[k], starts[k + 1]
        buffer = np.zeros((e - s + 2, num_strip), dtype=stored)
        if num_strip and s:
            for row, f in enumerate(slots(s)):
                buffer[row] = narrow(f.ravel()[lin_strip], stored)
        advance(s, e, buffer[2:] if num_strip else None)
        if "signal" not in state:
            state["cost"], dphi = objective(um)
            state["signal"] = adjoint_signal(sim, dphi, sensors)
        m0, m1 = sim.N - e, sim.N - s
        if not s:
            state["seed"] = float(np.linalg.norm(triplet(m0)[2] * valid))
        gradient_step(*triplet(m0), adjoint(m0, m0))
        src_rows = signal[s : e - 1][::-1]
        strip_rows = widen(buffer[: e - s - 1][::-1], sim.dtype) if num_strip else None
        for j in range(e - s - 1):
            m = m0 + j
            reverse(m, src_rows[j], None if strip_rows is None else strip_rows[j])
            p = adjoint(m + 1, m0 + 1 + j)
            gradient_step(*triplet(m + 1), p)

    for action, *segs in checkpoint_schedule(len(starts) - 1, checkpoints or 0):
        if action == "advance":
            advance(starts[segs[0]], starts[segs[1]])
        elif action == "store":
            store(segs[0])
        elif action == "restore":
            restore(segs[0])
        elif action == "drop":
            del snapshots[segs[0]]
        else:
            leaf(segs[0])

    grads = _finalize(sim, grads)
    for f in grads.values():
        f[~valid] = 0.0
    seed = state["seed"]
    a = triplet(sim.N - 1)[0]
    drift = float(np.linalg.norm(a * valid)) / seed if seed > 0.0 else float("inf")
    info = {"strip": num_strip, "drift": drift, "segment": segment,
            "sweeps": state["steps"] / sim.N, "valid": valid}
    return state["cost"], grads, um, info


# ------------------------------------------------------------- superposition
def superposition_sensitivity(sim: Sim, src_pos, src_sig, indicator, sensors, objective,
                              scale=1.0):
    """Lossless three-slot variant (cuwave/sensitivity.py:753-893), fused kernels."""
    require_full_domain(sim)
    if sim.pml is not None or sim.damping is not None:
        raise NotImplementedError("superposition needs a lossless operator")
    mat = sim.build_materials(indicator)
    st = Steps(sim, mat)
    src_lin = sim.flatten(src_pos)
    weight = sim.excitation_weights(mat, src_lin)
    signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
    rec_lin = sim.flatten(sensors)
    accs = {k: np.zeros(sim.Nx_padded, dtype=sim.dtype) for k in ("mass", "stiff")}
    factors = sim.step_factors()

    def superposed(sign):
        ratios = [-sign / (2.0 * d) ** 2 / float(f) for d, f in zip(sim.dx, factors)]
        ft = float(sim.dtype(sign / (2.0 * sim.dt) ** 2))
        fs = float(sim.dtype(ratios[0]))

        def step(u0, u1, u2, mass=True):
            st.fn("superposed")(ctypes.byref(st.g), _ptr(u0), _ptr(u1), _ptr(u2),
                                _ptr(accs["mass"]), _ptr(accs["stiff"]), _ptr(mat["stiff"]),
                                _ptr(st.minv), ctypes.c_int(sim.derive_inertia),
                                ctypes.c_double(ft if mass else 0.0), ctypes.c_double(fs),
                                st.f, ctypes.c_int(st.dirichlet))
        return step

    def frechet(sign, u0, u1, u2):
        fr = _dbl([float(sim.dtype(-sign * 0 / (2.0 * d) ** 2)) for d in sim.dx])
        st.fn("frechet")(ctypes.byref(st.g), _ptr(accs["mass"]), _ptr(accs["stiff"]),
                         _ptr(u0), _ptr(u1), _ptr(u2),
                         ctypes.c_double(float(sim.dtype(sign / (2.0 * sim.dt) ** 2))), fr)

    U = np.zeros((3, *sim.Nx_padded), dtype=sim.dtype)
    um = np.zeros((sim.N, sensors.shape[1]), dtype=sim.dtype)
    sub, add = superposed(-1.0), superposed(1.0)
    for t in range(sim.N):
        u0, u1, u2 = U[t % 3], U[(t + 1) % 3], U[(t + 2) % 3]
        sub(u0, u1, u2)
        st.excite(u2, signal[t], src_lin, weight)
        st.refill(u2)
        st.record(u2, um[t], rec_lin)
    t = sim.N - 1
    frechet(-1.0, U[t % 3], U[(t + 1) % 3], U[(t + 2) % 3])
    cost, dphi = objective(um)
    before = math.sqrt(sum(float(np.linalg.norm(f)) ** 2 for f in accs.values()))
    back_pos = np.concatenate((sensors, src_pos), axis=1)
    back_lin = sim.flatten(back_pos)
    back_weight = sim.excitation_weights(mat, back_lin)
    back_sig = np.ascontiguousarray(np.concatenate(
        (adjoint_signal(sim, dphi, sensors, scale, 1), signal[: sim.N][::-1]), axis=1))

    def backward(t, row, mass=True):
        u0, u1, u2 = (U[(sim.N + i - t) % 3] for i in (1, 0, 2))
        add(u0, u1, u2, mass)
        st.excite(u2, back_sig[row], back_lin, back_weight)
        st.refill(u2)

    backward(0, 0, mass=False)
    for t in range(sim.N - 1):
        backward(t + 1, t + 1)
    last = [U[(1 + i) % 3] for i in (1, 0, 2)]
    frechet(1.0, *last)
    after = math.sqrt(sum(float(np.linalg.norm(f)) ** 2 for f in accs.values()))
    cancellation = before / after if after > 0.0 else float("inf")
    norm = sim.dtype(1.0 / (2.0 * scale))
    accs = _finalize(sim, accs)
    for f in accs.values():
        f *= norm
    return cost, accs, um, {"scale": scale, "cancellation": cancellation}


# -------------------------------------------------------------------- misfit
def corner_weights(sim: Sim, coords):
    """Multilinear corners and weights (cuwave/utils.py:148-182), nodal."""
    coords = np.atleast_2d(np.asarray(coords, dtype=float))
    index = np.empty_like(coords)
    for d in range(sim.ndim):
        nodal = coords[:, d] / sim.dx[d] + 1.0
        index[:, d] = np.clip(nodal, 1.0, sim.Nx[d] - 2.0)
    base = np.stack([np.clip(np.floor(index[:, d]), 1, sim.Nx[d] - 3).astype(np.int32)
                     for d in range(sim.ndim)])
    offset = index.T - base
    corners = np.array(list(itertools.product((0, 1), repeat=sim.ndim)))
    nodes = base[:, :, None] + corners.T[:, None, :]
    w = np.ones((len(coords), len(corners)))
    for d in range(sim.ndim):
        o = offset[d][:, None]
        w *= np.where(corners[None, :, d] == 1, o, 1.0 - o)
    return nodes.reshape(sim.ndim, -1).astype(np.int32), w.astype(sim.dtype)


def point_source(sim: Sim, coords, signal):
    nodes, w = corner_weights(sim, coords)
    signal = np.asarray(signal, dtype=sim.dtype)
    if signal.ndim == 1:
        signal = np.broadcast_to(signal[:, None], (signal.shape[0], w.shape[0]))
    cols = (signal[:, :, None] * w).reshape(signal.shape[0], -1)
    return nodes, np.ascontiguousarray(cols / np.prod(sim.dx), dtype=sim.dtype)


class Sensors:
    def __init__(self, sim: Sim, coords):
        self.coords = np.atleast_2d(np.asarray(coords, dtype=float))
        self.count = len(self.coords)
        self.nodes, self.weights = corner_weights(sim, self.coords)

    def traces(self, record):
        return (record.reshape(record.shape[0], self.count, -1) * self.weights).sum(2)

    def scatter(self, dphi):
        return (dphi[:, :, None] * self.weights).reshape(dphi.shape[0], -1)

    def objective(self, objective):
        def wrapped(record):
            cost, dphi = objective(self.traces(record))
            return cost, self.scatter(dphi)
        return wrapped


def misfit_gradient(sim: Sim, sources, indicator, sensors: Sensors, observed,
                    adjoint=sensitivity):
    """Shot loop, chain rule onto the indicator (cuwave/utils.py:410-459)."""
    cost = 0.0
    gradient = np.zeros(sim.Nx_padded, dtype=sim.dtype)
    d_mass, d_stiff = sim.parametrization_jacobian()
    for (pos, sig), data in zip(sources, observed):
        c, grads, _, _ = adjoint(sim, pos, sig, indicator, sensors.nodes,
                                 sensors.objective(l2_misfit(data)))
        cost += c
        gradient += d_mass * grads["mass"] + d_stiff * grads["stiff"]
    return cost, gradient
