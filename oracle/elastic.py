"""CPU oracle of the elastic FD path: host orchestration over `elastic.c`.

TEST INFRASTRUCTURE ONLY: `tests/` and the `cpu_baseline` leg of `bench.py`
import this module, as the checker and as the CPU baseline; the product
(`cuwave_amd`) never does.

A numpy restatement of the reference's host code for `ElasticWave`, run on
the C kernels of `oracle/elastic.c`:

- the staggered closure tables and norms are the reference's own
  (`cuwave/stencils.py:75-213`), read from `tests/golden/el_stencils.npz`,
  which the reference wrote;
- material constants, factors, point inertia, excitation and adjoint weights:
  `cuwave/elastic.py:154-345, 627-704`;
- the gradient chain through the point and pair averages:
  `cuwave/elastic.py:705-729`, `cuwave/wave.py:386-475`;
- `simulate`: `cuwave/wave.py:1071-1163` (the pair: stress then update);
- `sensitivity`: `cuwave/sensitivity.py:384-497`, the fused adjoint branch
  (`cuwave/elastic.py:436-488`);
- `reconstruction_sensitivity`: `cuwave/sensitivity.py:500-750`, the strip
  with one row per component (`:160-226`), the gradient kernel
  (`cuwave/elastic.py:731-762`);
- `misfit_gradient` over shots: `cuwave/utils.py:410-459`.

Parity pins (tests/test_elastic_oracle.py): the golden vectors bit for bit,
the symmetry of the lossless operator, the 1D reduction to the scalar solver
and central finite differences in float64.
"""

from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import scipy.ndimage as ndi

from . import driver as D

HERE = Path(__file__).resolve().parent
TABLES = HERE.parent / "tests" / "golden" / "el_stencils.npz"
TRACTION, CLAMPED = 0, 1


class ElGeom(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("R", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("np", ctypes.c_int * 3),
        ("s", ctypes.c_long * 3),
        ("cs", ctypes.c_long),
        ("sg", ctypes.c_double * 4),
        ("nc", ctypes.c_int),
        ("hc", ctypes.c_int),
        ("nspan", ctypes.c_int),
        ("hspan", ctypes.c_int),
        ("node_rows", (ctypes.c_double * 16) * 12),
        ("half_rows", (ctypes.c_double * 16) * 12),
        ("lam", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("rho", ctypes.c_double),
        ("clamped", ctypes.c_int),
    ]


class ElPml(ctypes.Structure):
    _fields_ = [("aux", ctypes.c_void_p), ("profile", ctypes.c_void_p),
                ("lo", ctypes.c_int), ("hi", ctypes.c_int)]


def tables(order: int, stencil: str) -> dict:
    """The reference's closure rows, counts, full-radius weights and norms."""
    z = np.load(TABLES)
    key = f"o{order}_{stencil}"
    counts = z[f"{key}_counts"]
    b = {2: 1, 4: 4, 6: 6, 8: 8}[order]
    return {"node_rows": z[f"{key}_node_rows"], "half_rows": z[f"{key}_half_rows"],
            "counts": tuple(int(c) for c in counts), "stag": z[f"{key}_stag"],
            "node_norm": z[f"o{order}_node_norm"][:b], "half_norm": z[f"o{order}_half_norm"][:b]}


@dataclass
class ElSim:
    """The simulation `cuwave/elastic.py:154-345` describes (lossless, no damping)."""

    Nx: tuple
    dx: tuple
    N: int
    dt: float
    precision: str = "float64"
    space_order: int = 2
    stencil: str = "taylor"
    boundary: tuple = None  # ((low, high),) per axis, TRACTION or CLAMPED
    density: float = None
    wavespeed_p: float = None
    wavespeed_s: float = None
    plane: str = "strain"
    pml: D.Pml | None = None

    def __post_init__(self):
        self.Nx = tuple(int(n) for n in self.Nx)
        self.dx = tuple(float(d) for d in self.dx)
        self.ndim = len(self.Nx)
        self.Nx_padded = (*self.Nx[:-1], -(-self.Nx[-1] // 32) * 32)
        strides = [1] * self.ndim
        for d in range(self.ndim - 2, -1, -1):
            strides[d] = strides[d + 1] * self.Nx_padded[d + 1]
        self.strides = tuple(strides)
        self.comp_stride = int(np.prod(self.Nx_padded))
        self.dtype = np.float32 if self.precision == "float32" else np.float64
        if self.boundary is None:
            self.boundary = ((TRACTION, TRACTION),) * self.ndim
        self.boundary = tuple(tuple(int(c) for c in pair) for pair in self.boundary)
        self.R = self.space_order // 2
        self.ncomp = self.ndim
        self.nvoigt = self.ndim + self.ndim * (self.ndim - 1) // 2
        self.field_shape = (self.ncomp, *self.Nx_padded)
        self.node_rows = self.ndim + (self.ncomp > 1)
        self.tab = tables(self.space_order, self.stencil)

    @property
    def reach(self) -> int:
        return 2 * self.R - 1

    @property
    def shear(self) -> float:
        return self.density * self.wavespeed_s**2

    @property
    def lame(self) -> float:
        lam = self.density * (self.wavespeed_p**2 - 2.0 * self.wavespeed_s**2)
        if self.plane == "stress":
            return 2.0 * lam * self.shear / (lam + 2.0 * self.shear)
        return lam

    def clamped_mask(self) -> int:
        return sum(1 << (2 * d + s) for d, pair in enumerate(self.boundary)
                   for s, c in enumerate(pair) if c == CLAMPED)

    def geom(self) -> ElGeom:
        g = ElGeom()
        g.ndim, g.R = self.ndim, self.R
        for d in range(self.ndim):
            g.n[d], g.np[d], g.s[d] = self.Nx[d], self.Nx_padded[d], self.strides[d]
        g.cs = self.comp_stride
        for k, v in enumerate(self.tab["stag"]):
            g.sg[k] = float(v)
        g.nc, g.hc, g.nspan, g.hspan = self.tab["counts"]
        for table, name in ((g.node_rows, "node_rows"), (g.half_rows, "half_rows")):
            for i, row in enumerate(self.tab[name]):
                for j, v in enumerate(row):
                    table[i][j] = float(v)
        g.lam = float(self.dtype(self.lame))
        g.mu = float(self.dtype(self.shear))
        g.rho = float(self.dtype(self.density))
        g.clamped = self.clamped_mask()
        return g

    def step_factors(self):
        return [self.dtype(self.dt**2 / d) for d in self.dx]

    def inv_dx(self):
        return [self.dtype(1.0 / d) for d in self.dx]

    # -------------------------------------------------------------- norms
    def _norm_lines(self, half):
        """Per axis the closure norms along the padded extent (cuwave/elastic.py:290-303)."""
        b = len(self.tab["node_norm"])
        lines = []
        for d, n in enumerate(self.Nx):
            intervals = n - 3
            node, halfp = np.ones(intervals + 1), np.ones(intervals)
            node[:b], node[intervals + 1 - b:] = self.tab["node_norm"], self.tab["node_norm"][::-1]
            halfp[:b], halfp[intervals - b:] = self.tab["half_norm"], self.tab["half_norm"][::-1]
            line = np.ones(self.Nx_padded[d])
            if d in half:
                line[1:n - 2] = halfp
            else:
                line[1:n - 1] = node
            lines.append(line)
        return lines

    def point_weights(self, half):
        w = np.ones(self.Nx_padded, dtype=self.dtype)
        for d, line in enumerate(self._norm_lines(half)):
            shape = [1] * self.ndim
            shape[d] = -1
            w *= np.asarray(line.reshape(shape), dtype=self.dtype)
        return w

    def _point_norms(self, lin):
        component, node = np.divmod(np.asarray(lin, dtype=np.int64), self.comp_stride)
        nodal = self._norm_lines(())
        half = self._norm_lines(tuple(range(self.ndim)))
        weight = np.ones(len(node))
        for d in range(self.ndim):
            a = node // self.strides[d] % self.Nx_padded[d]
            weight *= np.where(component == d, half[d][a], nodal[d][a])
        return np.asarray(weight, dtype=self.dtype)

    # -------------------------------------------------------------- points
    def flatten(self, position) -> np.ndarray:
        position = np.asarray(position, dtype=np.int64)
        rows = position[-self.ndim:]
        lin = np.zeros(position.shape[1], dtype=np.int64)
        for d in range(self.ndim):
            lin += rows[d] * self.strides[d]
        if self.ncomp > 1:
            lin += position[0] * self.comp_stride
        return lin.astype(np.int32)

    def point_inverse_inertia(self, gamma, lin):
        """cuwave/elastic.py:313-332."""
        lin = np.asarray(lin, dtype=np.int64)
        component, node = lin // self.comp_stride, lin % self.comp_stride
        rows = [node // s % n for s, n in zip(self.strides, self.Nx_padded)]
        step = np.asarray(self.strides, dtype=np.int64)[component]
        flat = gamma.ravel()
        mean = 0.5 * (flat[node] + flat[node + step])
        mass = self.density * self._point_norms(lin) * mean
        minv = 1.0 / np.maximum(mass, np.finfo(self.dtype).tiny)
        for face in [2 * d + s for d, pair in enumerate(self.boundary)
                     for s, c in enumerate(pair) if c == CLAMPED]:
            d = face // 2
            wall = 1 if face % 2 == 0 else self.Nx[d] - 2
            minv = np.where((rows[d] == wall) & (component != d), 0.0, minv)
        return minv.astype(self.dtype)

    def excitation_weights(self, gamma, lin):
        volume = self.dtype(1.0 / float(np.prod(self.dx)))
        weight = self.point_inverse_inertia(gamma, lin) * volume
        return (self.dtype(self.dt**2 * 1.0) * weight).astype(self.dtype)

    def adjoint_weights(self, sensors):
        volume = self.dtype(1.0 / float(np.prod(self.dx)))
        return np.full(sensors.shape[1], volume, dtype=self.dtype)

    # ------------------------------------------------------------------ pml
    def pml_sim(self) -> D.Sim:
        """A pressure simulation of the same grid and layer: the profiles and slabs."""
        return D.Sim("ScalarWave", self.Nx, self.dx, self.N, self.dt, self.precision,
                     self.space_order, pml=self.pml, wavespeed=1.0, density=1.0,
                     pml_margin=self.R - 1)

    def pml_nodes(self):
        return self.pml_sim().pml_nodes()


def sim_from_npz(z) -> ElSim:
    pml = None
    if "pml" in z:
        w, c, r, f = z["pml"]
        pml = D.Pml(int(w), float(c), float(r), float(f), tuple(int(x) for x in z["pml_faces"]))
    rho, cp_, cs_ = (float(v) for v in z["elastic"])
    return ElSim(tuple(z["Nx"]), tuple(z["dx"]), int(z["N"]), float(z["dt"]),
                 str(z["precision"]), int(z["space_order"]), str(z["stencil"]),
                 tuple(map(tuple, z["boundary"])), rho, cp_, cs_, str(z["plane"]), pml)


# ---------------------------------------------------------------- the pair
class Pair:
    """The stress and update launches of one simulation, a pml folded into both."""

    def __init__(self, sim: ElSim, gamma, with_pml=True, adjoint=False):
        self.sim = sim
        self.lib = D.library()
        self.suf = "_f32" if sim.dtype == np.float32 else "_f64"
        self.g = sim.geom()
        self.gamma = np.ascontiguousarray(gamma, dtype=sim.dtype)
        self.sigma = np.zeros((sim.nvoigt, *sim.Nx_padded), dtype=sim.dtype)
        self.fs = (ctypes.c_double * 3)(*[float(v) for v in sim.inv_dx()] + [0.0] * (3 - sim.ndim))
        self.fu = (ctypes.c_double * 3)(*[float(v) for v in sim.step_factors()]
                                        + [0.0] * (3 - sim.ndim))
        self.pml = None
        self.memory = []
        self.phase = 1 if adjoint else 0
        if with_pml and sim.pml is not None:
            psim = sim.pml_sim()
            slabs = {d: rest for d, *rest in psim.pml_slabs()}
            self.pml = (ElPml * 3)()
            self.profiles = []
            rows = (3 if adjoint else 2) * sim.ndim
            for d in range(sim.ndim):
                if d not in slabs:
                    continue
                lo, hi, prof, plane = slabs[d]
                aux = np.zeros((rows, (lo + hi) * plane), dtype=sim.dtype)
                prof = np.ascontiguousarray(prof, dtype=sim.dtype)
                self.memory.append(aux)
                self.profiles.append(prof)
                self.pml[d].aux = aux.ctypes.data
                self.pml[d].profile = prof.ctypes.data
                self.pml[d].lo, self.pml[d].hi = lo, hi

    def fn(self, name):
        return getattr(self.lib, f"{name}{self.suf}")

    def stress(self, u1, gradient=None, v1=None):
        gn = gs = None
        if gradient is not None:
            gn, gs = gradient["normal"], gradient.get("shear")
        self.fn("el_stress")(ctypes.byref(self.g), D._ptr(u1), D._ptr(self.sigma),
                             D._ptr(self.gamma), self.fs, self.pml, ctypes.c_int(self.phase),
                             ctypes.c_int(gradient is not None), D._ptr(gn), D._ptr(gs),
                             D._ptr(v1))

    def update(self, u0, u1, u2, gradient=None, v1=None, mf=0.0):
        gm = None if gradient is None else gradient["mass"]
        self.fn("el_update")(ctypes.byref(self.g), D._ptr(u0), D._ptr(u1), D._ptr(u2),
                             D._ptr(self.sigma), D._ptr(self.gamma), self.fu, self.pml,
                             ctypes.c_int(self.phase), ctypes.c_int(gradient is not None),
                             D._ptr(gm), D._ptr(v1), ctypes.c_double(mf))

    def step(self, u0, u1, u2, gradient=None, v1=None, mf=0.0):
        self.stress(u1, gradient, v1)
        self.update(u0, u1, u2, gradient, v1, mf)
        if self.phase:  # the adjoint's strain memory swaps its two row blocks
            self.phase = 3 - self.phase
        return u2

    def gradient(self, grads, u0, u1, u2, l1, mf):
        self.fn("el_gradient")(ctypes.byref(self.g), D._ptr(grads["mass"]),
                               D._ptr(grads["normal"]), D._ptr(grads.get("shear")),
                               D._ptr(u0), D._ptr(u1), D._ptr(u2), D._ptr(l1),
                               ctypes.c_double(mf), self.fs)

    def excite(self, u, row, lin, weight):
        self.lib.__getattr__(f"cwo_excite{self.suf}")(
            D._ptr(u), D._ptr(np.ascontiguousarray(row)), D._ptr(lin), len(lin), D._ptr(weight))

    def record(self, u, row, lin):
        self.lib.__getattr__(f"cwo_record{self.suf}")(D._ptr(u), D._ptr(row), D._ptr(lin),
                                                      len(lin))

    def replay(self, u, row, lin):
        self.lib.__getattr__(f"cwo_replay{self.suf}")(
            D._ptr(u), D._ptr(np.ascontiguousarray(row)), D._ptr(lin), len(lin))

    def adjoint_excite(self, l2, row, lin, weight, g_mass, u1, mf):
        self.fn("el_adjoint_excite")(D._ptr(l2), D._ptr(np.ascontiguousarray(row)),
                                     D._ptr(lin), len(lin), D._ptr(weight), D._ptr(g_mass),
                                     D._ptr(u1), ctypes.c_double(mf))


# ---------------------------------------------------------------- gradients
def gradient_fields(sim: ElSim) -> dict:
    g = {"mass": np.zeros((sim.ncomp, *sim.Nx_padded), dtype=sim.dtype),
         "normal": np.zeros(sim.Nx_padded, dtype=sim.dtype)}
    if sim.ndim > 1:
        g["shear"] = np.zeros((sim.nvoigt - sim.ndim, *sim.Nx_padded), dtype=sim.dtype)
    return g


PAIRS = {1: ((0, 0),), 2: ((0, 0), (1, 1), (0, 1)),
         3: ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))}


def point_average_adjoint(sim, density, c):
    out = 0.5 * density
    to = [slice(None)] * sim.ndim
    fro = [slice(None)] * sim.ndim
    to[c], fro[c] = slice(1, None), slice(0, -1)
    out[tuple(to)] += 0.5 * density[tuple(fro)]
    return out


def pair_average(sim, field, axes):
    out = np.zeros(sim.Nx_padded, dtype=sim.dtype)
    safe = np.maximum(field, np.finfo(sim.dtype).tiny)
    inner = tuple(slice(0, n - 1) if d in axes else slice(0, n) for d, n in enumerate(sim.Nx))
    for bits in itertools.product((0, 1), repeat=2):
        shifted = tuple(slice(bits[axes.index(d)], n - 1 + bits[axes.index(d)])
                        if d in axes else slice(0, n) for d, n in enumerate(sim.Nx))
        out[inner] += 1.0 / safe[shifted]
    out[inner] = 4.0 / out[inner]
    return out


def pair_average_adjoint(sim, density, field, axes):
    safe = np.maximum(field, np.finfo(sim.dtype).tiny)
    mean = pair_average(sim, field, axes)
    scattered = density * mean * mean * 0.25
    out = np.zeros(sim.Nx_padded, dtype=sim.dtype)
    for bits in itertools.product((0, 1), repeat=2):
        to = [slice(None)] * sim.ndim
        fro = [slice(None)] * sim.ndim
        for d, b in zip(axes, bits):
            to[d], fro[d] = slice(b, None), slice(0, -b if b else None)
        out[tuple(to)] += scattered[tuple(fro)] / safe[tuple(to)] ** 2
    return out


def finalize_gradients(sim: ElSim, grads: dict, gamma) -> dict:
    """cuwave/elastic.py:716-729."""
    g_mass = np.zeros(sim.Nx_padded, dtype=sim.dtype)
    for c in range(sim.ncomp):
        density = grads["mass"][c] * sim.point_weights((c,)) * sim.density
        g_mass += point_average_adjoint(sim, density, c)
    g_stiff = grads["normal"] * sim.point_weights(())
    for p, axes in enumerate(PAIRS[sim.ndim][sim.ndim:]):
        density = grads["shear"][p] * sim.point_weights(axes)
        g_stiff += pair_average_adjoint(sim, density, gamma, axes)
    return {"mass": g_mass, "stiff": g_stiff}


def frozen(sim: ElSim, grads: dict) -> dict:
    if sim.pml is not None:
        layer = sim.pml_nodes()
        for f in grads.values():
            f[..., layer] = 0.0
    return grads


def adjoint_signal(sim: ElSim, dphi, sensors):
    signal = np.asarray(dphi, dtype=sim.dtype)[::-1] / sim.adjoint_weights(sensors)
    return np.ascontiguousarray(signal, dtype=sim.dtype)


# ------------------------------------------------------------------- drivers
def simulate(sim: ElSim, src_pos, src_sig, gamma, sensors=None):
    """Final (ncomp, *Nx) field and (N, num_sensors) record (cuwave/wave.py:1071-1163)."""
    gamma = np.ascontiguousarray(gamma, dtype=sim.dtype)
    pair = Pair(sim, gamma)
    lin = sim.flatten(src_pos)
    weight = sim.excitation_weights(gamma, lin)
    signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
    U = np.zeros((2, *sim.field_shape), dtype=sim.dtype)
    um = None
    if sensors is not None:
        rec = sim.flatten(sensors)
        um = np.zeros((sim.N, sensors.shape[1]), dtype=sim.dtype)
    for t in range(sim.N):
        # U[t % 2] takes u^t over u^(t - 2)
        u = pair.step(U[t % 2], U[1 - t % 2], U[t % 2])
        pair.excite(u, signal[t], lin, weight)
        if um is not None:
            pair.record(u, um[t], rec)
    field = U[(sim.N - 1) % 2][(Ellipsis, *(slice(0, n) for n in sim.Nx))].copy()
    if sim.ncomp == 1:
        field = field[0]
    return field, um


def sensitivity(sim: ElSim, src_pos, src_sig, gamma, sensors, objective):
    """Stored history, the fused adjoint-gradient pair (cuwave/sensitivity.py:384-497)."""
    gamma = np.ascontiguousarray(gamma, dtype=sim.dtype)
    pair = Pair(sim, gamma)
    lin = sim.flatten(src_pos)
    weight = sim.excitation_weights(gamma, lin)
    signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
    V = np.zeros((sim.N + 2, *sim.field_shape), dtype=sim.dtype)
    for t in range(sim.N):
        u = pair.step(V[t], V[t + 1], V[t + 2])
        pair.excite(u, signal[t], lin, weight)
    rec = sim.flatten(sensors)
    um = np.ascontiguousarray(V[2:].reshape(sim.N, -1)[:, rec])
    cost, dphi = objective(um)
    adj = adjoint_signal(sim, dphi, sensors)
    grads = gradient_fields(sim)
    back = Pair(sim, gamma, adjoint=sim.pml is not None)
    adj_weight = sim.excitation_weights(gamma, rec)
    mf = float(sim.dtype(1.0 / sim.dt**2))
    P = np.zeros((2, *sim.field_shape), dtype=sim.dtype)
    for m in range(sim.N):
        v1 = V[sim.N + 1 - m]
        l0, l1 = P[m % 2], P[1 - m % 2]
        back.step(l0, l1, l0, grads, v1, mf)
        back.adjoint_excite(l0, adj[m], rec, adj_weight, grads["mass"], v1, mf)
    out = frozen(sim, finalize_gradients(sim, grads, gamma))
    return cost, out, um, {}


def with_components(sim, nodes):
    if sim.ncomp == 1:
        return nodes
    tiled = np.tile(nodes, (1, sim.ncomp))
    row = np.repeat(np.arange(sim.ncomp, dtype=np.int32), nodes.shape[1])
    return np.ascontiguousarray(np.concatenate((row[None, :], tiled), axis=0))


def reconstruction_nodes(sim: ElSim):
    """cuwave/sensitivity.py:160-195 with the component rows of :208-215."""
    interior = np.zeros(sim.Nx_padded, dtype=bool)
    interior[tuple(slice(1, n - 1) for n in sim.Nx)] = True
    if sim.pml is None:
        return np.zeros((sim.node_rows, 0), dtype=np.int32), interior
    reversible = interior & ~sim.pml_nodes()
    square = ndi.generate_binary_structure(sim.ndim, sim.ndim)
    reach = ndi.binary_dilation(reversible, structure=square, iterations=sim.reach,
                                brute_force=True)
    strip = np.stack(np.nonzero(interior & ~reversible & reach)).astype(np.int32)
    return with_components(sim, strip), reversible


def reconstruction_sensitivity(sim: ElSim, src_pos, src_sig, gamma, sensors, objective,
                               checkpoints=None):
    """Strip-replayed reverse march (cuwave/sensitivity.py:500-750), serial order."""
    gamma = np.ascontiguousarray(gamma, dtype=sim.dtype)
    fwd = Pair(sim, gamma)
    back = Pair(sim, gamma, with_pml=False)
    adjp = Pair(sim, gamma, adjoint=sim.pml is not None)
    grads = gradient_fields(sim)
    lin = sim.flatten(src_pos)
    weight = sim.excitation_weights(gamma, lin)
    signal = np.ascontiguousarray(src_sig, dtype=sim.dtype)
    rec = sim.flatten(sensors)
    adj_weight = sim.excitation_weights(gamma, rec)
    U = np.zeros((3, *sim.field_shape), dtype=sim.dtype)
    snapshot_bytes = 2 * U[0].nbytes + sum(a.nbytes for a in fwd.memory)
    strip, valid = reconstruction_nodes(sim)
    segment = sim.N
    if checkpoints is not None and strip.shape[1]:
        row_bytes = 2 * strip.shape[1] * np.dtype(sim.dtype).itemsize
        segment = min(sim.N, max(2, snapshot_bytes // row_bytes))
    num_strip = strip.shape[1]
    lin_strip = sim.flatten(strip)
    starts = list(range(0, sim.N, segment)) + [sim.N]
    um = np.zeros((sim.N, sensors.shape[1]), dtype=sim.dtype)
    P = np.zeros((2, *sim.field_shape), dtype=sim.dtype)
    snapshots = {}
    state = {"recorded": 0, "steps": 0}
    mf = float(sim.dtype(1.0 / sim.dt**2))

    def forward(t, row_sig, um_row=None, strip_row=None):
        u = fwd.step(U[t % 3], U[(t + 1) % 3], U[(t + 2) % 3])
        fwd.excite(u, row_sig, lin, weight)
        if um_row is not None:
            fwd.record(u, um_row, rec)
        if strip_row is not None:
            fwd.record(u, strip_row, lin_strip)

    def advance(a, b, buffer=None):
        first = a == state["recorded"]
        if first:
            state["recorded"] = b
        for k in range(b - a):
            t = a + k
            forward(t, signal[t], um[t] if first else None,
                    buffer[k] if buffer is not None else None)
        state["steps"] += b - a

    def slots(t):
        return U[t % 3], U[(t + 1) % 3]

    def store(k):
        snapshots[k] = [f.copy() for f in (*slots(starts[k]), *fwd.memory)]

    def restore(k):
        for f, saved in zip((*slots(starts[k]), *fwd.memory), snapshots.get(k, ())):
            f[...] = saved
        if k == 0:
            for f in (*slots(0), *fwd.memory):
                f[...] = 0.0

    def triplet(m):
        return tuple(U[(sim.N + i - m) % 3] for i in (2, 0, 1))

    def adjoint(m, row):
        l0, l1 = P[m % 2], P[1 - m % 2]
        adjp.step(l0, l1, l0)
        adjp.excite(l0, state["signal"][row], rec, adj_weight)
        return l0

    def reverse(m, src_row, strip_row=None):
        a, b, c = triplet(m)
        c = back.step(b, a, c)
        back.excite(c, src_row, lin, weight)
        if strip_row is not None:
            back.replay(c, strip_row, lin_strip)

    def leaf(k):
        s, e = starts[k], starts[k + 1]
        buffer = np.zeros((e - s + 2, num_strip), dtype=sim.dtype)
        if num_strip and s:
            for row, f in enumerate(slots(s)):
                buffer[row] = f.ravel()[lin_strip]
        advance(s, e, buffer[2:] if num_strip else None)
        if "signal" not in state:
            state["cost"], dphi = objective(um)
            state["signal"] = adjoint_signal(sim, dphi, sensors)
        m0, m1 = sim.N - e, sim.N - s
        if not s:
            state["seed"] = float(np.linalg.norm(triplet(m0)[2] * valid))
        fwd.gradient(grads, *triplet(m0), adjoint(m0, m0), mf)
        src_rows = signal[s: e - 1][::-1]
        strip_rows = buffer[: e - s - 1][::-1] if num_strip else None
        for j in range(e - s - 1):
            m = m0 + j
            reverse(m, src_rows[j], None if strip_rows is None else strip_rows[j])
            p = adjoint(m + 1, m0 + 1 + j)
            fwd.gradient(grads, *triplet(m + 1), p, mf)

    for action, *segs in D.checkpoint_schedule(len(starts) - 1, checkpoints or 0):
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

    out = finalize_gradients(sim, grads, gamma)
    for f in out.values():
        f[~valid] = 0.0
    seed = state["seed"]
    a = triplet(sim.N - 1)[0]
    drift = float(np.linalg.norm(a * valid)) / seed if seed > 0.0 else float("inf")
    info = {"strip": num_strip, "drift": drift, "segment": segment,
            "sweeps": state["steps"] / sim.N, "valid": valid}
    return state["cost"], out, um, info


def misfit_gradient(sim: ElSim, shots, gamma, sensors, observed):
    """cuwave/utils.py:410-459 with the reconstruction adjoint; gamma's Jacobian is (1, 1)."""
    cost = 0.0
    gradient = np.zeros(sim.Nx_padded, dtype=sim.dtype)
    for (pos, sig), data in zip(shots, observed):
        c, g, _, _ = reconstruction_sensitivity(sim, pos, sig, gamma, sensors,
                                                D.l2_misfit(data))
        cost += c
        gradient += 1.0 * g["mass"] + 1.0 * g["stiff"]
    return cost, gradient
