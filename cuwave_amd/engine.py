"""Device state of one pressure simulation over a batch of shots, and its launches.

A `Propagator` binds a simulation and its material to the library: it prepares
the inverse inertia and the harmonic cell stiffness once (`cw_scalar_prepare`),
uploads the PML profiles, buckets source and sensor points by the kernels' tiles,
and launches the fused forward and adjoint steps on torch's current stream. It
plays the part of the reference's closure factories (`define_step`,
`define_excitation`, `define_get_signal`, `define_adjoint_gradient`,
`define_adjoint_pml`, cuwave/scalar.py:279-811, cuwave/wave.py:956-1033), minus
the per-step host work: a call advances a whole range of steps.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from .boundary import pml_profiles, pml_widths
from .stencils import cell_rows


@dataclass
class PointList:
    """Points bucketed by tile on the device, and the struct that names them."""

    points: _abi.Points
    keep: tuple  # the tensors the struct points into
    count: int  # columns of the signal / record table


class Propagator:
    """The kernels of one simulation bound to its material, for `nbatch` shots at once.

    Set `timings` to a list to have every forward and adjoint call bracketed by
    HIP events on the launch stream: entries (kind, launches, start, end).
    """

    timings: list | None = None

    def __init__(self, sim, mat: dict, nbatch: int):
        self.device = _abi.require_device()
        self.sim, self.nbatch = sim, int(nbatch)
        self.grid = sim.abi_grid(self.nbatch)
        self.lanes = self.grid.lanes
        dtype = sim.dtype
        self.stiff = mat["stiff"]
        self.minv = mat.get("minv")
        # the prepared materials: the harmonic cell stiffness of every axis,
        # then the inverse inertia, one block the kernels read through one
        # buffer resource
        self.prep = torch.empty((sim.ndim + 1, *sim.Nx_padded), dtype=dtype, device=self.device)
        self.damping = mat.get("damping")
        self.damping_reverse = None
        self.cells = mat.get("cells")  # a domain mask's packed cell radii, or None
        self.mat = _abi.Materials(_abi.ptr(self.stiff), _abi.ptr(self.minv), _abi.ptr(self.prep),
                                  _abi.ptr(self.damping), None, float(sim.dt),
                                  _abi.ptr(self.cells))
        for r, row in enumerate(cell_rows(sim.space_order, sim.stencil)):
            for k, w in enumerate(row):
                self.mat.rows[r][k] = float(w)
        _abi.call("cw_scalar_prepare", ctypes.byref(self.grid), ctypes.byref(self.mat), _abi.stream())
        tiles = _abi.I32()
        _abi.call("cw_tile_count", ctypes.byref(self.grid), ctypes.byref(tiles))
        self.ntiles = tiles.value
        self.field = int(np.prod(sim.Nx_padded))
        # the layer: widths, profiles and slab sizes per lined axis
        self.pml = _abi.Pml()
        self.slabs = []  # (axis, slab nodes)
        self.profiles = []
        if sim.pml is not None:
            for d, ((lo, hi), prof) in enumerate(zip(pml_widths(sim), pml_profiles(sim))):
                if prof is None:
                    continue
                profile = torch.as_tensor(prof, device=self.device)
                self.profiles.append(profile)
                self.pml.lo[d], self.pml.hi[d] = lo, hi
                self.pml.profile[d] = _abi.ptr(profile)
                plane = self.field // sim.Nx_padded[d]
                self.slabs.append((d, (lo + hi) * plane))

    def reverse_damping(self, field: torch.Tensor | None) -> None:
        """The damping the reverse march of a reconstruction steps with (the field
        negated where it runs backwards), or None for the forward field."""
        if field is not None:
            self._require_shape("reverse damping", field, tuple(self.sim.Nx_padded))
        self.damping_reverse = field
        self.mat.damping_reverse = _abi.ptr(field)

    # ------------------------------------------------------------ memory
    # Per-shot arrays end in the lanes axis: one node's shots side by side.
    def _zeros(self, *shape) -> torch.Tensor:
        return torch.zeros((*shape, self.lanes), dtype=self.sim.dtype, device=self.device)

    def levels_memory(self, slots: int) -> torch.Tensor:
        """Zeroed field levels: (slots, *Nx_padded, lanes)."""
        return self._zeros(slots, *self.sim.Nx_padded)

    def aux_memory(self) -> list[torch.Tensor]:
        """Zeroed forward layer memory, (3, slab, lanes) per lined axis."""
        return [self._zeros(3, size) for _, size in self.slabs]

    def eta_memory(self) -> list[torch.Tensor]:
        """Zeroed adjoint layer memory, two alternating buffers per lined axis.

        A buffer holds (5, slab, lanes): the recursion's three variables and
        the two tap fields the neighbours' transposed taps read.
        """
        return [self._zeros(2, 5, size) for _, size in self.slabs]

    def record_memory(self, sensors: PointList) -> torch.Tensor:
        """Zeroed sensor record: (N, sensors, lanes)."""
        return self._zeros(self.sim.N, sensors.count)

    def accumulator(self) -> torch.Tensor:
        """Zeroed gradient accumulator over the padded grid; the kernels sum the shots into it."""
        return torch.zeros(self.sim.Nx_padded, dtype=self.sim.dtype, device=self.device)

    def shots(self, t: torch.Tensor) -> torch.Tensor:
        """The (nbatch, ...) view of a per-shot array whose last axis is the lanes axis."""
        return torch.movedim(t, -1, 0)[: self.nbatch]

    def lane_table(self, per_shot: torch.Tensor) -> torch.Tensor:
        """A (nbatch, ...) per-shot table as the kernels read it: (..., lanes), padding zero."""
        out = torch.zeros((*per_shot.shape[1:], self.lanes), dtype=self.sim.dtype,
                          device=self.device)
        out[..., : self.nbatch] = torch.movedim(per_shot.to(self.device, self.sim.dtype), 0, -1)
        return out

    def _layer(self, aux=None, eta=None) -> _abi.Pml:
        layer = _abi.Pml()
        ctypes.memmove(ctypes.byref(layer), ctypes.byref(self.pml), ctypes.sizeof(layer))
        for k, (d, size) in enumerate(self.slabs):
            if aux is not None:
                self._require_shape("pml memory", aux[k], (3, size, self.lanes))
                layer.aux[d] = _abi.ptr(aux[k])
            if eta is not None:
                self._require_shape("adjoint pml memory", eta[k], (2, 5, size, self.lanes))
                layer.eta[d][0] = _abi.ptr(eta[k][0])
                layer.eta[d][1] = _abi.ptr(eta[k][1])
        return layer

    def _require_shape(self, what: str, tensor: torch.Tensor | None, shape: tuple) -> None:
        # the kernels index raw pointers: a wrong shape would read past a buffer
        if tensor is not None and tuple(tensor.shape) != tuple(shape):
            raise ValueError(f"{what}: shape {tuple(tensor.shape)}, the kernels need {tuple(shape)}")
        if tensor is not None and tensor.dtype not in (self.sim.dtype, torch.int16):
            raise ValueError(f"{what}: dtype {tensor.dtype}, the kernels need {self.sim.dtype}")

    def levels(self, tensor: torch.Tensor, offset: int = 0) -> _abi.Levels:
        """Levels over a (slots, *Nx_padded, lanes) tensor, level k in slot (k + offset) % slots."""
        slots = tensor.shape[0]
        self._require_shape("levels", tensor, (slots, *self.sim.Nx_padded, self.lanes))
        return _abi.Levels(_abi.ptr(tensor), int(np.prod(tensor.shape[1:])), slots, offset)

    def _strip(self, strip: dict) -> _abi.Strip:
        self._require_shape("strip record", strip["record"],
                            (strip["rows"], strip["count"], self.lanes))
        if tuple(strip["map"].shape) != tuple(self.sim.Nx_padded):
            raise ValueError("strip map: one entry per node of the padded grid")
        return _abi.Strip(_abi.ptr(strip["map"]), _abi.ptr(strip["record"]), strip["count"],
                          strip["first"], strip["rows"], int(strip.get("bf16", 0)))

    # ------------------------------------------------------------ points
    def point_list(self, lin, shot, col, columns: int, strides: tuple[int, int, int],
                   weight=None) -> PointList:
        """Bucket points by the kernels' tiles and upload them (cw_point_lists).

        `strides` are the (row, column, shot) element strides of the signal or
        record table the points read or write.
        """
        lin = np.ascontiguousarray(lin, dtype=np.int32)
        count = len(lin)
        starts = np.zeros(self.ntiles + 1, dtype=np.int32)
        order = np.zeros(max(count, 1), dtype=np.int32)
        _abi.call("cw_point_lists", ctypes.byref(self.grid), lin.ctypes.data, count,
                  starts.ctypes.data, order.ctypes.data)
        order = order[:count]
        dev = self.device

        def upload(a, dtype):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)

        t_starts = upload(starts, torch.int32)
        t_lin = upload(lin[order], torch.int32)
        t_shot = upload(np.broadcast_to(np.asarray(shot, dtype=np.int32), (count,))[order], torch.int32)
        t_col = upload(np.asarray(col, dtype=np.int32)[order], torch.int32)
        keep = [t_starts, t_lin, t_shot, t_col]
        pts = _abi.Points(_abi.ptr(t_starts), _abi.ptr(t_lin), _abi.ptr(t_shot), _abi.ptr(t_col),
                          None, *map(int, strides))
        if weight is not None:
            t_w = upload(np.asarray(weight, dtype=self.sim.np_dtype)[order], self.sim.dtype)
            keep.append(t_w)
            pts.weight = _abi.ptr(t_w)
        return PointList(pts, tuple(keep), int(columns))

    # ------------------------------------------------------------ launches
    def _timed(self, kind: str, launches: int):
        import contextlib

        timings = Propagator.timings
        if timings is None or launches <= 0:
            return contextlib.nullcontext()

        @contextlib.contextmanager
        def bracket():
            stream = torch.cuda.current_stream()
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record(stream)
            yield
            end.record(stream)
            timings.append((kind, launches, start, end))

        return bracket()

    def forward(self, u: torch.Tensor, t0: int, t1: int, **kw) -> None:
        """Steps t0 .. t1 - 1 into the levels of `u` (cw_scalar_forward); keywords as
        `forward_args` takes them."""
        a = self.forward_args(u, **kw)
        with self._timed("forward", t1 - t0):
            _abi.call("cw_scalar_forward", ctypes.byref(a), int(t0), int(t1), _abi.stream())

    def march(self, u: torch.Tensor, t0: int, t1: int, window: int, **kw) -> None:
        """Steps t0 .. t1 - 1 as `forward` takes them, in windows of `window` steps
        replayed from one captured graph (cw_scalar_forward_clocked).

        The step index lives in a device counter that each window advances, so
        one capture serves every window: the first window runs eagerly (it warms
        whatever the library sets up on first use), one is captured, the graph is
        replayed for every further whole window, and the remainder runs eagerly.
        A short march, or `window` < 1, is one `forward` call."""
        if window < 1 or t1 - t0 < 3 * window:
            return self.forward(u, t0, t1, **kw)
        a = self.forward_args(u, **kw)
        clock = torch.full((1,), t0, dtype=torch.int32, device=self.device)

        def window_of(steps: int) -> None:
            _abi.call("cw_scalar_forward_clocked", ctypes.byref(a), _abi.ptr(clock), int(steps),
                      int(t1), _abi.stream())

        with self._timed("forward", t1 - t0):
            window_of(window)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                window_of(window)
            whole = (t1 - t0) // window - 1
            for _ in range(whole):
                graph.replay()
            rest = t1 - t0 - window * (whole + 1)
            if rest:
                window_of(rest)
            torch.cuda.current_stream().synchronize()  # the graph dies with this call
        del graph

    def forward_args(self, u: torch.Tensor, *, offset: int = 0, aux=None,
                     sources: PointList | None = None, signal: torch.Tensor | None = None,
                     sensors: PointList | None = None, record: torch.Tensor | None = None,
                     strip: dict | None = None,
                     illum: torch.Tensor | None = None) -> _abi.ForwardArgs:
        """The argument block of a forward sweep over the levels of `u`; `illum`, a
        per-node accumulator, gathers the illumination sum over steps and shots of u^2."""
        a = _abi.ForwardArgs()
        a.grid = self.grid
        a.mat = self.mat
        a.pml = self._layer(aux=aux)
        a.u = self.levels(u, offset)
        if sources is not None:
            self._require_shape("source signal", signal, (self.sim.N, sources.count))
            a.sources = sources.points
            a.signal = _abi.ptr(signal)
        if sensors is not None:
            self._require_shape("sensor record", record, (self.sim.N, sensors.count, self.lanes))
            a.sensors = sensors.points
            a.record = _abi.ptr(record)
        if strip is not None:
            a.strip = self._strip(strip)
        if illum is not None:
            self._require_shape("illumination", illum, tuple(self.sim.Nx_padded))
            a.illum = _abi.ptr(illum)
        a.steps = self.sim.N
        return a

    def backward(self, lam: torch.Tensor, u: torch.Tensor, n0: int, n1: int, *, eta,
                 g_mass: torch.Tensor, g_stiff: torch.Tensor | None = None, u_offset: int = 0,
                 reverse: bool = False, sensors: PointList | None = None,
                 adjoint: torch.Tensor | None = None, sources: PointList | None = None,
                 signal: torch.Tensor | None = None, strip: dict | None = None,
                 combine: tuple[float, float] | None = None, skip_layer: bool = False) -> None:
        """Adjoint levels n1 - 1 down to n0 (cw_scalar_backward)."""
        a = _abi.BackwardArgs()
        a.grid = self.grid
        a.mat = self.mat
        a.pml = self._layer(eta=eta)
        a.lam = self.levels(lam, 0)
        a.u = self.levels(u, u_offset)
        a.reverse = int(reverse)
        if sensors is not None:
            self._require_shape("adjoint signal", adjoint, (self.sim.N, sensors.count, self.lanes))
            a.sensors = sensors.points
            a.adjoint_signal = _abi.ptr(adjoint)
        if sources is not None:
            self._require_shape("source signal", signal, (self.sim.N, sources.count))
            a.sources = sources.points
            a.signal = _abi.ptr(signal)
        if strip is not None:
            a.strip = self._strip(strip)
        for what, acc in (("mass gradient", g_mass), ("stiffness gradient", g_stiff)):
            self._require_shape(what, acc, tuple(self.sim.Nx_padded))
        a.g_mass = _abi.ptr(g_mass)
        a.g_stiff = _abi.ptr(g_stiff)
        if combine is not None:
            a.combine = 1
            a.jac_mass, a.jac_stiff = float(combine[0]), float(combine[1])
        a.skip_layer = int(skip_layer)
        a.mass_factor = float(self.sim.np_dtype(self.sim.adjoint_mass_factor()))
        a.steps = self.sim.N
        with self._timed("backward", n1 - n0):
            _abi.call("cw_scalar_backward", ctypes.byref(a), int(n0), int(n1), _abi.stream())

    def field_gradient(self, u0, u1, u2, l1, g_mass, g_stiff) -> None:
        """Subtract the densities of the forward triplet (u0, u1, u2) against the
        adjoint field l1 from the accumulators (cw_scalar_gradient); one shot, every
        field over the padded grid."""
        if self.lanes != 1:
            raise ValueError("the field gradient takes one shot")
        sim = self.sim
        a = _abi.GradientArgs()
        a.grid = self.grid
        a.mat = self.mat
        for what, t in (("u0", u0), ("u1", u1), ("u2", u2), ("l1", l1),
                        ("mass gradient", g_mass), ("stiffness gradient", g_stiff)):
            self._require_shape(what, t, tuple(sim.Nx_padded))
        a.u0, a.u1, a.u2, a.l1 = _abi.ptr(u0), _abi.ptr(u1), _abi.ptr(u2), _abi.ptr(l1)
        a.g_mass, a.g_stiff = _abi.ptr(g_mass), _abi.ptr(g_stiff)
        # the operator L rather than dt^2 L (cuwave/scalar.py:775-778)
        a.inv_dt2 = float(sim.np_dtype(1.0 / sim.dt**2))
        for d, f in enumerate(sim.step_factors()):
            a.factor[d] = float(sim.np_dtype(float(f) / sim.dt**2))
        _abi.call("cw_scalar_gradient", ctypes.byref(a), _abi.stream())

    def superposed(self, u0, u1, u2, acc_mass, acc_stiff, ft: float, fs: float, step: bool) -> None:
        """One superposition step, or the Frechet densities of a triplet (cw_scalar_superposed)."""
        if self.lanes != 1:
            raise ValueError("the superposition step takes one shot")
        a = _abi.SuperposedArgs()
        a.grid = self.grid
        a.mat = self.mat
        for what, t in (("u0", u0), ("u1", u1), ("u2", u2), ("mass density", acc_mass),
                        ("stiffness density", acc_stiff)):
            self._require_shape(what, t, tuple(self.sim.Nx_padded))
        a.u0, a.u1, a.u2 = _abi.ptr(u0), _abi.ptr(u1), _abi.ptr(u2)
        a.acc_mass, a.acc_stiff = _abi.ptr(acc_mass), _abi.ptr(acc_stiff)
        a.ft, a.fs, a.step = float(ft), float(fs), int(step)
        _abi.call("cw_scalar_superposed", ctypes.byref(a), _abi.stream())


# ------------------------------------------------------------------ shot points
def source_points(prop: Propagator, mat: dict, sources,
                  damping="sim") -> tuple[PointList, torch.Tensor]:
    """Every shot's source points and one (N, columns) signal table, shot b's columns its own.

    The points of all shots share one list: each entry names its shot, so a
    launch adds shot b's sources to shot b's field only (cuwave/wave.py:985-993).
    """
    from .wave import as_numpy, require_stepped

    sim = prop.sim
    lins, shots, cols, weights, signals = [], [], [], [], []
    column = 0
    for b, source in enumerate(sources):
        pos = np.asarray(as_numpy(source.position))
        lin = sim.flatten_indices(pos)
        require_stepped(sim, pos, "a source")
        sig = as_numpy(source.signal)
        if sig.ndim != 2 or sig.shape[0] < sim.N or sig.shape[1] != len(lin):
            raise ValueError(f"shot {b}: signal needs ({sim.N}, {len(lin)}), not {sig.shape}")
        lins.append(lin)
        shots.append(np.full(len(lin), b, dtype=np.int32))
        cols.append(column + np.arange(len(lin), dtype=np.int32))
        weights.append(sim.excitation_weights(mat, lin, damping=damping))
        signals.append(np.asarray(sig[: sim.N], dtype=sim.np_dtype))
        column += len(lin)
    table = np.ascontiguousarray(np.concatenate(signals, axis=1), dtype=sim.np_dtype)
    pts = prop.point_list(np.concatenate(lins), np.concatenate(shots), np.concatenate(cols),
                          column, (column, 1, 0), weight=np.concatenate(weights))
    return pts, torch.as_tensor(table, device=prop.device)


def sensor_points(prop: Propagator, sensors, weight=None) -> PointList:
    """Sensor nodes every shot shares, column k the k-th sensor.

    They write (or, weighted, read) an (N, sensors, lanes) table: the sensor
    record, or the adjoint signal `lane_table` lays out.
    """
    from .wave import require_stepped

    lin = prop.sim.flatten_indices(sensors)
    require_stepped(prop.sim, sensors, "a sensor")
    count, lanes = len(lin), prop.lanes
    return prop.point_list(lin, -1, np.arange(count, dtype=np.int32), count,
                           (count * lanes, lanes, 1), weight=weight)
