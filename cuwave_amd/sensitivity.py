"""Adjoint sensitivities, same arguments and same return as the reference's.

`sensitivity` stores the forward field (N + 2 grids) and steps the exact
transpose of the discretisation backwards (cuwave/sensitivity.py:384-497).
`reconstruction_sensitivity` stores only the strip of nodes the PML makes
irreversible and rebuilds the forward field by a reverse march fused into the
adjoint step itself (cuwave/sensitivity.py:500-750); `checkpoints` cuts the run
into segments replayed from snapshots in treeverse order, so the memory no longer
grows with N. `superposition_sensitivity` reads the cross term off the diagonal of
B(u + k lambda) in three slots, for a lossless operator (cuwave/sensitivity.py:
753-893).

All return `(cost, {"mass": ..., "stiff": ...}, traces, info)`: the gradients
with respect to the inertia and stiffness coefficients over the padded grid,
as torch tensors on the GPU.
"""

from __future__ import annotations

import itertools
import math
import warnings

import numpy as np
import scipy.ndimage as ndi
import torch

from .boundary import pml_nodes
from .engine import Propagator, sensor_points, source_points
from .wave import Simulation, Source, as_numpy, family_drivers

# cancellation past which B(w, w) - B(u, u) has eaten too much of the mantissa
CANCELLATION_LIMIT = {"float32": 1e5, "float64": 1e12}
DECAY_LIMIT = 1e-2  # the share of a spectrum the last period may still add
FIT_LIMIT = {"float32": 1e-4, "float64": 1e-10}  # how far dcost/dtraces may sit off the sinusoids
PULSE_PERIODS = 1.0  # the adjoint pulse's Gaussian width, in periods
ADJOINT_DELAY = 1  # lines the superposition adjoint up with the forward triplet


# ------------------------------------------------------------------ objectives
def l2_misfit(observed):
    """Objective factory: J = 1/2 sum (traces - observed)^2, and its derivative."""

    def objective(traces):
        data = torch.as_tensor(as_numpy(observed) if not isinstance(observed, torch.Tensor)
                               else observed, device=traces.device, dtype=traces.dtype)
        residual = traces - data
        return 0.5 * float(torch.sum(residual**2)), residual

    return objective


def windowed_misfit(observed, window):
    """`l2_misfit` restricted to a time window, its derivative carrying the window twice."""

    def objective(traces):
        data = torch.as_tensor(as_numpy(observed), device=traces.device, dtype=traces.dtype)
        w = torch.as_tensor(as_numpy(window), device=traces.device, dtype=traces.dtype)
        residual = w * (traces - data)
        return 0.5 * float(torch.sum(residual**2)), w * residual

    return objective


def adjoint_signal(sim: Simulation, dphi, sensors, scale: float = 1.0, delay: int = 0) -> torch.Tensor:
    """The adjoint excitation: `dphi` reversed in time, over W and the source factor.

    (cuwave/sensitivity.py:272-292). `delay` drops entries off the front and pads
    zeros at the back; `scale` is the superposition k.
    """
    dphi = torch.as_tensor(dphi, dtype=sim.dtype)
    w = torch.as_tensor(sim.adjoint_weights(sensors), device=dphi.device)
    signal = torch.flip(dphi, (0,)) / w
    if delay:
        signal = torch.cat((signal[delay:], torch.zeros_like(signal[:delay])))
    if scale != 1.0:
        signal = signal * torch.tensor(scale, dtype=sim.dtype, device=signal.device)
    return signal.contiguous()


# --------------------------------------------------------------------- helpers
def _elastic(sim: Simulation) -> bool:
    """Whether `sim` is of a vector family (elastic, cell-assembled elastic, vector
    Maxwell), which the pressure family's drivers do not step."""
    return sim.ncomp > 1 or type(sim).__name__ == "ElasticWave" or family_drivers(sim) is not None


def _require(sim: Simulation, source: Source, sensors) -> None:
    """The reference's checks before a sensitivity run: points off the ghosts
    (cuwave/sensitivity.py:430-432, 546-548)."""
    sim.require_interior(sensors, "sensor")
    sim.require_interior(source.position, "source")


def require_full_domain(sim: Simulation) -> None:
    """Raise if `sim.domain` or `sim.dirichlet` is set: the adjoint kernels grade
    no domain wall (cuwave/sensitivity.py:121-126)."""
    if sim.masked:
        raise NotImplementedError(
            "the material gradients step every node: unset domain and dirichlet")


def _finalize(sim: Simulation, grads: dict, device) -> dict:
    """Weight the accumulators by the cell weights W (cuwave/wave.py:336-344)."""
    w = torch.as_tensor(sim.cell_weights(), device=device)
    for f in grads.values():
        f.mul_(w)
    return grads


def _zero_outside(grads: dict, mask: np.ndarray, device) -> dict:
    """Assign zero where `mask` holds, so a NaN there cannot survive as 0 * NaN."""
    m = torch.as_tensor(mask, device=device)
    for f in grads.values():
        f[m] = 0.0
    return grads


def reconstruction_nodes(sim: Simulation, indicator=None, span: int | None = None):
    """The nodes a reverse march replays from a record, and those it rebuilds exactly.

    Run backwards, the lossless step of a node is exact; a layer node's memory
    terms decay forwards and cannot be run back. A reversible node's reverse step
    reads every node within `sim.reach` of it along each axis and across the
    diagonals (a box of side 2 reach + 1), so the record must hold each layer node
    inside such a box around some reversible node (cuwave/sensitivity.py:160-195).

    Returns the (ndim, num) record nodes in row-major order and the mask of the
    reversible interior nodes, where the rebuilt field is exact.
    """
    inside = np.zeros(sim.Nx_padded, dtype=bool)
    inside[tuple(slice(1, n - 1) for n in sim.Nx)] = True
    if sim.pml is None and sim.damping is None:
        return np.zeros((sim.ndim, 0), dtype=np.int32), inside
    exact = inside.copy()
    if sim.pml is not None:
        exact &= ~pml_nodes(sim)
    if sim.damping is not None:
        # a damped node runs back only where its growth stays bounded over the span
        span = sim.N if span is None else span
        exact &= ~(sim.damping > 0) | reversible_damping(sim, indicator, span)
    # a box maximum is the dilation by the full square, taken `reach` times over
    seen = ndi.maximum_filter(exact.astype(np.uint8), size=2 * sim.reach + 1,
                              mode="constant", cval=0).astype(bool)
    record = np.argwhere(inside & ~exact & seen).T.astype(np.int32)
    return record, exact


REVERSIBLE_GROWTH = {"float32": 1e2, "float64": 1e6}  # how far a reverse march may amplify


def reversible_damping(sim: Simulation, indicator, span: int) -> np.ndarray:
    """Damped nodes a reverse march of `span` steps can run backwards
    (cuwave/sensitivity.py:139-157). Solved for its trailing slot the damped update
    is the update with the damping negated, exactly invertible; what running it back
    costs is the decay turned growth, (1 + beta) / (1 - beta) a step for the static
    mode, which must stay under REVERSIBLE_GROWTH over the span. beta is taken at the
    largest inverse inertia of the node's 3^ndim box."""
    if sim.damping is None:
        return np.zeros(sim.Nx_padded, dtype=bool)
    minv = ndi.maximum_filter(sim.inverse_inertia(indicator), size=3)
    beta = 0.5 * sim.dt * sim.damping * minv
    tiny = np.finfo(sim.np_dtype).tiny
    with np.errstate(over="ignore"):  # beta >= 1 overflows to inf: irreversible
        growth = span * np.log((1.0 + beta) / np.maximum(1.0 - beta, tiny))
    bound = math.log(REVERSIBLE_GROWTH[sim.precision])
    return (sim.damping > 0) & (beta < 1.0) & (growth <= bound)


def _reversible_segments(snapshots: int, replays: int) -> int:
    """Segments a binomial treeverse reverses with `snapshots` spare snapshots when no
    segment is marched more than `replays` extra times: C(snapshots + replays, snapshots)
    (Griewank & Walther 2000, https://doi.org/10.1145/347837.347846)."""
    return math.comb(snapshots + replays, snapshots)


def segment_split(lo: int, hi: int, free: int) -> int:
    """The segment at which treeverse over segments lo .. hi - 1 places its next snapshot.

    The replay count is the smallest that `free` snapshots allow for the span; the
    segments after the snapshot get exactly what one snapshot fewer reverses at that
    count, and the ones before it the remainder (never fewer than one).
    """
    replays = next(r for r in itertools.count() if _reversible_segments(free, r) >= hi - lo)
    return max(lo + 1, hi - _reversible_segments(free - 1, replays))


def checkpoint_schedule(segments: int, slots: int) -> list[tuple]:
    """The reversal order of `segments` with `slots` snapshots besides the zero start.

    Built with an explicit stack of pending ranges (lo, hi, free snapshots, the
    snapshot to release when the range is done). Actions: ("restore", k) loads the
    start state of segment k, ("store", k) snapshots it, ("drop", k) releases it,
    ("advance", a, b) marches from segment a's start to segment b's, ("leaf", k)
    marches segment k recording its strip and then reverses it. The schedule is the
    reference's (cuwave/sensitivity.py:243-269): the `info` sweep count depends on it.
    """
    actions: list[tuple] = []
    pending = [[0, segments, slots, None]]
    while pending:
        top = pending[-1]
        lo, hi, free, _ = top
        if hi - lo > 1 and free > 0:
            mid = segment_split(lo, hi, free)
            actions += [("restore", lo), ("advance", lo, mid), ("store", mid)]
            top[1] = mid  # what is left of this range once the part past mid is reversed
            pending.append([mid, hi, free - 1, mid])
            continue
        # no snapshot to spare: every segment past lo replays from lo, the last first
        for k in range(hi - 1, lo, -1):
            actions += [("restore", lo), ("advance", lo, k), ("leaf", k)]
        actions += [("restore", lo), ("leaf", lo)]
        pending.pop()
        if top[3] is not None:
            actions.append(("drop", top[3]))
    return actions


# ------------------------------------------------------------------ sensitivity
def sensitivity(sim: Simulation, source: Source, indicator, sensors, objective):
    """Cost and d(cost)/d(mass, stiff) by the exact adjoint over a stored history.

    The forward field is kept for all N + 2 levels; the adjoint steps the exact
    transpose with both gradient densities fused in and the transposed PML, the
    layer then frozen out of the gradient (cuwave/sensitivity.py:384-497).
    """
    own = family_drivers(sim)
    if own is not None:
        return own.sensitivity(sim, source, indicator, sensors, objective)
    if _elastic(sim):
        from . import elastic

        return elastic.sensitivity(sim, source, indicator, sensors, objective)
    _require(sim, source, sensors)
    require_full_domain(sim)
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, 1)
    dev = prop.device
    src, signal = source_points(prop, mat, [source])
    rec = sensor_points(prop, sensors)
    V = prop.levels_memory(sim.N + 2)
    record = prop.record_memory(rec)
    prop.forward(V, 0, sim.N, offset=2, aux=prop.aux_memory(), sources=src, signal=signal,
                 sensors=rec, record=record)
    um = prop.shots(record)[0]
    cost, dphi = objective(um)
    adjoint = prop.lane_table(adjoint_signal(sim, dphi, sensors)[None])
    lin = sim.flatten_indices(sensors)
    adj = sensor_points(prop, sensors, weight=sim.excitation_weights(mat, lin))
    P = prop.levels_memory(2)
    acc = {k: prop.accumulator() for k in ("mass", "stiff")}
    prop.backward(P, V, 0, sim.N, u_offset=2, eta=prop.eta_memory(), g_mass=acc["mass"],
                  g_stiff=acc["stiff"], sensors=adj, adjoint=adjoint)
    grads = _finalize(sim, acc, dev)
    if sim.pml is not None:
        grads = _zero_outside(grads, pml_nodes(sim), dev)
    return cost, grads, um, {}


# ------------------------------------------------------------ source gradient
def source_sensitivity(sim: Simulation, source: Source, indicator, sensors, objective):
    """Cost and d(cost)/d(source.signal): the adjoint field read at the source.

    The traces are linear in the signal, so the gradient pairs no forward field
    with the adjoint and neither history is kept (cuwave/sensitivity.py:1054-1111).
    The adjoint marches alone in its two slots, its gradient densities switched
    off, one level a call, and each level is read at the source nodes before
    the next overwrites it (a launch a level, as the reference's probe is; this
    is not a batched hot path). The rows are then weighted as `adjoint_signal`
    divides (its transpose).

    Returns (cost, gradient, traces, {}), the gradient (N, source points) as
    `source.signal` is laid out.
    """
    own = family_drivers(sim)
    if own is not None:
        # the vector Maxwell and cell-assembled adjoints are the forward step itself
        # (no pml; damped or not): the load marched forward from the sensors and read
        # at the sources, reversed in time
        _require(sim, source, sensors)
        _, um = own.simulate(sim, source, indicator, sensors)
        cost, dphi = objective(um)
        dphi = torch.as_tensor(dphi, dtype=sim.dtype, device=um.device)
        w_in = torch.as_tensor(np.asarray(sim.adjoint_weights(sensors)), device=um.device)
        load = Source(sensors, torch.flip(dphi, (0,)) / w_in)
        _, lam = own.simulate(sim, load, indicator, source.position)
        w = torch.as_tensor(np.asarray(sim.adjoint_weights(source.position)), device=lam.device)
        return cost, lam.flip(0) * w, um, {}
    if _elastic(sim):
        raise NotImplementedError("source_sensitivity is built for the scalar equations")
    _require(sim, source, sensors)
    from .wave import simulate

    _, um = simulate(sim, source, indicator, sensors)
    cost, dphi = objective(um)
    if sim.masked:
        # a mask grades its cells per node, which the transposed kernel does not
        # read: the recursion is stepped by the forward step itself, exact where
        # the graded operator is symmetric (order 2) (cuwave/scalar.py:321-323),
        # the adjoint load injected at the sensors and lambda read at the
        # sources as a forward run records
        load = Source(sensors, adjoint_signal(sim, dphi, sensors))
        _, lam = simulate(sim, load, indicator, source.position)
        w = torch.as_tensor(np.asarray(sim.adjoint_weights(source.position)), device=lam.device)
        return cost, lam.flip(0) * w, um, {}
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, 1)
    dev = prop.device
    adjoint = prop.lane_table(adjoint_signal(sim, dphi, sensors)[None])
    lin = sim.flatten_indices(sensors)
    adj = sensor_points(prop, sensors, weight=sim.excitation_weights(mat, lin))
    P = prop.levels_memory(2)
    U = prop.levels_memory(2)  # zero: the densities the kernel adds are zero
    acc = prop.accumulator()
    eta = prop.eta_memory()
    at = torch.as_tensor(sim.flatten_indices(source.position), device=dev, dtype=torch.int64)
    slots = P.view(2, -1, prop.lanes)
    lam = torch.empty((sim.N, len(at)), dtype=sim.dtype, device=dev)
    for n in range(sim.N - 1, -1, -1):
        prop.backward(P, U, n, n + 1, eta=eta, g_mass=acc, combine=(0.0, 0.0), sensors=adj,
                      adjoint=adjoint)
        lam[n] = slots[n % 2, at, 0]
    w = torch.as_tensor(np.asarray(sim.adjoint_weights(source.position)), device=dev)
    return cost, lam * w, um, {}


# ------------------------------------------------------ reconstruction (strip)
class _Reconstruction:
    """The forward/backward machinery of the strip-replayed reverse march over a batch.

    One object drives `reconstruction_sensitivity` (one shot) and the batched
    misfit gradient: a forward sweep that records the traces and the strip, then
    adjoint levels that rebuild the forward field one level ahead of the adjoint
    in the same kernel.
    """

    def __init__(self, sim, mat, sources, sensors, strip_dtype=None, checkpoints=None,
                 illum=None, indicator=None):
        if sim.damping is not None and indicator is None:
            raise ValueError("a damped reconstruction needs the indicator (its inertia "
                             "scales how far the damping can be run back)")
        self.sim, self.mat = sim, mat
        self.illum = illum  # a per-node accumulator the first forward pass adds u^2 into
        self.prop = prop = Propagator(sim, mat, len(sources))
        self.dev = dev = prop.device
        self.B = len(sources)
        self.src, self.signal = source_points(prop, mat, sources)
        self.rec = sensor_points(prop, sensors)
        self.sensors = sensors
        self.nrec = self.rec.count
        self.bf16 = strip_dtype == "bfloat16"
        if strip_dtype not in (None, "bfloat16"):
            raise ValueError(f"strip_dtype {strip_dtype!r}: None or 'bfloat16'")
        if self.bf16 and sim.precision != "float32":
            raise NotImplementedError("a bfloat16 strip is built for float32 runs")
        strip, self.valid = reconstruction_nodes(sim, indicator)
        self.U = prop.levels_memory(2)
        self.aux = prop.aux_memory()
        itemsize = 2 if self.bf16 else np.dtype(sim.np_dtype).itemsize
        snapshot = 2 * self.U[0].nbytes + sum(a.nbytes for a in self.aux)
        segment = sim.N
        if checkpoints is not None and strip.shape[1]:
            # the strip of a segment costs about one snapshot (cuwave/sensitivity.py:575-583)
            segment = min(sim.N, max(2, snapshot // (2 * strip.shape[1] * itemsize * prop.lanes)))
        if segment < sim.N:  # a damped node reverses over a segment more easily
            strip, self.valid = reconstruction_nodes(sim, indicator, segment)
        if not self.valid.any():
            raise ValueError("damping too strong to reverse covers every interior node, so "
                             "there is nothing to reconstruct; set checkpoints, damp only the "
                             "faces with boundary.sponge, or use sensitivity")
        self.count = strip.shape[1]
        smap = np.full(sim.Nx_padded, -1, dtype=np.int32)
        if self.count:
            smap.ravel()[sim.flatten_indices(strip)] = np.arange(self.count, dtype=np.int32)
        self.strip_map = torch.as_tensor(smap, device=dev)
        self.src_back = self.src
        if sim.damping is not None:
            # backwards a reversible damping is negated (cuwave/sensitivity.py:592-598);
            # the reverse march's sources share its divisor
            back = np.where(self.valid, -sim.damping, sim.damping).astype(sim.np_dtype)
            prop.reverse_damping(torch.as_tensor(back, device=dev))
            self.src_back, _ = source_points(prop, mat, sources, damping=back)
        self.segment = segment
        self.checkpoints = checkpoints or 0
        self.starts = list(range(0, sim.N, segment)) + [sim.N]
        self.record = prop.record_memory(self.rec)
        self.recorded = 0
        self.steps = 0
        self.snapshots = {}

    def strip_window(self, s: int, e: int) -> dict | None:
        if not self.count:
            return None
        first = -1 if s == 0 else s  # a zero row stands for step -1
        rows = e - first
        dtype = torch.int16 if self.bf16 else self.sim.dtype
        buf = torch.zeros((rows, self.count, self.prop.lanes), dtype=dtype, device=self.dev)
        return {"map": self.strip_map, "record": buf, "count": self.count, "first": first,
                "rows": rows, "bf16": int(self.bf16)}

    def advance(self, a: int, b: int, strip=None) -> None:
        first = a == self.recorded
        if first:
            self.recorded = b
        self.prop.forward(self.U, a, b, aux=self.aux, sources=self.src, signal=self.signal,
                          sensors=self.rec if first else None,
                          record=self.record if first else None, strip=strip,
                          illum=self.illum if first else None)
        self.steps += b - a

    def store(self, k: int) -> None:
        self.snapshots[k] = [f.clone() for f in (self.U, *self.aux)]

    def restore(self, k: int) -> None:
        fields = (self.U, *self.aux)
        if k == 0:
            for f in fields:
                f.zero_()
            return
        for f, saved in zip(fields, self.snapshots[k]):
            f.copy_(saved)

    def run(self, objectives, gradient_mode):
        """Forward, objectives, adjoint; returns per-shot costs.

        `objectives(record)` takes the (B, N, nrec) record and returns a list of
        costs and the (B, N, nrec) derivative; `gradient_mode` is a dict of the
        accumulators and how the kernel adds into them.
        """
        sweep = self.sweep(gradient_mode)
        costs, dphi = objectives(next(sweep))
        try:
            sweep.send(dphi)
        except StopIteration:
            pass
        return costs

    def sweep(self, gradient_mode):
        """The whole schedule as a generator: it yields the (B, N, nrec) record
        once, when the adjoint first needs its source, and takes back the
        (B, N, nrec) derivative of the cost with respect to it (so the forward
        half can run in one call and the adjoint half in a later one)."""
        self.lam = self.prop.levels_memory(2)
        self.eta = self.prop.eta_memory()
        self.adjoint = None
        self.seed = None
        self.gradient_mode = gradient_mode
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

    def leaf(self, k: int):
        sim, prop = self.sim, self.prop
        s, e = self.starts[k], self.starts[k + 1]
        strip = self.strip_window(s, e)
        self.advance(s, e, strip)
        if self.adjoint is None:
            dphi = yield prop.shots(self.record)
            w = torch.as_tensor(sim.adjoint_weights(self.sensors), device=self.dev)
            dphi = torch.as_tensor(dphi, dtype=sim.dtype, device=self.dev)
            self.adjoint = prop.lane_table(torch.flip(dphi, (1,)) / w)
            lin = sim.flatten_indices(self.sensors)
            self.adj = sensor_points(prop, self.sensors, weight=sim.excitation_weights(self.mat, lin))
        if s == 0:
            self.seed = self._valid_norm(self.U[(e - 1) % 2])
        common = dict(eta=self.eta, sensors=self.adj, adjoint=self.adjoint,
                      sources=self.src_back, signal=self.signal, strip=strip,
                      **self.gradient_mode)
        # the top level reads u^(e - 1) as the forward left it; below it the
        # reverse march rebuilds u^(n - 1) one level ahead
        prop.backward(self.lam, self.U, e - 1, e, reverse=False, **common)
        low = s if s == 0 else s + 1
        if e - 1 > low:
            prop.backward(self.lam, self.U, low, e - 1, reverse=True, **common)
        if s > 0:
            prop.backward(self.lam, self.U, s, s + 1, reverse=False, **common)

    def _valid_norm(self, level: torch.Tensor) -> float:
        valid = torch.as_tensor(self.valid, device=self.dev)
        return float(torch.linalg.vector_norm(self.prop.shots(level) * valid))

    def drift(self) -> float:
        """Norm of the rebuilt u^(-1), which is zero, over that of the top state."""
        rebuilt = self._valid_norm(self.U[1])
        return rebuilt / self.seed if self.seed else float("inf")


def reconstruction_sensitivity(sim: Simulation, source: Source, indicator, sensors, objective,
                               strip_dtype: str | None = None, checkpoints: int | None = None):
    """Cost and gradients as `sensitivity`, storing a boundary strip and not the field.

    The lossless step run with its outer slots swapped marches the forward field
    backwards; the nodes the PML makes irreversible are recorded as a strip and
    replayed (cuwave/sensitivity.py:500-750). Here the reverse step rides inside
    the adjoint kernel, one level ahead of the adjoint. `info` carries the
    `strip` node count, the `drift` of the rebuilt initial state, the steps per
    `segment` and the forward `sweeps`.
    """
    own = family_drivers(sim)
    if own is not None:
        return own.reconstruction_sensitivity(sim, source, indicator, sensors, objective,
                                              strip_dtype, checkpoints)
    if _elastic(sim):
        from . import elastic

        return elastic.reconstruction_sensitivity(sim, source, indicator, sensors, objective,
                                                  strip_dtype, checkpoints)
    _require(sim, source, sensors)
    require_full_domain(sim)
    if checkpoints is not None and checkpoints < 0:
        raise ValueError(f"checkpoints must be non-negative: {checkpoints}")
    mat = sim.build_materials(indicator)
    run = _Reconstruction(sim, mat, [source], sensors, strip_dtype, checkpoints,
                          indicator=indicator)
    dev = run.dev
    acc = {k: run.prop.accumulator() for k in ("mass", "stiff")}

    def objectives(record):
        cost, dphi = objective(record[0])
        return [cost], torch.as_tensor(dphi, dtype=sim.dtype, device=dev)[None]

    costs = run.run(objectives, dict(g_mass=acc["mass"], g_stiff=acc["stiff"], skip_layer=True))
    grads = _finalize(sim, acc, dev)
    grads = _zero_outside(grads, ~run.valid, dev)
    info = {"strip": run.count, "drift": run.drift(), "segment": run.segment,
            "sweeps": run.steps / sim.N}
    return costs[0], grads, run.prop.shots(run.record)[0], info


# ---------------------------------------------------------------- superposition
def _refill(sim: Simulation, u: torch.Tensor) -> None:
    """The ghost refill of every face, as the reference's refill kernels (scalar.cu:535-555)."""
    from .boundary import Dirichlet

    for d, pair in enumerate(sim.boundary):
        for side, cond in enumerate(pair):
            ghost, mirror, wall = (0, 2, 1) if side == 0 else (sim.Nx[d] - 1, sim.Nx[d] - 3, sim.Nx[d] - 2)
            inner = [slice(1, n - 1) for n in sim.Nx]
            g_ix, m_ix, w_ix = list(inner), list(inner), list(inner)
            g_ix[d], m_ix[d], w_ix[d] = ghost, mirror, wall
            if cond == Dirichlet:
                u[tuple(g_ix)] = -u[tuple(m_ix)]
                u[tuple(w_ix)] = 0.0
            else:
                u[tuple(g_ix)] = u[tuple(m_ix)]


def superposition_sensitivity(sim: Simulation, source: Source, indicator, sensors, objective,
                              scale: float = 1.0):
    """Cost and gradients in three field slots, for a lossless operator.

    B(u, lambda) = [B(w, w) - B(u, u)] / 2k with w = u + k lambda rebuilt by
    running the closed lossless domain backwards (cuwave/sensitivity.py:753-893).
    `info` carries `scale` and the `cancellation` the subtraction cost.
    """
    if _elastic(sim):
        raise NotImplementedError("superposition_sensitivity is built for the pressure family; "
                                  "use sensitivity or reconstruction_sensitivity")
    if sim.pml is not None or sim.damping is not None:
        raise NotImplementedError(
            "superposition_sensitivity runs the field backwards, which only a lossless "
            "operator allows (no pml, no damping); use sensitivity or "
            "reconstruction_sensitivity"
        )
    _require(sim, source, sensors)
    require_full_domain(sim)
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, 1)
    dev = prop.device
    npd = sim.np_dtype
    lin_src = torch.as_tensor(sim.flatten_indices(source.position).astype(np.int64), device=dev)
    w_src = torch.as_tensor(sim.excitation_weights(mat, sim.flatten_indices(source.position)), device=dev)
    sig = torch.as_tensor(np.asarray(as_numpy(source.signal), dtype=npd)[: sim.N], device=dev)
    lin_rec = torch.as_tensor(sim.flatten_indices(sensors).astype(np.int64), device=dev)
    accs = {k: torch.zeros(sim.Nx_padded, dtype=sim.dtype, device=dev) for k in ("mass", "stiff")}
    factors = sim.step_factors()

    def coefficients(sign):
        ratios = [-sign / (2.0 * d) ** 2 / float(f) for d, f in zip(sim.dx, factors)]
        if not np.allclose(ratios, ratios[0]):
            raise ValueError(f"step factors {factors} are not proportional to 1 / dx**2")
        return float(npd(sign / (2.0 * sim.dt) ** 2)), float(npd(ratios[0]))

    def frechet(sign, u0, u1, u2):
        # the mass density alone: the stiffness factors are zero (cuwave/scalar.py:836-856)
        saved = [prop.grid.f[d] for d in range(sim.ndim)]
        for d in range(sim.ndim):
            prop.grid.f[d] = float(npd(-sign * 0 / (2.0 * sim.dx[d]) ** 2))
        prop.superposed(u0, u1, u2, accs["mass"], accs["stiff"], float(npd(sign / (2.0 * sim.dt) ** 2)),
                        0.0, step=False)
        for d in range(sim.ndim):
            prop.grid.f[d] = saved[d]

    def excite(u, lin, w, row):
        u.view(-1).index_add_(0, lin, w * row)

    U = torch.zeros((3, *sim.Nx_padded), dtype=sim.dtype, device=dev)
    um = torch.zeros((sim.N, len(lin_rec)), dtype=sim.dtype, device=dev)
    ft, fs = coefficients(-1.0)
    for t in range(sim.N):
        u0, u1, u2 = U[t % 3], U[(t + 1) % 3], U[(t + 2) % 3]
        prop.superposed(u0, u1, u2, accs["mass"], accs["stiff"], ft, fs, step=True)
        excite(u2, lin_src, w_src, sig[t])
        _refill(sim, u2)
        um[t] = u2.view(-1)[lin_rec]
    t = sim.N - 1
    frechet(-1.0, U[t % 3], U[(t + 1) % 3], U[(t + 2) % 3])
    cost, dphi = objective(um)
    before = math.sqrt(sum(float(torch.linalg.vector_norm(f)) ** 2 for f in accs.values()))
    back_lin = torch.cat((lin_rec, lin_src))
    back_w = torch.cat((torch.as_tensor(sim.excitation_weights(mat, sim.flatten_indices(sensors)), device=dev), w_src))
    back_sig = torch.cat((adjoint_signal(sim, dphi, sensors, scale, ADJOINT_DELAY),
                          torch.flip(sig, (0,))), dim=1).contiguous()
    ft, fs = coefficients(1.0)

    def backward(t, row, mass=True):
        u0, u1, u2 = (U[(sim.N + i - t) % 3] for i in (1, 0, 2))
        prop.superposed(u0, u1, u2, accs["mass"], accs["stiff"], ft if mass else 0.0, fs, step=True)
        excite(u2, back_lin, back_w, back_sig[row])
        _refill(sim, u2)

    backward(0, 0, mass=False)
    for t in range(sim.N - 1):
        backward(t + 1, t + 1)
    last = [U[(1 + i) % 3] for i in (1, 0, 2)]
    frechet(1.0, *last)
    after = math.sqrt(sum(float(torch.linalg.vector_norm(f)) ** 2 for f in accs.values()))
    cancellation = before / after if after > 0.0 else float("inf")
    limit = CANCELLATION_LIMIT[sim.precision]
    if cancellation > limit:
        warnings.warn(
            f"superposition scale={scale:g} leaves a cancellation of {cancellation:.1e} in "
            f"{sim.precision}, past the usable {limit:.0e}: the gradient is largely round-off. "
            f"Raise scale by about {cancellation / limit:.0e}.",
            RuntimeWarning,
            stacklevel=2,
        )
    accs = _finalize(sim, accs, dev)
    norm = torch.tensor(npd(1.0 / (2.0 * scale)), device=dev)
    for f in accs.values():
        f.mul_(norm)
    return cost, accs, um, {"scale": scale, "cancellation": cancellation}


# ----------------------------------------------------------- spectral gradient
def phase_table(sim: Simulation, frequencies) -> np.ndarray:
    """(N, 2F) float64: cos(w t) and -sin(w t) per frequency, interleaved; the
    running DFT of a field is its sum over t against these rows
    (cuwave/sensitivity.py:295-303). Built in double: the phase reaches 1e5 rad."""
    w = 2.0 * np.pi * np.asarray(frequencies, dtype=np.float64)
    phase = np.outer(np.arange(sim.N, dtype=np.float64) * sim.dt, w)
    return np.stack((np.cos(phase), -np.sin(phase)), axis=2).reshape(sim.N, -1)


def spectral_load(sim: Simulation, dphi, table: np.ndarray):
    """Least-squares fit dphi = sum_f a_f cos(w_f t) + b_f sin(w_f t) per sensor
    (cuwave/sensitivity.py:306-323). Returns the (F, sensors) complex loads
    a + i b and the relative residual of the fit."""
    basis = table * np.tile([1.0, -1.0], table.shape[1] // 2)  # cos, sin columns
    target = np.asarray(as_numpy(dphi), dtype=np.float64)
    coef = np.linalg.lstsq(basis, target, rcond=None)[0]
    norm = float(np.linalg.norm(target))
    fit = float(np.linalg.norm(target - basis @ coef)) / norm if norm > 0.0 else 0.0
    return coef[0::2] + 1j * coef[1::2], fit


def spectral_pulse(sim: Simulation, frequency: float, load) -> np.ndarray:
    """(N, columns) Gaussian-windowed carriers whose DTFT at `frequency` is `load`
    (cuwave/sensitivity.py:326-346): a cos and a sin carrier under one envelope
    of width PULSE_PERIODS periods centred at four widths, mixed per column so
    the pair's transform at w equals the complex load. Raises when eight widths
    outlast the run."""
    sigma = PULSE_PERIODS / frequency
    t = np.arange(sim.N, dtype=np.float64) * sim.dt
    if 8.0 * sigma > t[-1]:
        raise ValueError(f"the adjoint pulse spans {8.0 * sigma:.3g}, longer than the run "
                         f"{t[-1]:.3g}")
    w = 2.0 * np.pi * frequency
    window = np.exp(-0.5 * ((t - 4.0 * sigma) / sigma) ** 2)
    carriers = np.stack((window * np.cos(w * t), window * np.sin(w * t)), axis=1)  # (N, 2)
    at_w = np.exp(-1j * w * t) @ carriers  # each carrier's DTFT at w
    load = np.atleast_1d(np.asarray(load))
    mix = np.linalg.solve(np.array([at_w.real, at_w.imag]), np.stack((load.real, load.imag)))
    return carriers @ mix


def settled(sim: Simulation, spectrum, last, frequency: float):
    """The DFT continued past step N as the constant `last` the run settled to: the
    geometric tail sum_{t >= N} z^t = z^N / (1 - z) (cuwave/sensitivity.py:349-359)."""
    z = np.exp(-2j * np.pi * frequency * sim.dt)
    return spectrum + last * complex(z**sim.N / (1.0 - z))


def spectra(sim: Simulation, acc, last, frequency: float, steps: int):
    """One frequency's settled spectrum from its (4, num) accumulator (the DFT over
    all steps, re and im, then over the last `steps`), and the share of it the
    last `steps` add beyond the settled constant (cuwave/sensitivity.py:362-388)."""
    whole = acc[0].to(torch.complex128) + 1j * acc[1].to(torch.float64)
    tail = acc[2].to(torch.complex128) + 1j * acc[3].to(torch.float64)
    z = np.exp(-2j * np.pi * frequency * sim.dt)
    tail = tail - last * complex(np.sum(z ** np.arange(sim.N - steps, sim.N)))
    norm = float(torch.linalg.vector_norm(whole))
    share = float(torch.linalg.vector_norm(tail)) / norm if norm > 0.0 else 0.0
    return settled(sim, whole, last, frequency), share


def spectral_sensitivity(sim: Simulation, source: Source, indicator, sensors, objective,
                         frequencies, region=None):
    """Cost and gradients as `sensitivity`, for a cost that reads the traces only
    through their DFT at `frequencies` (cuwave/sensitivity.py:896-1051).

    Both passes run forwards in time and keep a running DFT at the nodes the
    gradient reads (`region`, or every node, widened by the stencil's reach)
    instead of a history. dcost/dtraces is fitted by the sinusoids at the
    frequencies (a `ValueError` past FIT_LIMIT: the cost is not spectral); each
    frequency then drives one causal adjoint run by a decaying pulse with the
    same spectral load, and the gradient pairs the forward and adjoint spectra
    (the forward triplet at z^-2, z^-1, 1 against the adjoint) through the field
    gradient kernel, real and imaginary halves. It is exact once both runs have
    decayed; `info["decay"]` is the share the last period still added (a
    `RuntimeWarning` past DECAY_LIMIT), `info["fit"]` the fit residual.

    A forward step and an adjoint level each make one launch here, each followed
    by the DFT's device update: this variant trades launches for memory and is
    not a batched path.
    """
    from .boundary import require_off_held_walls

    if _elastic(sim):
        raise NotImplementedError("spectral_sensitivity is built for the scalar equations")
    _require(sim, source, sensors)
    require_full_domain(sim)
    require_off_held_walls(sim, np.asarray(as_numpy(source.position)))
    frequencies = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
    logical = np.zeros(sim.Nx_padded, dtype=bool)
    logical[tuple(slice(0, n) for n in sim.Nx)] = True
    region = logical if region is None else np.asarray(as_numpy(region), dtype=bool)
    # the gradient at a node reads its neighbours, ghosts included: so does the DFT
    square = ndi.generate_binary_structure(sim.ndim, sim.ndim)
    kept = ndi.binary_dilation(region, structure=square, iterations=sim.reach + 1) & logical
    nodes = np.stack(np.nonzero(kept)).astype(np.int32)

    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, 1)
    dev = prop.device
    at = torch.as_tensor(sim.flatten_indices(nodes), device=dev, dtype=torch.int64)
    src, signal = source_points(prop, mat, [source])
    rec = sensor_points(prop, sensors)
    record = prop.record_memory(rec)
    host_table = phase_table(sim, frequencies)
    F = len(frequencies)
    period = math.ceil(1.0 / (float(frequencies.min()) * sim.dt))
    tail = host_table.copy()
    tail[: sim.N - period] = 0.0  # the same factors over the last period alone
    table = torch.as_tensor(np.concatenate((host_table, tail), axis=1), dtype=sim.dtype,
                            device=dev)

    # forward: the field's running DFT at the kept nodes
    U = prop.levels_memory(2)
    aux = prop.aux_memory()
    slots = U.view(2, -1)
    u_hat = torch.zeros((4 * F, len(at)), dtype=sim.dtype, device=dev)
    for t in range(sim.N):
        prop.forward(U, t, t + 1, aux=aux, sources=src, signal=signal, sensors=rec,
                     record=record)
        u_hat += table[t][:, None] * slots[t % 2, at][None, :]
    u_last = slots[(sim.N - 1) % 2, at].to(torch.float64)
    um = prop.shots(record)[0]
    cost, dphi = objective(um)

    load, fit = spectral_load(sim, dphi, host_table)
    if fit > FIT_LIMIT[sim.precision]:
        raise ValueError(f"dcost/dtraces is off the sinusoids at {frequencies} by {fit:.1e}: "
                         f"the cost is not spectral there, use sensitivity")
    weights = np.asarray(sim.adjoint_weights(sensors), dtype=np.float64)
    lin = sim.flatten_indices(sensors)
    adj = sensor_points(prop, sensors, weight=sim.excitation_weights(mat, lin))
    acc = {k: prop.accumulator() for k in ("mass", "stiff")}
    zero_u = prop.levels_memory(2)
    dummy = prop.accumulator()
    decay = 0.0
    shape = tuple(sim.Nx_padded)
    for f, frequency in enumerate(frequencies):
        # the causal adjoint run driven at this frequency: level n is its step N - 1 - n
        pulse = spectral_pulse(sim, frequency, load[f]) / weights
        adjoint = prop.lane_table(torch.as_tensor(pulse, dtype=sim.dtype)[None])
        P = prop.levels_memory(2)
        ring = P.view(2, -1)
        eta = prop.eta_memory()  # the layer memory starts from rest each run
        cols = [2 * f, 2 * f + 1, 2 * F + 2 * f, 2 * F + 2 * f + 1]
        lam_hat = torch.zeros((4, len(at)), dtype=sim.dtype, device=dev)
        for t in range(sim.N):
            n = sim.N - 1 - t
            prop.backward(P, zero_u, n, n + 1, eta=eta, g_mass=dummy, combine=(0.0, 0.0),
                          sensors=adj, adjoint=adjoint)
            lam_hat += table[t, cols][:, None] * ring[n % 2, at][None, :]
        lam_last = ring[0, at].to(torch.float64)
        spectrum, share = spectra(sim, u_hat[cols], u_last, frequency, period)
        adjoint_hat, adjoint_share = spectra(sim, lam_hat, lam_last, frequency, period)
        decay = max(decay, share, adjoint_share)
        # lambda^t pairs with (u^(t - 2), u^(t - 1), u^t): the triplet is z^-2, z^-1, 1
        shift = complex(np.exp(-2j * np.pi * frequency * sim.dt))
        triplet = (spectrum * shift**2, spectrum * shift, spectrum)
        # Re(a b) = Re a Re b - Im a Im b: the bilinear form in two real halves
        for part, sign in ((torch.real, 1.0), (torch.imag, -1.0)):
            fields = [torch.zeros(shape, dtype=sim.dtype, device=dev) for _ in range(4)]
            for field, value in zip(fields, (*triplet, sign * adjoint_hat)):
                field.view(-1)[at] = part(value).to(sim.dtype)
            prop.field_gradient(*fields, acc["mass"], acc["stiff"])

    grads = _finalize(sim, acc, dev)
    if sim.pml is not None:
        grads = _zero_outside(grads, pml_nodes(sim), dev)
    grads = _zero_outside(grads, ~region, dev)
    if decay > DECAY_LIMIT:
        warnings.warn(f"the last period still adds more than {DECAY_LIMIT:.0e} of the "
                      f"spectrum: the gradient misses the tail the run cut off "
                      f"(info['decay'] says how much). Lengthen the run or damp the faces.",
                      RuntimeWarning, stacklevel=2)
    return cost, grads, um, {"decay": decay, "fit": fit}
