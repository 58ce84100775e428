"""Write the golden input/output vectors of the reference's FD path.

Run in the build container only (it reads the reference checkout, which never
travels to a GPU box):

    python tests/golden/make_golden.py [/path/to/reference]

The reference has no GPU here, so its own CPU backend runs (the unchanged kernel
sources compiled by the host C++ compiler behind its `cpu.hpp`, IEEE arithmetic,
`cuwave/cpu.py:22`). Each case stores its inputs (simulation parameters, the
indicator, source and sensors) beside the outputs, so the tests rebuild nothing
from a recipe. What is pinned here and nowhere else: the composed outputs of
`simulate`, `sensitivity`, `reconstruction_sensitivity`,
`superposition_sensitivity` and `misfit_gradient` as the reference computes them
(`cuwave/wave.py:1071`, `cuwave/sensitivity.py:384, 500, 753`,
`cuwave/utils.py:410`), and the exact optimized stencil rows
(`cuwave/stencils.py:40-58`).
"""

from __future__ import annotations

import math
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REFERENCE = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
CACHE = tempfile.mkdtemp(prefix="cuwave-golden-")  # the reference's CPU build cache

os.environ["CUWAVE_DEVICE"] = "cpu"
os.environ["XDG_CACHE_HOME"] = CACHE
sys.path.insert(0, str(REFERENCE))

from cuwave import stencils  # noqa: E402
from cuwave.boundary import Dirichlet, Neumann, Pml  # noqa: E402
from cuwave.scalar import AcousticWave, ScalarWave  # noqa: E402
from cuwave.sensitivity import (  # noqa: E402
    l2_misfit,
    reconstruction_nodes,
    reconstruction_sensitivity,
    sensitivity,
    superposition_sensitivity,
)
from cuwave.signals import ricker  # noqa: E402
from cuwave.utils import Sensors, misfit_gradient, point_source  # noqa: E402
from cuwave.wave import Source, simulate, stable_dt  # noqa: E402

CODES = {Neumann: 0, Dirichlet: 1}


def sim_params(sim) -> dict:
    """The constructor arguments of a pressure simulation, as arrays."""
    p = {
        "kind": np.array(type(sim).__name__),
        "Nx": np.array(sim.Nx, dtype=np.int64),
        "dx": np.array(sim.dx, dtype=np.float64),
        "N": np.array(sim.N),
        "dt": np.array(sim.dt, dtype=np.float64),
        "threads": np.array(sim.threads),
        "precision": np.array(sim.precision),
        "space_order": np.array(sim.space_order),
        "stencil": np.array(sim.stencil),
        "boundary": np.array([[CODES[c] for c in pair] for pair in sim.boundary]),
    }
    if isinstance(sim, ScalarWave):
        p["wavespeed"] = np.array(sim.wavespeed)
        p["density"] = np.array(sim.density)
    else:
        p["materials"] = np.array([sim.rho1, sim.rho2, sim.kappa1, sim.kappa2])
    if sim.pml is not None:
        faces = range(2 * sim.ndim) if sim.pml.faces is None else sim.pml.faces
        p["pml"] = np.array(
            [sim.pml.width, sim.pml.wavespeed, sim.pml.reflection, sim.pml.frequency]
        )
        p["pml_faces"] = np.array(list(faces), dtype=np.int64)
    return p


def save(name: str, **arrays) -> None:
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **arrays)
    print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")


def inputs(sim, source, indicator, sensors) -> dict:
    return {
        **sim_params(sim),
        "indicator": np.asarray(indicator),
        "src_pos": np.asarray(source.position),
        "src_sig": np.asarray(source.signal),
        "sensors": np.asarray(sensors),
    }


# ------------------------------------------------------------------ stencils
def stencil_case() -> None:
    rows = {}
    for R in range(1, 5):
        for kind in ("taylor", "optimized"):
            rows[f"cell_{kind}_{R}"] = stencils.cell_coefficients(R, kind)
            rows[f"stag_{kind}_{R}"] = stencils.staggered_weights(R, kind)
    for order in (2, 4, 6, 8):
        for kind in ("taylor", "optimized"):
            rows[f"stable_dt_{kind}_{order}"] = np.array(
                stable_dt((1e-3, 2e-3), 1500.0, order, kind)
            )
    save("stencils", **rows)


# ---------------------------------------------------------- problem builders
def matched(order=2, kind="scalar", boundary=None, faces=None, N=160, ndim=2,
            precision="float64"):
    """A sim with a six-node pml behind `faces`, its region 26 nodes wide."""
    res, width = 26, 6
    Nx = (res + 2 * width, res + 2 * width - 3, res + 2 * width - 1)[:ndim]
    dx = (1.0 / (res - 3),) * ndim
    threads = {2: (4, 32), 3: (2, 4, 32)}[ndim]
    speed = math.sqrt(5.0) if kind == "acoustic" else 1.0
    common = dict(precision=precision, space_order=order, boundary=boundary,
                  pml=Pml(width, speed, faces=faces))
    dt = 0.85 * stable_dt(dx, speed, order)
    if kind == "scalar":
        return ScalarWave(Nx, dx, N, dt, threads, wavespeed=1.0, density=1.3, **common)
    return AcousticWave(Nx, dx, N, dt, threads, rho1=1.0, rho2=3.0, kappa1=1.0,
                        kappa2=5.0, **common)


def matched_case(sim, seed=0):
    """Two sources, sensors deep in the layer, on its front and inside."""
    t = np.arange(sim.N) * sim.dt
    signal = ricker(t, 1.0, 6.0) / np.prod(sim.dx)
    pos = np.array([[9, 12], [10, 25], [20, 14]][: sim.ndim], dtype=np.int32)
    columns = np.stack((signal, -0.6 * signal), axis=1).astype(sim.dtype)
    source = Source(pos, np.ascontiguousarray(columns))
    rows = [[8, 20, 29, 4], [7, 15, 28, 3], [9, 16, 7, 3]]
    sensors = np.ascontiguousarray(np.array(rows[: sim.ndim]), dtype=np.int32)
    base = 0.2 if isinstance(sim, AcousticWave) else 1.0
    rng = np.random.default_rng(seed)
    indicator = (base + 0.3 * rng.random(sim.Nx_padded)).astype(sim.dtype)
    return source, sensors, indicator


def closed(order=2, N=120, res=30, precision="float64", boundary=None, kind="scalar"):
    Nx, dx = (res, res - 2), (1.0 / (res - 3),) * 2
    if kind == "scalar":
        return ScalarWave(Nx, dx, N, 0.85 * stable_dt(dx, 1.0, order), (4, 32),
                          precision=precision, space_order=order, wavespeed=1.0,
                          density=1.3, boundary=boundary)
    speed = math.sqrt(6.87e8 / 2643.0)
    return AcousticWave(Nx, dx, N, 0.85 * stable_dt(dx, speed, order), (4, 32),
                        precision=precision, space_order=order, rho1=1.204,
                        rho2=2643.0, kappa1=1.419e5, kappa2=6.87e8,
                        boundary=boundary)


def closed_case(sim, base=1.0, row=1):
    """Source on wall node `row` (2 past a held wall); sensors on a wall, a corner and inside."""
    t = np.linspace(0, (sim.N - 1) * sim.dt, sim.N)
    signal = ricker(t, 1.0, 5.0) / np.prod(sim.dx)
    source = Source(np.array([[row], [sim.Nx[1] // 2]], dtype=np.int32),
                    np.ascontiguousarray(signal[:, None], dtype=sim.dtype))
    sensors = np.array([[1, sim.Nx[0] - 2, sim.Nx[0] // 3],
                        [5, sim.Nx[1] - 2, sim.Nx[1] // 3]], dtype=np.int32)
    x, y = np.meshgrid(*[(np.arange(n) - 1) * d for n, d in
                         zip(sim.Nx_padded, sim.dx)], indexing="ij")
    indicator = np.ascontiguousarray(
        base + 0.25 * np.sin(2 * math.pi * x) * np.cos(3 * math.pi * y),
        dtype=sim.dtype)
    return source, sensors, indicator


# ------------------------------------------------------------------- forward
def forward_cases() -> None:
    cases = {
        "fwd_closed_o2_f64": (closed(2), 1.0),
        "fwd_closed_o8_f32": (closed(8, precision="float32"), 1.0),
        "fwd_dirichlet_o4_f64": (closed(4, boundary=((Dirichlet, Neumann),
                                                     (Neumann, Dirichlet))), 1.0),
        "fwd_acoustic_o6_f32": (closed(6, precision="float32", kind="acoustic"), 0.5),
    }
    for name, (sim, base) in cases.items():
        row = 2 if "dirichlet" in name else 1
        source, sensors, indicator = closed_case(sim, base, row)
        field, um = simulate(sim, source, indicator, sensors)
        save(name, **inputs(sim, source, indicator, sensors), field=field, traces=um)
    pml_cases = {
        "fwd_pml_o2_f64": matched(2),
        "fwd_pml_o8_dirichlet_f32": matched(8, boundary=Dirichlet,
                                            precision="float32"),
        "fwd_pml_o4_mixed_f64": matched(4, boundary=((Dirichlet, Neumann),
                                                     (Neumann, Dirichlet)),
                                        faces=(0, 3)),
        "fwd_pml_acoustic_o8_f32": matched(8, kind="acoustic", precision="float32"),
        "fwd_pml_3d_o4_f32": matched(4, boundary=Dirichlet, ndim=3, N=100,
                                     precision="float32"),
        "fwd_pml_3d_o8_f64": matched(8, ndim=3, N=80),
    }
    for name, sim in pml_cases.items():
        source, sensors, indicator = matched_case(sim)
        field, um = simulate(sim, source, indicator, sensors)
        save(name, **inputs(sim, source, indicator, sensors), field=field, traces=um)


def disk_forward() -> None:
    """The plain 2D forward: 256^2 region, 20-cell layer, order 8, 2000 steps."""
    roi, pml, dx, c1, c2 = 256, 21, 1e-3, 1500.0, 2000.0
    n = roi + 2 * (pml - 1) + 2
    dt = 0.4 * dx / c2
    sim = AcousticWave((n, n), (dx, dx), 2000, dt, (8, 32), precision="float32",
                       space_order=8, rho1=1.0, rho2=1.0, kappa1=c1**2,
                       kappa2=c2**2, pml=Pml(pml, c2, 1e-6))
    centre = np.random.default_rng(1).integers(64, 192, 2) + pml
    a0, a1 = np.meshgrid(*[np.arange(k) for k in sim.Nx_padded], indexing="ij")
    disk = (a0 - centre[0]) ** 2 + (a1 - centre[1]) ** 2 <= 30**2
    indicator = disk.astype(np.float32)
    t = np.arange(sim.N) * dt
    signal = ricker(t, 1.0, 1e5) / np.prod(sim.dx)
    source = Source(np.array([[pml], [128 + pml]], dtype=np.int32),
                    np.ascontiguousarray(signal[:, None], dtype=np.float32))
    cols = np.round(np.linspace(0, roi - 1, 16)).astype(np.int32) + pml
    sensors = np.stack([np.full(16, pml, dtype=np.int32), cols])
    field, um = simulate(sim, source, indicator, sensors)
    save("disk_298x298_o8_forward", **inputs(sim, source, indicator, sensors),
         field=field, traces=um)


# ------------------------------------------------------------------- adjoint
def adjoint_cases() -> None:
    cases = {
        "adj_pml_o2_f64": matched(2),
        "adj_pml_o8_dirichlet_f64": matched(8, boundary=Dirichlet),
        "adj_pml_o4_mixed_f64": matched(4, boundary=((Dirichlet, Neumann),
                                                     (Neumann, Dirichlet)),
                                        faces=(0, 3)),
        "adj_pml_acoustic_o8_f32": matched(8, kind="acoustic", precision="float32"),
        "adj_pml_acoustic_o8_f64": matched(8, kind="acoustic"),
        "adj_pml_3d_o4_f64": matched(4, boundary=Dirichlet, ndim=3, N=100),
    }
    for name, sim in cases.items():
        source, sensors, indicator = matched_case(sim)
        observed = 0.7 * simulate(sim, source, indicator, sensors)[1]
        objective = l2_misfit(observed)
        cost, grads, um, _ = sensitivity(sim, source, indicator, sensors, objective)
        rcost, rgrads, rum, info = reconstruction_sensitivity(
            sim, source, indicator, sensors, objective)
        strip, valid = reconstruction_nodes(sim, indicator)
        out = dict(observed=observed, cost=np.array(cost), g_mass=grads["mass"],
                   g_stiff=grads["stiff"], traces=um, r_cost=np.array(rcost),
                   r_g_mass=rgrads["mass"], r_g_stiff=rgrads["stiff"],
                   r_strip=np.array(info["strip"]), r_drift=np.array(info["drift"]),
                   valid=valid, strip_nodes=strip)
        if sim.ndim == 2:
            ccost, cgrads, _, cinfo = reconstruction_sensitivity(
                sim, source, indicator, sensors, objective, checkpoints=2)
            out.update(c_cost=np.array(ccost), c_g_mass=cgrads["mass"],
                       c_g_stiff=cgrads["stiff"], c_segment=np.array(cinfo["segment"]),
                       c_sweeps=np.array(cinfo["sweeps"]))
        if sim.precision == "float32":
            bcost, bgrads, _, _ = reconstruction_sensitivity(
                sim, source, indicator, sensors, objective, strip_dtype="bfloat16")
            out.update(b_g_mass=bgrads["mass"], b_g_stiff=bgrads["stiff"])
        save(name, **inputs(sim, source, indicator, sensors), **out)

    # lossless closed boxes: all three variants
    lossless = {
        "adj_closed_o2_f64": closed(2),
        "adj_closed_o4_dirichlet_f64": closed(4, boundary=((Dirichlet, Neumann),
                                                           (Neumann, Dirichlet))),
        "adj_closed_o8_f32": closed(8, precision="float32"),
        "adj_closed_acoustic_o4_f64": closed(4, kind="acoustic"),
    }
    for name, sim in lossless.items():
        source, sensors, indicator = closed_case(
            sim, 0.5 if isinstance(sim, AcousticWave) else 1.0,
            2 if "dirichlet" in name else 1)
        objective = l2_misfit(np.zeros((sim.N, sensors.shape[1]), dtype=sim.dtype))
        cost, grads, um, _ = sensitivity(sim, source, indicator, sensors, objective)
        rcost, rgrads, _, rinfo = reconstruction_sensitivity(
            sim, source, indicator, sensors, objective)
        scale = 1e3 if sim.precision == "float32" else 1.0
        scost, sgrads, _, sinfo = superposition_sensitivity(
            sim, source, indicator, sensors, objective, scale=scale)
        save(name, **inputs(sim, source, indicator, sensors), cost=np.array(cost),
             g_mass=grads["mass"], g_stiff=grads["stiff"], traces=um,
             r_g_mass=rgrads["mass"], r_g_stiff=rgrads["stiff"],
             r_drift=np.array(rinfo["drift"]), s_cost=np.array(scost),
             s_g_mass=sgrads["mass"], s_g_stiff=sgrads["stiff"],
             s_scale=np.array(scale), s_cancellation=np.array(sinfo["cancellation"]))


def misfit_case() -> None:
    """Three shots through `Sensors`, coordinates off the nodes."""
    for precision in ("float64", "float32"):
        sim = matched(4, kind="acoustic", precision=precision, N=140)
        rng = np.random.default_rng(5)
        indicator = (0.2 + 0.3 * rng.random(sim.Nx_padded)).astype(sim.dtype)
        truth = (0.35 + 0.0 * indicator).astype(sim.dtype)
        t = np.arange(sim.N) * sim.dt
        wavelet = ricker(t, 1.0, 6.0)
        L = [(n - 3) * d for n, d in zip(sim.Nx, sim.dx)]
        coords = [(0.31 * L[0], f * L[1]) for f in (0.3, 0.5, 0.72)]
        sources = [point_source(sim, c, wavelet) for c in coords]
        rec = np.stack([np.full(7, 0.7 * L[0]), np.linspace(0.2, 0.8, 7) * L[1]], 1)
        sensors = Sensors(sim, rec + 0.013)
        observed = [sensors.traces(simulate(sim, s, truth, sensors.nodes)[1])
                    for s in sources]
        cost, gradient = misfit_gradient(sim, sources, indicator, sensors, observed)
        rcost, rgradient = misfit_gradient(sim, sources, indicator, sensors, observed,
                                           adjoint=reconstruction_sensitivity)
        tag = "f64" if precision == "float64" else "f32"
        save(f"misfit_acoustic_o4_{tag}", **sim_params(sim), indicator=indicator,
             truth=truth, src_pos=np.stack([s.position for s in sources]),
             src_sig=np.stack([s.signal for s in sources]),
             coords=np.array(coords), rec_coords=rec + 0.013,
             sensor_nodes=sensors.nodes, sensor_weights=sensors.weights,
             observed=np.stack(observed), cost=np.array(cost), gradient=gradient,
             r_cost=np.array(rcost), r_gradient=rgradient)


if __name__ == "__main__":
    which = set(sys.argv[2:]) or {"stencils", "forward", "disk", "adjoint",
                                  "misfit"}
    if "stencils" in which:
        stencil_case()
    if "forward" in which:
        forward_cases()
    if "disk" in which:
        disk_forward()
    if "adjoint" in which:
        adjoint_cases()
    if "misfit" in which:
        misfit_case()
