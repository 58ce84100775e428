"""Write the golden input/output vectors of the reference's elastic FD path.

Run in the build container only (it reads the reference checkout, which never
travels to a GPU box):

    python tests/golden/make_golden_elastic.py [/path/to/reference]

As `make_golden.py`, it runs the reference's own CPU backend (its unchanged
kernel sources compiled by the host C++ compiler, IEEE arithmetic,
`cuwave/cpu.py:22`), its build cache in a temporary directory. Each case stores
its inputs beside the outputs. Pinned here: the staggered SBP closure tables
(`cuwave/stencils.py:75-213`), the composed outputs of `simulate`,
`sensitivity` and `reconstruction_sensitivity` for `ElasticWave`
(`cuwave/elastic.py`, `cuwave/kernels/elastic.cu`,
`cuwave/kernels/elastic_sensitivity.cu`), and `misfit_gradient` over shots.
"""

from __future__ import annotations

import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REFERENCE = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
CACHE = tempfile.mkdtemp(prefix="cuwave-golden-")

os.environ["CUWAVE_DEVICE"] = "cpu"
os.environ["XDG_CACHE_HOME"] = CACHE
sys.path.insert(0, str(REFERENCE))

from cuwave import stencils  # noqa: E402
from cuwave.boundary import Clamped, Pml, Traction  # noqa: E402
from cuwave.elastic import ElasticWave  # noqa: E402
from cuwave.sensitivity import (  # noqa: E402
    l2_misfit,
    reconstruction_nodes,
    reconstruction_sensitivity,
    sensitivity,
)
from cuwave.signals import ricker  # noqa: E402
from cuwave.utils import misfit_gradient  # noqa: E402
from cuwave.wave import Source, simulate, stable_dt  # noqa: E402

CODES = {Traction: 0, Clamped: 1}
CP, CS, RHO = 2.0, 1.0, 1.3


def save(name: str, **arrays) -> None:
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **arrays)
    print("wrote", path.name, sorted(arrays))


def sim_params(sim) -> dict:
    p = {
        "kind": np.array("ElasticWave"),
        "Nx": np.array(sim.Nx, dtype=np.int64),
        "dx": np.array(sim.dx, dtype=np.float64),
        "N": np.array(sim.N),
        "dt": np.array(sim.dt, dtype=np.float64),
        "precision": np.array(sim.precision),
        "space_order": np.array(sim.space_order),
        "stencil": np.array(sim.stencil),
        "boundary": np.array([[CODES[c] for c in pair] for pair in sim.boundary]),
        "elastic": np.array([sim.density, sim.wavespeed_p, sim.wavespeed_s]),
        "plane": np.array(sim.plane),
    }
    if sim.pml is not None:
        faces = range(2 * sim.ndim) if sim.pml.faces is None else sim.pml.faces
        p["pml"] = np.array([sim.pml.width, sim.pml.wavespeed, sim.pml.reflection,
                             sim.pml.frequency])
        p["pml_faces"] = np.array(list(faces), dtype=np.int64)
    return p


def make(Nx, N, order, precision="float64", boundary=None, pml=None, stencil="taylor",
         plane="strain"):
    dx = tuple(1.0 / (n - 3) for n in Nx)
    threads = {1: (128,), 2: (8, 8), 3: (4, 4, 8)}[len(Nx)]
    return ElasticWave(Nx, dx, N, 0.8 * stable_dt(dx, CP, order), threads, precision=precision,
                       space_order=order, stencil=stencil, density=RHO, wavespeed_p=CP,
                       wavespeed_s=CS, boundary=boundary, pml=pml, plane=plane)


def problem(sim, seed=0, row=None):
    """One source on component `ndim - 1`, receivers on every component, a rough model."""
    rng = np.random.default_rng(seed)
    t = np.arange(sim.N) * sim.dt
    signal = ricker(t, 1.0, 6.0)[:, None] / np.prod(sim.dx)
    mid = [n // 2 for n in sim.Nx]
    if row is not None:
        mid[0] = row
    comp = [[sim.ncomp - 1]] if sim.ncomp > 1 else []
    pos = np.array(comp + [[m - 2] for m in mid], dtype=np.int32)
    source = Source(pos, np.ascontiguousarray(signal, dtype=sim.dtype))
    comps = [list(range(sim.ncomp)) + [0]] if sim.ncomp > 1 else []
    sensors = np.array(comps + [[m + 2] * sim.ncomp + [m + 1] for m in mid], dtype=np.int32)
    indicator = np.asarray(0.6 + 0.3 * rng.random(sim.Nx_padded), dtype=sim.dtype)
    return source, sensors, indicator


def inputs(sim, source, indicator, sensors) -> dict:
    return {**sim_params(sim), "indicator": np.asarray(indicator),
            "src_pos": np.asarray(source.position), "src_sig": np.asarray(source.signal),
            "sensors": np.asarray(sensors)}


def stencil_case() -> None:
    out = {}
    for order in (2, 4, 6, 8):
        for stencil in ("taylor", "optimized"):
            tables, counts = stencils._closure_tables(order, stencil)
            key = f"o{order}_{stencil}"
            out[f"{key}_node_rows"] = tables["NODE_ROWS"]
            out[f"{key}_half_rows"] = tables["HALF_ROWS"]
            out[f"{key}_counts"] = np.array(counts)
            out[f"{key}_stag"] = stencils.staggered_weights(order // 2, stencil)
        node, half = stencils.closure_norms(order, 40)
        out[f"o{order}_node_norm"] = node
        out[f"o{order}_half_norm"] = half
    save("el_stencils", **out)


FREE_TOP = ((Traction, Clamped), (Clamped, Clamped))
LINED = dict(faces=(1, 2, 3))  # every face but the free top


def forward_cases() -> None:
    cases = {
        "el_fwd_2d_o4_pml_f32": (make((44, 46), 120, 4, "float32", FREE_TOP,
                                      Pml(7, CP, frequency=2.0, **LINED)), 3),
        "el_fwd_2d_o2_traction_f64": (make((30, 33), 90, 2), None),
        "el_fwd_2d_o8_mixed_f32": (make((44, 40), 100, 8, "float32",
                                        ((Clamped, Traction), (Traction, Clamped)),
                                        stencil="optimized"), None),
        "el_fwd_2d_o6_stress_f64": (make((42, 41), 80, 6, boundary=FREE_TOP,
                                         plane="stress"), None),
        "el_fwd_3d_o2_pml_f64": (make((20, 19, 21), 60, 2, boundary=(
            (Traction, Clamped), (Clamped, Clamped), (Clamped, Clamped)),
            pml=Pml(5, CP, frequency=2.0, faces=(1, 2, 3, 4, 5))), 4),
        "el_fwd_3d_o4_f32": (make((24, 23, 22), 60, 4, "float32"), None),
        "el_fwd_1d_o4_f64": (make((60,), 150, 4, boundary=((Traction, Clamped),)), None),
    }
    for name, (sim, row) in cases.items():
        source, sensors, indicator = problem(sim, row=row)
        field, traces = simulate(sim, source, indicator, sensors)
        save(name, **inputs(sim, source, indicator, sensors), field=np.asarray(field),
             traces=np.asarray(traces))


def adjoint_cases() -> None:
    cases = {
        "el_adj_2d_o4_pml_f64": make((40, 42), 90, 4, "float64", FREE_TOP,
                                     Pml(7, CP, frequency=2.0, **LINED)),
        "el_adj_2d_o4_pml_f32": make((40, 42), 90, 4, "float32", FREE_TOP,
                                     Pml(7, CP, frequency=2.0, **LINED)),
        "el_adj_2d_o2_closed_f64": make((26, 27), 70, 2),
        "el_adj_2d_o8_pml_f64": make((48, 46), 80, 8, "float64",
                                     ((Clamped, Clamped), (Clamped, Clamped)),
                                     Pml(9, CP, frequency=2.0)),
        "el_adj_3d_o2_pml_f64": make((20, 19, 21), 50, 2, "float64", (
            (Traction, Clamped), (Clamped, Clamped), (Clamped, Clamped)),
            Pml(5, CP, frequency=2.0, faces=(1, 2, 3, 4, 5))),
        "el_adj_1d_o4_f64": make((50,), 100, 4, boundary=((Traction, Clamped),)),
    }
    for name, sim in cases.items():
        row = (sim.pml.width + 2) if sim.pml is not None else None
        source, sensors, indicator = problem(sim, seed=1, row=row)
        _, observed = simulate(sim, source, indicator, sensors)
        observed = 0.8 * np.asarray(observed)
        objective = l2_misfit(observed)
        out = inputs(sim, source, indicator, sensors)
        out["observed"] = observed
        cost, g, traces, _ = sensitivity(sim, source, indicator, sensors, objective)
        out.update(s_cost=np.array(cost), s_mass=np.asarray(g["mass"]),
                   s_stiff=np.asarray(g["stiff"]), s_traces=np.asarray(traces))
        cost, g, traces, info = reconstruction_sensitivity(sim, source, indicator, sensors,
                                                           objective)
        out.update(r_cost=np.array(cost), r_mass=np.asarray(g["mass"]),
                   r_stiff=np.asarray(g["stiff"]), r_strip=np.array(info["strip"]),
                   r_drift=np.array(info["drift"]))
        if sim.pml is not None:
            cost, g, _, info = reconstruction_sensitivity(sim, source, indicator, sensors,
                                                          objective, checkpoints=2)
            out.update(c_cost=np.array(cost), c_mass=np.asarray(g["mass"]),
                       c_stiff=np.asarray(g["stiff"]), c_segment=np.array(info["segment"]),
                       c_sweeps=np.array(info["sweeps"]))
            strip, valid = reconstruction_nodes(sim, indicator)
            out.update(strip_nodes=np.asarray(strip), valid=np.asarray(valid))
        save(name, **out)


def misfit_case() -> None:
    """Three shots of one model, summed by misfit_gradient through the reconstruction."""
    for precision in ("float32", "float64"):
        sim = make((40, 42), 80, 4, precision, FREE_TOP, Pml(7, CP, frequency=2.0, **LINED))
        _, sensors, indicator = problem(sim, seed=2, row=9)
        t = np.arange(sim.N) * sim.dt
        shots, observed = [], []
        truth = np.asarray(indicator) * 1.1
        for k, col in enumerate((12, 21, 30)):
            sig = (ricker(t, 1.0, 6.0 + k)[:, None] / np.prod(sim.dx)).astype(sim.dtype)
            src = Source(np.array([[0], [9], [col]], dtype=np.int32), sig)
            shots.append(src)
            observed.append(np.asarray(simulate(sim, src, truth, sensors)[1]))

        cost, gradient = _misfit(sim, shots, indicator, sensors, observed)
        save(f"el_misfit_o4_{'f32' if precision == 'float32' else 'f64'}",
             **sim_params(sim), indicator=np.asarray(indicator),
             src_pos=np.stack([s.position for s in shots]),
             src_sig=np.stack([np.asarray(s.signal) for s in shots]),
             sensors=sensors, observed=np.stack(observed), cost=np.array(cost),
             gradient=np.asarray(gradient))


def _misfit(sim, shots, indicator, sensors, observed):
    """misfit_gradient with receivers on unknowns: the record is the traces."""
    return misfit_gradient(sim, shots, indicator, _Receivers(sensors), observed,
                           adjoint=reconstruction_sensitivity)


class _Receivers:
    """Receivers sitting on unknowns: the traces are the record (cuwave/utils.py:303-365)."""

    def __init__(self, nodes):
        self.nodes = nodes

    def objective(self, objective):
        return objective


if __name__ == "__main__":
    stencil_case()
    forward_cases()
    adjoint_cases()
    misfit_case()
