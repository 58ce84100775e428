"""Shared builders for the tests: simulations from the golden vectors, both for the
product (`cuwave_amd`) and for the CPU oracle (`oracle/driver.py`)."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def golden(name: str):
    return np.load(GOLDEN / f"{name}.npz")


def golden_names(prefix: str) -> list[str]:
    return sorted(p.stem for p in GOLDEN.glob(f"{prefix}*.npz"))


def product_sim(z, **override):
    """The `cuwave_amd` simulation a golden case stores."""
    from cuwave_amd import boundary, scalar

    bmap = {0: boundary.Neumann, 1: boundary.Dirichlet}
    bnd = tuple((bmap[int(a)], bmap[int(b)]) for a, b in z["boundary"])
    pml = None
    if "pml" in z:
        w, c, r, f = (float(v) for v in z["pml"])
        pml = boundary.Pml(int(w), c, r, f, tuple(int(x) for x in z["pml_faces"]))
    common = dict(precision=str(z["precision"]), space_order=int(z["space_order"]),
                  stencil=str(z["stencil"]), boundary=bnd, pml=pml)
    common.update(override)
    args = (tuple(int(n) for n in z["Nx"]), tuple(float(d) for d in z["dx"]), int(z["N"]),
            float(z["dt"]))
    if str(z["kind"]) == "ScalarWave":
        return scalar.ScalarWave(*args, wavespeed=float(z["wavespeed"]),
                                 density=float(z["density"]), **common)
    r1, r2, k1, k2 = (float(v) for v in z["materials"])
    return scalar.AcousticWave(*args, rho1=r1, rho2=r2, kappa1=k1, kappa2=k2, **common)


def elastic_sim(z, **override):
    """The `cuwave_amd` elastic simulation an elastic golden case stores (faces coded
    0 traction, 1 clamped)."""
    from cuwave_amd import boundary, elastic

    bmap = {0: boundary.Traction, 1: boundary.Clamped}
    bnd = tuple((bmap[int(a)], bmap[int(b)]) for a, b in z["boundary"])
    pml = None
    if "pml" in z:
        w, c, r, f = (float(v) for v in z["pml"])
        pml = boundary.Pml(int(w), c, r, f, tuple(int(x) for x in z["pml_faces"]))
    rho, cp, cs = (float(v) for v in z["elastic"])
    kw = dict(precision=str(z["precision"]), space_order=int(z["space_order"]),
              stencil=str(z["stencil"]), boundary=bnd, pml=pml, density=rho, wavespeed_p=cp,
              wavespeed_s=cs, plane=str(z["plane"]))
    kw.update(override)
    return elastic.ElasticWave(tuple(int(n) for n in z["Nx"]), tuple(float(d) for d in z["dx"]),
                               int(z["N"]), float(z["dt"]), **kw)


def oracle_sim(z):
    from oracle import driver

    return driver.sim_from_npz(z)


def rel_max(a, b) -> float:
    """max |a - b| / max |b|."""
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b) -> float:
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def host(a) -> np.ndarray:
    return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)


def magnetodielectric_class():
    """A DielectricWave that designs the permeability too, to reach the designed-permeability
    path (cuwave tests/maxwell_test.py:102-119): both materials interpolate."""
    from dataclasses import dataclass

    from cuwave_amd.maxwell import DielectricWave

    @dataclass
    class MagnetodielectricWave(DielectricWave):
        magnetic = True
        permeability2: float = 2.0

        def parametrization(self, indicator):
            eps = self.permittivity1 + indicator * (self.permittivity2 - self.permittivity1)
            mu = self.permeability + indicator * (self.permeability2 - self.permeability)
            return eps.astype(self.np_dtype), (1.0 / mu).astype(self.np_dtype)

        def parametrization_jacobian(self, indicator):
            contrast = self.permeability2 - self.permeability
            mu = self.permeability + indicator * contrast
            return self.permittivity2 - self.permittivity1, -contrast / mu**2

    return MagnetodielectricWave


class Receivers:
    """Receivers on unknowns for `misfit_gradient`: the record is the traces."""

    def __init__(self, nodes):
        self.nodes = nodes

    def objective(self, objective):
        return objective

    def traces(self, record):
        return record
