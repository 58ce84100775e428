"""Scalar pressure wave equations: the material hooks the kernels are driven by.

`PressureWave` holds the nodal stiffness `stiff` and inverse inertia `minv` the
step reads (cuwave/scalar.py:140-160, 739-745); `ScalarWave` scales both with one
indicator gamma, so `minv` is never formed (`derive_inertia`, cuwave/scalar.py:
875-908); `AcousticWave` interpolates inverse density and inverse bulk modulus
between two phases (cuwave/scalar.py:911-950). `BulkDensityWave` is new: its
indicator stacks independent bulk modulus and density fields, so stiff = 1/rho
and minv = kappa, the form a pressure-velocity acoustic medium takes.

The parametrizations run on the host in the simulation's precision with numpy's
scalar rules, as the reference computes them, so the material the kernels read
is the reference's to the bit.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from .wave import Simulation, as_numpy, domain_cells


def mirror_ghosts(sim: Simulation, field: np.ndarray) -> np.ndarray:
    """Copy each ghost layer from the node two in, for homogeneous Neumann (cuwave/wave.py:125-134)."""
    for d in range(sim.ndim):
        for ghost, mirror in ((0, 2), (sim.Nx[d] - 1, sim.Nx[d] - 3)):
            dst = [slice(None)] * sim.ndim
            dst[d] = ghost
            src = [slice(None)] * sim.ndim
            src[d] = mirror
            field[tuple(dst)] = field[tuple(src)]
    return field


@dataclass
class PressureWave(Simulation):
    """Scalar wave equation base: nodal `stiff` and `minv` fields."""

    derive_inertia = False  # set where m == k: minv derived from stiff
    gradient_names = ("mass", "stiff")
    accepts_damping = True
    accepts_domain = True

    def parametrization(self, indicator: np.ndarray):
        raise NotImplementedError

    def parametrization_jacobian(self, indicator=None):
        raise NotImplementedError

    def indicator_shape(self) -> tuple[int, ...]:
        return self.Nx_padded

    def build_materials(self, indicator) -> dict:
        """The material the kernels read, on the host and on the GPU, ghosts mirrored."""
        ind = np.asarray(as_numpy(indicator), dtype=self.np_dtype)
        if ind.shape != self.indicator_shape():
            raise ValueError(f"indicator of shape {ind.shape}, expected {self.indicator_shape()}")
        stiff, minv = self.parametrization(ind)
        stiff = mirror_ghosts(self, np.array(stiff, dtype=self.np_dtype))
        minv = None if minv is None else mirror_ghosts(self, np.array(minv, dtype=self.np_dtype))
        return self.device_materials(stiff, minv)

    def device_materials(self, stiff: np.ndarray, minv: np.ndarray | None) -> dict:
        device = _abi.require_device()
        mat = {
            "stiff_host": stiff,
            "minv_host": minv,
            "stiff": torch.as_tensor(np.ascontiguousarray(stiff), device=device),
            "minv": None if minv is None else torch.as_tensor(np.ascontiguousarray(minv), device=device),
        }
        if self.damping is not None:  # the field as given (cuwave/scalar.py:407-415)
            mat["damping_host"] = self.damping
            mat["damping"] = torch.as_tensor(self.damping, device=device)
        if self.masked:  # the packed cell radii the masked step reads
            mat["cells"] = torch.as_tensor(domain_cells(self), device=device)
        return mat

    def inverse_inertia(self, indicator) -> np.ndarray:
        """Nodal 1 / m for `indicator`."""
        stiff, minv = self.parametrization(np.asarray(as_numpy(indicator), dtype=self.np_dtype))
        return 1.0 / stiff if minv is None else minv

    def excitation_weights(self, mat: dict, lin: np.ndarray, damping="sim") -> np.ndarray:
        """Source weights dt^2 source_factor / inertia at `lin`, over 1 + beta where the
        step is damped, since the damped update divides by it (cuwave/scalar.py:857-872);
        `damping` overrides the simulation's field (the reverse march's)."""
        damping = self.damping if isinstance(damping, str) else damping
        field = mat["stiff_host"] if self.derive_inertia else mat["minv_host"]
        weight = field.ravel()[lin]
        if self.derive_inertia:
            weight = 1.0 / weight
        if damping is not None:
            beta = 0.5 * weight * np.asarray(damping).ravel()[lin] * self.dt
            weight = weight / (1.0 + beta)
        return (self.np_dtype(self.dt**2 * self.source_factor()) * weight).astype(self.np_dtype)

    def adjoint_weights(self, sensors) -> np.ndarray:
        """The divisor of the adjoint source: cell weights over the source factor."""
        return self.sensor_cell_weights(sensors) * self.source_factor()


@dataclass
class ScalarWave(PressureWave):
    """Constant-speed scalar wave equation; gamma scales density and stiffness alike."""

    wavespeed: float = None
    density: float = None

    derive_inertia = True

    def __post_init__(self) -> None:
        super().__post_init__()
        if self.wavespeed is None or self.density is None:
            raise ValueError("ScalarWave requires wavespeed and density")

    def parametrization(self, indicator):
        return indicator, None

    def parametrization_jacobian(self, indicator=None):
        return 1.0, 1.0

    def step_factors(self) -> list:
        return [self.np_dtype(2.0 * self.wavespeed**2 * self.dt**2 / d**2) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0 / self.density


@dataclass
class AcousticWave(PressureWave):
    """Two-phase acoustic medium: gamma interpolates 1/rho and 1/kappa between the phases."""

    rho1: float = None
    rho2: float = None
    kappa1: float = None
    kappa2: float = None

    def __post_init__(self) -> None:
        super().__post_init__()
        if None in (self.rho1, self.rho2, self.kappa1, self.kappa2):
            raise ValueError("AcousticWave requires rho1, rho2, kappa1, kappa2")

    def parametrization(self, indicator):
        rho_inv = 1 / self.rho1 + indicator * (1 / self.rho2 - 1 / self.rho1)
        kappa_inv = 1 / self.kappa1 + indicator * (1 / self.kappa2 - 1 / self.kappa1)
        return rho_inv, 1 / kappa_inv

    def parametrization_jacobian(self, indicator=None):
        return 1 / self.kappa2 - 1 / self.kappa1, 1 / self.rho2 - 1 / self.rho1

    def step_factors(self) -> list:
        return [self.np_dtype(2.0 * self.dt**2 / d**2) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0

    def indicator_of_velocity(self, velocity) -> np.ndarray:
        """gamma giving wave speed `velocity` at rho1 = rho2 = 1: 1/c^2 interpolated.

        With rho1 = rho2 = 1 the mass 1/kappa is affine in gamma, so any positive
        speed field is reached exactly: gamma = (1/c^2 - 1/c1^2) / (1/c2^2 - 1/c1^2).
        """
        if self.rho1 != 1.0 or self.rho2 != 1.0:
            raise ValueError("a velocity model needs rho1 = rho2 = 1")
        c = np.asarray(as_numpy(velocity), dtype=np.float64)
        g = (1 / c**2 - 1 / self.kappa1) / (1 / self.kappa2 - 1 / self.kappa1)
        return g.astype(self.np_dtype)


@dataclass
class BulkDensityWave(PressureWave):
    """Independent bulk modulus and density: the indicator is the (2, *Nx_padded) stack (kappa, rho).

    stiff = 1 / rho and minv = kappa, so the mass 1 / kappa and the stiffness each
    follow one field, and dJ/dkappa = -g_mass / kappa^2, dJ/drho = -g_stiff / rho^2.
    """

    def indicator_shape(self) -> tuple[int, ...]:
        return (2, *self.Nx_padded)

    def parametrization(self, indicator):
        kappa, rho = indicator[0], indicator[1]
        return 1 / rho, kappa

    def parametrization_jacobian(self, indicator):
        kappa, rho = np.asarray(as_numpy(indicator), dtype=self.np_dtype)
        return -1 / kappa**2, -1 / rho**2

    def step_factors(self) -> list:
        return [self.np_dtype(2.0 * self.dt**2 / d**2) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0
