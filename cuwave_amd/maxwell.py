"""The out-of-plane reductions of Maxwell's equations on the scalar kernels.

In 1D and 2D one field component is out of plane and the curl-curl form reduces
to a flux divergence, so both polarisations are `PressureWave`s with the two
coefficient fields assigned differently (cuwave/maxwell.py:72-160):

- `ElectricWave` carries E_z: the permittivity is its inertia and 1/mu its
  stiffness, so the design (which interpolates the permittivity) enters the
  inertia, with a constant Jacobian.
- `MagneticWave` carries H_z: mu is its inertia and 1/eps its stiffness, so the
  design enters the stiffness as 1/eps(gamma), with a nodal Jacobian.

Both step with the factors 2 dt^2 / dx^2 and an unscaled source. The 3D vector
form (`MaxwellWave`, the staggered curl-curl) is not part of this build.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .scalar import PressureWave
from .wave import as_numpy


@dataclass
class PolarizedWave(PressureWave):
    """Two-phase 2D Maxwell base: one out-of-plane component, gamma interpolating eps."""

    permittivity1: float = None  # gamma = 0
    permittivity2: float = None  # gamma = 1
    permeability: float = 1.0  # non-magnetic media: not designed

    def __post_init__(self) -> None:
        super().__post_init__()
        if None in (self.permittivity1, self.permittivity2):
            raise ValueError(f"{type(self).__name__} requires permittivity1 and permittivity2")
        if min(self.permittivity1, self.permittivity2, self.permeability) <= 0.0:
            raise ValueError(
                f"permittivity and permeability must be positive: {self.permittivity1}, "
                f"{self.permittivity2}, {self.permeability}")
        if self.ndim > 2:
            raise ValueError(f"an out-of-plane reduction is 1D or 2D, not {self.ndim}D")

    @property
    def light_speed(self) -> float:
        """The fastest speed on the grid, 1 / sqrt(min(eps) mu), which bounds dt."""
        return 1.0 / math.sqrt(min(self.permittivity1, self.permittivity2) * self.permeability)

    def permittivity(self, indicator):
        """eps1 + gamma (eps2 - eps1): linear in the design."""
        return self.permittivity1 + indicator * (self.permittivity2 - self.permittivity1)

    def _constant(self, value: float) -> np.ndarray:
        return np.full(self.Nx_padded, value, dtype=self.np_dtype)

    def step_factors(self) -> list:
        return [self.np_dtype(2.0 * self.dt**2 / d**2) for d in self.dx]

    def source_factor(self) -> float:
        return 1.0


@dataclass
class ElectricWave(PolarizedWave):
    """E_z: the permittivity is the inertia, 1/mu the stiffness."""

    def parametrization(self, indicator):
        return self._constant(1.0 / self.permeability), 1.0 / self.permittivity(indicator)

    def parametrization_jacobian(self, indicator=None):
        return self.permittivity2 - self.permittivity1, 0.0


@dataclass
class MagneticWave(PolarizedWave):
    """H_z: mu is the inertia, 1/eps the stiffness (a nodal Jacobian)."""

    def parametrization(self, indicator):
        return 1.0 / self.permittivity(indicator), self._constant(1.0 / self.permeability)

    def parametrization_jacobian(self, indicator):
        ind = np.asarray(as_numpy(indicator), dtype=self.np_dtype)
        contrast = self.permittivity2 - self.permittivity1
        return 0.0, -contrast / self.permittivity(ind) ** 2
