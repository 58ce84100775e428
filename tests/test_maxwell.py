"""The out-of-plane Maxwell reductions on the host (cuwave/maxwell.py:72-160).

What the reference's own ReductionTest states (tests/maxwell_test.py:202-272),
where it needs no GPU: the light speed is the faster phase's, three
dimensions and a missing permittivity are refused, and under the change of
variables rho = mu, kappa = 1/eps, `ElectricWave`'s coefficient fields are
`AcousticWave`'s to the bit. `MagneticWave` takes the design in the
stiffness, and its nodal Jacobian matches a central difference of 1/eps.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from cuwave_amd.maxwell import ElectricWave, MagneticWave
from cuwave_amd.scalar import AcousticWave
from cuwave_amd.stencils import stable_dt

INDEX = 3.48
EPS1, EPS2, MU = 1.0, INDEX**2, 1.0


def _sim(cls, Nx=(32, 32), **kw):
    dx = tuple(1.0 / (n - 3) for n in Nx)
    return cls(Nx, dx, 10, 0.7 * stable_dt(dx, 1.0 / math.sqrt(EPS1 * MU), 2),
               precision="float64", permittivity1=EPS1, permittivity2=EPS2,
               permeability=MU, **kw)


def test_the_light_speed_is_the_faster_phase():
    assert _sim(ElectricWave).light_speed == pytest.approx(1.0 / math.sqrt(EPS1 * MU))


def test_three_dimensions_and_a_missing_permittivity_are_refused():
    with pytest.raises(ValueError):
        ElectricWave((16, 16, 16), (0.1,) * 3, 10, 0.01, permittivity1=EPS1,
                     permittivity2=EPS2)
    with pytest.raises(ValueError):
        ElectricWave((32, 32), (0.1, 0.1), 10, 0.01, permittivity1=1.0)
    with pytest.raises(ValueError):
        MagneticWave((32, 32), (0.1, 0.1), 10, 0.01, permittivity1=1.0, permittivity2=-2.0)


def test_the_electric_wave_is_the_acoustic_wave_under_the_change_of_variables():
    em = _sim(ElectricWave)
    ac = AcousticWave(em.Nx, em.dx, em.N, em.dt, precision="float64",
                      space_order=em.space_order, rho1=MU, rho2=MU, kappa1=1.0 / EPS1,
                      kappa2=1.0 / EPS2)
    gamma = np.random.default_rng(1).random(em.Nx_padded)
    for a, b in zip(em.parametrization(gamma), ac.parametrization(gamma)):
        np.testing.assert_array_equal(np.broadcast_to(a, gamma.shape),
                                      np.broadcast_to(b, gamma.shape))
    assert em.step_factors() == ac.step_factors()
    assert em.source_factor() == ac.source_factor()
    jm, js = em.parametrization_jacobian()
    assert jm == EPS2 - EPS1 and js == 0.0


def test_the_magnetic_wave_takes_the_design_in_the_stiffness():
    sim = _sim(MagneticWave)
    gamma = np.random.default_rng(2).random(sim.Nx_padded)
    stiff, minv = sim.parametrization(gamma)
    np.testing.assert_allclose(stiff, 1.0 / (EPS1 + gamma * (EPS2 - EPS1)), rtol=1e-15)
    assert np.all(minv == 1.0 / MU)
    jm, js = sim.parametrization_jacobian(gamma)
    h = 1e-6
    fd = (sim.parametrization(gamma + h)[0] - sim.parametrization(gamma - h)[0]) / (2 * h)
    assert jm == 0.0
    np.testing.assert_allclose(js, fd, rtol=1e-7)
