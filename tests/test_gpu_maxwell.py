"""The out-of-plane Maxwell reductions through the HIP kernels.

As the reference's own contracts state them (tests/maxwell_test.py:1-12,
160-337): `ElectricWave` is the acoustic solver under rho = mu, kappa = 1/eps,
unknown for unknown, so the two fields agree bit for bit; a plane wave travels
at the speed the index sets; a quarter-wave slab transmits what Fabry-Perot
says it does, and the error shrinks as the grid is refined; and the gradients
of both polarisations (the design in the inertia for `ElectricWave`, in the
stiffness through a nodal Jacobian for `MagneticWave`) match central
differences of the cost, for the stored-history and the reverse-march adjoint
and the batched misfit gradient. An L2 misfit stands in for the reference's
spectral `intensity` objective.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from helpers import host  # noqa: E402

from cuwave_amd import sensitivity as S  # noqa: E402
from cuwave_amd import utils  # noqa: E402
from cuwave_amd.maxwell import ElectricWave, MagneticWave  # noqa: E402
from cuwave_amd.scalar import AcousticWave  # noqa: E402
from cuwave_amd.signals import ricker  # noqa: E402
from cuwave_amd.stencils import stable_dt  # noqa: E402
from cuwave_amd.wave import grid_coords, simulate  # noqa: E402

INDEX = 3.48
EPS1, EPS2, MU = 1.0, INDEX**2, 1.0


def _sim(cls, Nx, N=260, space_order=2):
    dx = tuple(1.0 / (n - 3) for n in Nx)
    return cls(Nx, dx, N, 0.7 * stable_dt(dx, 1.0 / math.sqrt(EPS1 * MU), space_order),
               precision="float64", space_order=space_order, permittivity1=EPS1,
               permittivity2=EPS2, permeability=MU)


def test_the_electric_wave_is_the_acoustic_solver_bit_for_bit():
    em = _sim(ElectricWave, (64, 64))
    ac = AcousticWave(em.Nx, em.dx, em.N, em.dt, precision="float64",
                      space_order=em.space_order, rho1=MU, rho2=MU, kappa1=1.0 / EPS1,
                      kappa2=1.0 / EPS2)
    print("case electric vs acoustic 64x64", flush=True)
    gamma = np.random.default_rng(1).random(em.Nx_padded)
    signal = ricker(np.arange(em.N) * em.dt, 1.0, 8.0)
    one = simulate(em, utils.point_source(em, [(0.4, 0.5)], signal), gamma)[0]
    other = simulate(ac, utils.point_source(ac, [(0.4, 0.5)], signal), gamma)[0]
    assert float(np.max(np.abs(host(one) - host(other)))) == 0.0


def test_a_plane_wave_travels_at_the_speed_the_index_sets():
    for eps in (EPS1, EPS2):
        index = math.sqrt(eps * MU)
        dx = (1.0 / 1997,)
        dt = 0.4 * stable_dt(dx, 1.0 / index, 2)
        sim = ElectricWave((2000,), dx, int(0.6 * index / dt), dt, precision="float64",
                           permittivity1=eps, permittivity2=eps, permeability=MU)
        print(f"case plane wave eps {eps}", flush=True)
        source = utils.point_source(sim, [(0.2,)], ricker(np.arange(sim.N) * dt, 1.0, 30.0))
        probes = np.array([[round(0.5 / dx[0]) + 1, round(0.7 / dx[0]) + 1]], dtype=np.int32)
        record = host(simulate(sim, source, np.zeros(sim.Nx_padded), sensors=probes)[1])
        peaks = np.argmax(np.abs(record), axis=0)
        speed = 0.2 / ((peaks[1] - peaks[0]) * dt)
        assert speed * index == pytest.approx(1.0, abs=0.01)


def _slab_transmission(points_per_wavelength, frequencies):
    """Measured and Fabry-Perot power transmission of a quarter-wave slab in 1D."""
    thickness = 0.25 / INDEX
    length, x_src, x_slab, x_probe = 20.0, 8.0, 10.0, 12.0
    dx = (1.0 / (INDEX * points_per_wavelength),)
    Nx = (int(length / dx[0]) + 3,)
    dt = 0.95 * stable_dt(dx, 1.0, 2)
    N = math.ceil(19.0 / dt)  # both wall echoes reach the probe at t = 20
    sim = ElectricWave(Nx, dx, N, dt, precision="float64", permittivity1=EPS1,
                       permittivity2=EPS2, permeability=MU)
    x = grid_coords(Nx, dx)[0]
    slab = (np.abs(x - (x_slab + 0.5 * thickness)) <= 0.5 * thickness).astype(np.float64)
    t = np.arange(N) * dt
    source = utils.point_source(sim, [(x_src,)], ricker(t, 1.0, 1.2))
    probe = np.array([[round(x_probe / dx[0]) + 1]], dtype=np.int32)
    reference = host(simulate(sim, source, np.zeros(sim.Nx_padded), sensors=probe)[1])[:, 0]
    through = host(simulate(sim, source, slab, sensors=probe)[1])[:, 0]
    basis = np.exp(-2j * np.pi * np.outer(frequencies, t))
    measured = np.abs((basis @ through) / (basis @ reference)) ** 2
    beta = 2.0 * np.pi * INDEX * float(slab.sum()) * dx[0] * np.asarray(frequencies)
    contrast = (EPS1 - EPS2) ** 2 / (4.0 * EPS1 * EPS2)
    return measured, 1.0 / (1.0 + contrast * np.sin(beta) ** 2)


def test_a_quarter_wave_slab_transmits_what_fabry_perot_says():
    frequencies = (0.5, 1.0, 1.5, 2.0)
    print("case fabry-perot slab", flush=True)
    coarse, analytic = _slab_transmission(20, frequencies)
    np.testing.assert_allclose(coarse / analytic, 1.0, atol=0.05)
    fine, _ = _slab_transmission(80, frequencies)

    def worst(measured):
        return float(np.max(np.abs(measured - analytic) / analytic))

    assert worst(fine) < 0.25 * worst(coarse)  # first order: the interface limits it


def _problem(sim, seed=0):
    rng = np.random.default_rng(seed)
    source = utils.point_source(sim, [(0.3, 0.5)], ricker(np.arange(sim.N) * sim.dt, 1.0, 8.0))
    sensors = np.array([[30, 30, 32], [24, 26, 25]], dtype=np.int32)
    return source, sensors, rng.random(sim.Nx_padded)


def _fd(cost_of, indicator, nodes, h=1e-7):
    out = []
    for node in nodes:
        plus, minus = indicator.copy(), indicator.copy()
        plus[node] += h
        minus[node] -= h
        out.append((cost_of(plus) - cost_of(minus)) / (2 * h))
    return out


@pytest.mark.parametrize("cls", [ElectricWave, MagneticWave])
def test_both_polarisations_match_finite_differences(cls):
    sim = _sim(cls, (48, 48))
    source, sensors, indicator = _problem(sim)
    print(f"case {cls.__name__} gradient 48x48", flush=True)
    observed = np.zeros((sim.N, sensors.shape[1]))
    rng = np.random.default_rng(5)
    nodes = [(int(rng.integers(12, 34)), int(rng.integers(12, 34))) for _ in range(4)]

    def cost_of(field):
        return S.sensitivity(sim, source, field, sensors, S.l2_misfit(observed))[0]

    expect = _fd(cost_of, indicator, nodes)
    for adjoint in (S.sensitivity, S.reconstruction_sensitivity):
        _, g, _, _ = adjoint(sim, source, indicator, sensors, S.l2_misfit(observed))
        jm, js = sim.parametrization_jacobian(indicator)
        chained = jm * host(g["mass"]) + js * host(g["stiff"])
        for node, ref in zip(nodes, expect):
            assert chained[node] / ref == pytest.approx(1.0, abs=1e-4), (adjoint.__name__, node)
    # the batched path: a nodal Jacobian chained after the sweep
    sens = utils.Sensors(sim, (sensors.T - 1) * np.array(sim.dx))
    _, gradient = utils.misfit_gradient(sim, [source], indicator, sens, [observed])
    gradient = host(gradient)
    nodes_fd = _fd(lambda f: utils.misfit(sim, [source], f, sens, [observed]), indicator, nodes)
    for node, ref in zip(nodes, nodes_fd):
        assert gradient[node] / ref == pytest.approx(1.0, abs=1e-4), node
