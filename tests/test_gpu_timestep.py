"""stable_timestep: the power iteration through the step kernels (cuwave/wave.py:56-116).

- On a uniform medium the largest mode is the grid's Nyquist mode, whose
  symbol is the closed form `stable_dt` evaluates (for the kernels' own cell
  row): the measured bound agrees with it within a percent or so (the finite
  box's highest mode sits just under the infinite grid's; sixty iterations
  of the power method approach it from below).
- With a fast inclusion the fast phase sets the bound: within a few percent of
  `stable_dt` at the fast speed, far under the slow one's.
- Elastic, as the reference's own TimestepTest states
  (tests/elastic_test.py:862-889): a uniform medium allows more than the
  cheap bound at orders 2 and 4; a void of 1e-4 density costs order 4
  timestep (0.3 < ratio < 1) and not order 2.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from cuwave_amd import elastic, scalar, wave  # noqa: E402
from cuwave_amd.stencils import flux_row, stable_dt  # noqa: E402


def _nyquist_dt(dx, c, order):
    """The closed form for the kernels' own cell row (at radius 1 the kernels hold the
    weight at exactly one, where `stable_dt` uses the fitted row)."""
    row = flux_row(order, "optimized")
    lam = 4.0 * abs(float(row @ (-1.0) ** np.arange(4)))
    return 2.0 / (c * float(np.sqrt(lam * sum(1.0 / d**2 for d in dx))))


def test_a_uniform_medium_gives_the_closed_form_bound():
    for order in (2, 4, 8):
        dx = (1e-3, 1e-3)
        dt = 0.8 * stable_dt(dx, 1500.0, order)
        sim = scalar.ScalarWave((96, 96), dx, 1, dt, precision="float64", space_order=order,
                                wavespeed=1500.0, density=1.0)
        print(f"case uniform order {order}", flush=True)
        found = wave.stable_timestep(sim, np.ones(sim.Nx_padded))
        ratio = found / (0.95 * _nyquist_dt(dx, 1500.0, order))
        assert 0.97 < ratio < 1.01, (order, ratio)


def test_a_fast_inclusion_sets_the_bound():
    dx = (1e-3, 1e-3)
    c1, c2 = 1500.0, 3000.0
    sim = scalar.AcousticWave((96, 96), dx, 1, 0.5 * stable_dt(dx, c2, 8), precision="float64",
                              space_order=8, rho1=1.0, rho2=1.0, kappa1=c1**2, kappa2=c2**2)
    a0, a1 = np.meshgrid(*(np.arange(n) for n in sim.Nx_padded), indexing="ij")
    disk = ((a0 - 48) ** 2 + (a1 - 48) ** 2 <= 15**2).astype(np.float64)
    found = wave.stable_timestep(sim, disk)
    fast, slow = 0.95 * _nyquist_dt(dx, c2, 8), 0.95 * _nyquist_dt(dx, c1, 8)
    assert 0.9 < found / fast < 1.01
    assert found < 0.6 * slow


def _elastic(res, order):
    cp_, cs, rho = 2.0, 1.0, 1.3
    dx = (1.0 / (res - 3),) * 2
    return elastic.ElasticWave((res, res), dx, 1, 0.8 * stable_dt(dx, cp_, order),
                               precision="float64", space_order=order, stencil="optimized",
                               density=rho, wavespeed_p=cp_, wavespeed_s=cs)


def test_elastic_bounds_behave_as_the_reference_states():
    for order in (2, 4):
        sim = _elastic(64, order)
        assert wave.stable_timestep(sim, np.ones(sim.Nx_padded)) / sim.dt > 1.0
    ratios = {}
    for order in (2, 4):
        sim = _elastic(96, order)
        axes = np.meshgrid(*[(np.arange(n) - 1) * h for n, h in zip(sim.Nx_padded, sim.dx)],
                           indexing="ij")
        void = (axes[0] - 0.5) ** 2 + (axes[1] - 0.5) ** 2 < 0.1**2
        indicator = np.where(void, 1e-4, 1.0)
        ratios[order] = wave.stable_timestep(sim, indicator) / sim.dt
    assert ratios[2] > 1.0
    assert 0.3 < ratios[4] < 1.0
