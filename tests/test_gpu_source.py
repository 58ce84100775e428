"""`source_sensitivity`: d(cost)/d(source.signal) from the adjoint alone
(cuwave/sensitivity.py:1054-1111).

The cost is quadratic in the signal, so a central difference of it is exact
but for rounding: the gradient at a few (step, point) entries matches it in
float64, with a CPML and without. The traces it returns are `simulate`'s.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from helpers import host  # noqa: E402

from cuwave_amd import boundary, scalar, wave  # noqa: E402
from cuwave_amd import sensitivity as S  # noqa: E402
from cuwave_amd.signals import ricker  # noqa: E402


def _case(pml):
    dx = (1e-3, 1e-3)
    n = 60
    dt = 0.5 * wave.stable_dt(dx, 2500.0, 4)
    sim = scalar.AcousticWave((n, n + 4), dx, 300, dt, precision="float64", space_order=4,
                              rho1=1.0, rho2=1.3, kappa1=1500.0**2, kappa2=2500.0**2,
                              pml=boundary.Pml(10, 2500.0, 1e-6) if pml else None)
    rng = np.random.default_rng(3)
    indicator = rng.random(sim.Nx_padded)
    sig = ricker(np.arange(sim.N) * dt, 1.0, 1e5)
    position = np.array([[25, 26], [30, 31]])
    signal = np.stack([sig, 0.5 * np.roll(sig, 7)], axis=1) / (dx[0] * dx[1])
    sensors = np.array([[14, 15, 16, 40], [20, 34, 45, 30]])
    observed = rng.standard_normal((sim.N, sensors.shape[1])) * 1e-3
    return sim, wave.Source(position, signal), indicator, sensors, S.l2_misfit(observed)


@pytest.mark.parametrize("pml", [True, False])
def test_the_source_gradient_matches_central_differences(pml):
    sim, source, indicator, sensors, objective = _case(pml)
    print(f"case source sensitivity pml={pml}", flush=True)
    cost, gradient, traces, info = S.source_sensitivity(sim, source, indicator, sensors,
                                                        objective)
    assert info == {} and gradient.shape == (sim.N, 2)
    _, expect = wave.simulate(sim, source, indicator, sensors)
    np.testing.assert_array_equal(host(traces), host(expect))
    gradient = host(gradient)
    h = 1e-3 * float(np.abs(source.signal).max())
    for n, j in ((5, 0), (40, 1), (120, 0), (250, 1), (299, 0)):
        def cost_at(delta):
            signal = source.signal.copy()
            signal[n, j] += delta
            return _cost(sim, wave.Source(source.position, signal), indicator, sensors,
                         objective)

        fd = (cost_at(h) - cost_at(-h)) / (2 * h)
        scale = max(abs(fd), 1e-6 * float(np.abs(gradient).max()))
        assert abs(gradient[n, j] - fd) <= 1e-6 * scale + 1e-9 * float(np.abs(gradient).max()), \
            (n, j, gradient[n, j], fd)


def _cost(sim, source, indicator, sensors, objective):
    _, traces = wave.simulate(sim, source, indicator, sensors)
    return objective(traces)[0]
