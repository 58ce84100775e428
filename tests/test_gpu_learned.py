"""Learned FWI end to end on a small grid, through the HIP solver.

Training samples come from the solver: traces on the true model, the
batched misfit gradient and the illumination map at the start model. The
preconditioned iteration may only accept steps that lower the exact misfit,
so its misfit never rises. The plain iteration (Adam on the indicator) is the
baseline it is timed against.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

pytestmark = pytest.mark.gpu

from cuwave_amd import boundary, learned, scalar, utils, wave  # noqa: E402
from cuwave_amd.signals import ricker  # noqa: E402

ROI, LAYER, DX = 40, 8, 1e-3


def _setup(steps=160, shots=4):
    width = LAYER + 1
    n = ROI + 2 * LAYER + 2
    dt = 0.4 * DX / 2500.0
    sim = scalar.AcousticWave((n, n), (DX, DX), steps, dt, precision="float32", space_order=8,
                              rho1=1.0, rho2=1.0, kappa1=1500.0**2, kappa2=2500.0**2,
                              pml=boundary.Pml(width, 2500.0, 1e-6))

    def indicator(v):
        full = np.pad(v, ((width, n - width - ROI),
                          (width, sim.Nx_padded[1] - width - ROI)), mode="edge")
        return sim.indicator_of_velocity(full)

    wavelet = (ricker(np.arange(steps) * dt, 1.0, 1e5) / (DX * DX))[:, None]
    cols = np.round(np.linspace(4, ROI - 4, shots)).astype(int) + width
    sources = [wave.Source(np.array([[width], [c]]), wavelet) for c in cols]
    coords = np.stack([np.full(ROI // 2, width * DX),
                       (np.arange(0, ROI, 2) + width) * DX], axis=1)
    return sim, indicator, sources, utils.Sensors(sim, coords)


def test_learned_fwi_trains_and_its_steps_lower_the_misfit():
    sim, indicator, sources, sensors = _setup()
    print("case learned fwi 58x64 four shots", flush=True)
    samples = []
    for seed in (0, 1):
        true = learned.layered_model(seed, (ROI, ROI))
        start = gaussian_filter(true, 6)
        samples.append(learned.training_sample(sim, sources, sensors, indicator(true),
                                               indicator(start)))
    x, y = samples[0]
    assert x.shape == (2, *sim.Nx_padded) and y.shape == (1, *sim.Nx_padded)
    net = learned.UNet()
    losses = learned.train(net, samples, epochs=20)
    assert np.isfinite(losses).all() and losses[-1] < losses[0]

    true = learned.layered_model(1000, (ROI, ROI))
    start = indicator(gaussian_filter(true, 6))
    observed = utils.measure(sim, sources, indicator(true), sensors)
    _, plain = learned.fwi(sim, sources, sensors, observed, start, 3, step=0.02)
    _, pre = learned.fwi(sim, sources, sensors, observed, start, 3, step=0.05, net=net)
    costs = [c for _, c in pre]
    assert np.isfinite([c for _, c in plain]).all() and np.isfinite(costs).all()
    assert all(b <= a for a, b in zip(costs, costs[1:]))  # accepted steps only
    assert plain[0][1] == pytest.approx(pre[0][1])  # the same start
