"""Learned FWI end to end, through the HIP solver and the matrix-core network.

Training samples come from the solver (traces on the true model, the batched
misfit gradient and the illumination map at the start model); the U-Net is
trained with the library's kernels. On held-out models the learned run (the
network's step first, then the normalised gradient) is timed against the
plain normalised-gradient run to the plain run's final misfit: the gain the
network brings is that it gets there faster, on most models and at the median.
The preconditioned run accepts only steps that lower the exact misfit, so its
misfit never rises.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

pytestmark = pytest.mark.gpu

from cuwave_amd import boundary, learned, scalar, utils, wave  # noqa: E402
from cuwave_amd.signals import ricker  # noqa: E402

ROI, LAYER, DX = 96, 16, 1e-3


def _setup(steps=600, shots=6):
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
    cols = np.round(np.linspace(8, ROI - 8, shots)).astype(int) + width
    sources = [wave.Source(np.array([[width], [c]]), wavelet) for c in cols]
    coords = np.stack([np.full(ROI // 2, width * DX),
                       (np.arange(0, ROI, 2) + width) * DX], axis=1)
    return sim, indicator, sources, utils.Sensors(sim, coords)


def test_the_trained_network_reaches_the_plain_misfit_faster():
    sim, indicator, sources, sensors = _setup()
    print("case learned fwi 130x130, 6 shots, 48 models", flush=True)

    def pair(seed):
        true = learned.layered_model(seed, (ROI, ROI))
        return true, gaussian_filter(true, ROI / 32)

    samples = [learned.training_sample(sim, sources, sensors, indicator(t), indicator(s))
               for t, s in map(pair, range(48))]
    x, y = samples[0]
    assert x.shape == (2, *sim.Nx_padded) and y.shape == (1, *sim.Nx_padded)
    import torch

    torch.manual_seed(0)
    net = learned.UNet().cuda()
    losses = learned.train(net, samples, epochs=120, lr=1e-3, batch=16)
    assert np.isfinite(losses).all() and losses[-1] < 0.7 * losses[0]
    held = {1000 + k: pair(1000 + k) for k in range(4)}
    results = learned.evaluate(sim, sources, sensors, indicator, held, net, iterations=8,
                               step=0.2)
    ratios = [r["ratio"] if r["ratio"] is not None else np.inf for r in results]
    print("time-to-misfit ratios", ratios, flush=True)
    for r in results:
        costs = [c for _, c in r["learned"]["history"]]
        assert all(b <= a for a, b in zip(costs, costs[1:]))  # accepted steps only
        assert r["plain"]["history"][0][1] == pytest.approx(r["learned"]["history"][0][1])
    assert sum(x < 1.0 for x in ratios) >= 3
    assert float(np.median(ratios)) < 1.0
