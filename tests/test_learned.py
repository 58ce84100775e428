"""The learned-FWI pieces that run without a GPU: the U-Net, its training, the features.

The network and its optimiser are torch's own (the CPU run is the library
oracle); these tests pin the shapes, the parameter count of the 3-level
32/64/128 network, that training fits a target, and the feature scaling.
"""

from __future__ import annotations

import numpy as np
import torch

from cuwave_amd import learned


def test_the_unet_has_the_three_level_shape():
    net = learned.UNet()
    count = sum(p.numel() for p in net.parameters())
    assert 440_000 < count < 500_000
    x = torch.randn(2, 2, 24, 36)
    assert net(x).shape == (2, 1, 24, 36)
    with torch.autocast("cpu", dtype=torch.bfloat16):
        assert net(x).dtype == torch.bfloat16
    net(x).square().mean().backward()
    assert all(p.grad is not None for p in net.parameters())


def test_training_fits_a_target():
    torch.manual_seed(0)
    x = torch.randn(4, 2, 16, 16)
    y = 0.5 * x[:, :1] - 0.25 * x[:, 1:]  # a linear map the network can learn
    losses = learned.train(learned.UNet(base=8), list(zip(x, y)), epochs=60, lr=3e-3)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.3 * losses[0]


def test_features_are_the_normalised_gradient_and_illumination():
    g = torch.tensor([[1.0, -4.0], [2.0, 0.5]], dtype=torch.float64)
    il = torch.tensor([[1.0, 3.0], [0.0, 2.0]], dtype=torch.float64)
    n = learned.normalised_gradient(g, il, floor=0.5)
    np.testing.assert_allclose(n.numpy(), g.numpy() / (il.numpy() + 0.5 * 3.0))
    f = learned.features(g, il, floor=0.5)
    assert f.shape == (2, 2, 2)
    assert float(f[0].abs().max()) == 1.0 and float(f[1].abs().max()) == 1.0


def test_the_models_are_seeded_and_in_range():
    a = learned.layered_model(3, (40, 50))
    b = learned.layered_model(3, (40, 50))
    c = learned.layered_model(4, (40, 50))
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.min() >= 1500.0 and a.max() <= 2500.0
