"""The learned-FWI pieces that run without a GPU: the network's shape, the features,
the models.

The network's forward and backward run on the GPU's matrix cores only
(tests/test_gpu_unet.py checks them against the oracle network); here the
product's parameters go through the oracle network (`oracle/unet.py`, torch's
CPU library) to pin the shapes and the parameter count of the 3-level
32/64/128 U-Net, and the product refuses to run on the CPU.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cuwave_amd import learned, unet
from oracle import unet as O


def test_the_unet_has_the_three_level_shape():
    net = unet.UNet()
    count = sum(p.numel() for p in net.parameters())
    assert 440_000 < count < 500_000
    params = {k: v.detach() for k, v in net.named_parameters()}
    x = torch.randn(2, 24, 36, 2)
    assert O.unet(params, x, net.cin_k).shape == (2, 24, 36, 1)
    assert net.cin_k == 16


def test_the_unet_runs_on_the_gpu_only():
    with pytest.raises(RuntimeError, match="GPU"):
        unet.UNet()(torch.randn(1, 8, 8, 2))


def test_the_oracle_network_reads_the_weight_layouts_as_stated():
    # a 3x3 kernel with one non-zero tap moves the image by that tap's offset
    cin = 16
    w = torch.zeros(1, 9 * cin)
    tap, ci = 3 * 0 + 2, 5  # dh = -1, dw = +1
    w[0, tap * cin + ci] = 1.0
    x = torch.randn(1, 6, 7, cin)
    y = O.conv3x3(x, w, None, False)[0, :, :, 0]
    np.testing.assert_allclose(y[1:, :-1].numpy(), x[0, :-1, 1:, ci].numpy())
    # the up-convolution: column (2 a + b) cout + co feeds output pixel (2i + a, 2j + b)
    wu = torch.zeros(3, 4 * 2)
    wu[1, (2 * 1 + 0) * 2 + 1] = 1.0
    xu = torch.randn(1, 2, 2, 3)
    yu = O.upconv(xu, wu, torch.zeros(2))
    np.testing.assert_allclose(yu[0, 1::2, 0::2, 1].numpy(), xu[0, :, :, 1].numpy())


def test_features_are_the_normalised_gradient_and_illumination():
    g = torch.tensor([[1.0, -4.0], [2.0, 0.5]], dtype=torch.float64)
    il = torch.tensor([[1.0, 3.0], [0.0, 2.0]], dtype=torch.float64)
    n = learned.normalised_gradient(g, il, floor=0.5)
    np.testing.assert_allclose(n.numpy(), g.numpy() / (il.numpy() + 0.5 * 3.0))
    f = learned.features(g, il, floor=0.5)
    assert f.shape == (2, 2, 2)
    # the gradient over its 99th percentile of magnitude, clipped to +-3
    q = np.quantile(np.abs(n.numpy()).ravel(), 0.99)
    np.testing.assert_allclose(f[0].numpy(), np.clip(n.numpy() / q, -3, 3), rtol=1e-6)
    # the illumination on a log scale: 1 at its maximum, 0 at the floor
    assert float(f[1].max()) == pytest.approx(1.0) and float(f[1].min()) >= 0.0


def test_the_models_are_seeded_and_in_range():
    a = learned.layered_model(3, (40, 50))
    b = learned.layered_model(3, (40, 50))
    c = learned.layered_model(4, (40, 50))
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.min() >= 1500.0 and a.max() <= 2500.0
