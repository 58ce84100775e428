"""`cuwave_amd.autograd.traces`: the gradient of any torch loss of the traces.

The incoming gradient of the traces is the adjoint source. These tests pin it:
- against the batched misfit gradient for the L2 misfit;
- against the oracle's stored-history adjoint for a loss that is not
  quadratic, whose derivative no built-in objective supplies;
- against central finite differences in float64, for a scalar indicator and
  for the two-field (kappa, rho) indicator of `BulkDensityWave`.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import golden, host, oracle_sim, product_sim, rel_max

pytestmark = pytest.mark.gpu

from cuwave_amd import autograd, scalar, utils, wave  # noqa: E402

F32_GRAD = 1e-4


def _shots(z):
    return [wave.Source(p, s) for p, s in zip(z["src_pos"], z["src_sig"])]


def test_an_l2_loss_gives_the_misfit_gradient():
    z = golden("misfit_acoustic_o4_f32")
    sim = product_sim(z)
    print("case autograd l2 misfit_acoustic_o4_f32", flush=True)
    sensors = utils.Sensors(sim, z["rec_coords"])
    sources = _shots(z)
    ind = torch.tensor(z["indicator"], requires_grad=True)
    tr = autograd.traces(ind, sim, sources, sensors)
    obs = torch.as_tensor(np.stack(z["observed"]), device=tr.device, dtype=tr.dtype)
    loss = 0.5 * torch.sum((tr - obs) ** 2)
    loss.backward()
    cost, gradient = utils.misfit_gradient(sim, sources, z["indicator"], sensors,
                                           list(z["observed"]))
    assert abs(loss.item() - cost) <= 1e-5 * abs(cost)
    assert rel_max(ind.grad, host(gradient)) < F32_GRAD
    assert rel_max(ind.grad, z["r_gradient"]) < F32_GRAD
    # a batch of two shots a launch gives the same
    ind2 = torch.tensor(z["indicator"], requires_grad=True)
    tr2 = autograd.traces(ind2, sim, sources, sensors, batch=2)
    (0.5 * torch.sum((tr2 - obs) ** 2)).backward()
    assert rel_max(ind2.grad, host(ind.grad)) < F32_GRAD


def test_a_loss_of_any_form_matches_the_oracle():
    from oracle import driver

    z = golden("misfit_acoustic_o4_f64")
    sim, osim = product_sim(z), oracle_sim(z)
    print("case autograd quartic misfit_acoustic_o4_f64", flush=True)
    sensors = utils.Sensors(sim, z["rec_coords"])
    osensors = driver.Sensors(osim, z["rec_coords"])
    sources = _shots(z)
    wt = np.linspace(0.5, 1.5, sim.N)[:, None]

    def quartic(t):  # a loss no built-in objective gives
        return 0.25 * (wt * t**4).sum()

    ind = torch.tensor(z["indicator"], requires_grad=True)
    tr = autograd.traces(ind, sim, sources, sensors)
    loss = 0.25 * torch.sum(torch.as_tensor(wt, device=tr.device) * tr**4)
    loss.backward()
    d_mass, d_stiff = osim.parametrization_jacobian()
    expect, ecost = np.zeros(sim.Nx_padded), 0.0
    for s in sources:
        def objective(t):
            return quartic(t), wt * t**3
        c, g, _, _ = driver.sensitivity(osim, s.position, s.signal, z["indicator"],
                                        osensors.nodes, osensors.objective(objective))
        ecost += c
        expect += d_mass * g["mass"] + d_stiff * g["stiff"]
    assert abs(loss.item() - ecost) <= 1e-10 * abs(ecost)
    assert rel_max(ind.grad, expect) < 1e-10


def _fd_check(sim, indicator, sources, sensors, loss_of, nodes, h=1e-6):
    ind = torch.tensor(indicator, requires_grad=True)
    loss_of(autograd.traces(ind, sim, sources, sensors)).backward()
    grad = ind.grad.numpy()

    def cost(x):
        with torch.no_grad():
            return float(loss_of(autograd.traces(torch.tensor(x), sim, sources, sensors)))

    x = np.array(indicator)
    for node in nodes:
        base = x[node]
        x[node] = base + h
        plus = cost(x)
        x[node] = base - h
        minus = cost(x)
        x[node] = base
        fd = (plus - minus) / (2 * h)
        assert abs(grad[node] - fd) <= 1e-6 * abs(fd), (node, grad[node], fd)


def test_the_gradient_matches_finite_differences_in_float64():
    z = golden("adj_pml_o8_dirichlet_f64")
    sim = product_sim(z)
    sim.N = 90
    print("case autograd fd adj_pml_o8_dirichlet_f64", flush=True)
    coords = np.array([[0.6, 0.9], [0.9, 0.6], [1.1, 1.2]])
    sensors = utils.Sensors(sim, coords)
    sources = [wave.Source(z["src_pos"], z["src_sig"][: sim.N])]

    def loss_of(tr):
        return torch.sum(tr**2) + 0.1 * torch.sum(tr**3)

    _fd_check(sim, z["indicator"], sources, sensors, loss_of,
              [(15, 15), (20, 12), (12, 22)])


def test_a_two_field_indicator_gets_both_gradients():
    # BulkDensityWave: the indicator stacks (kappa, rho) and the jacobian is a field
    base = golden("adj_pml_o8_dirichlet_f64")
    ref = product_sim(base)
    sim = scalar.BulkDensityWave(ref.Nx, ref.dx, 70, ref.dt, precision="float64",
                                 space_order=ref.space_order, boundary=ref.boundary)
    print("case autograd fd bulk-density f64", flush=True)
    rng = np.random.default_rng(5)
    kappa = 1.0 + 0.2 * rng.random(sim.Nx_padded)
    rho = 1.0 + 0.2 * rng.random(sim.Nx_padded)
    indicator = np.stack([kappa, rho])
    sensors = utils.Sensors(sim, np.array([[0.6, 0.9], [0.9, 0.6]]))
    sources = [wave.Source(base["src_pos"], base["src_sig"][: sim.N])]

    def loss_of(tr):
        return torch.sum(tr**2)

    _fd_check(sim, indicator, sources, sensors, loss_of,
              [(0, 15, 15), (1, 15, 15), (0, 20, 12), (1, 12, 22)])
