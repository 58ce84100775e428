"""`spectral_sensitivity` through the HIP kernels (cuwave/sensitivity.py:896-1051).

As the reference's own tests pin it (tests/sensitivity_test.py:672-685,
1146-1185):

- through an open domain (a six-node PML round a 26-node region, order 4,
  4800 steps so both runs decay) it matches the stored-history `sensitivity`
  of the same `intensity` cost to 1e-3 relative, with the decay it reports
  under 1e-3, and the same cost;
- a `region` keeps the gradient inside it, equal there to the whole-grid
  one, zero outside;
- a run that has not decayed warns;
- a cost that is not spectral is refused.

The field gradient kernel it pairs the spectra with is checked on its own
against the fused adjoint's densities of one level: with the forward triplet
and the adjoint level the adjoint kernel saw, the mass density by parts in
time and the stiffness density agree.
"""

from __future__ import annotations

import warnings

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from helpers import host  # noqa: E402

from cuwave_amd import boundary, scalar, utils, wave  # noqa: E402
from cuwave_amd import sensitivity as S  # noqa: E402
from cuwave_amd.signals import ricker  # noqa: E402

FREQS = (5.0, 7.0)


def _matched(N, order=4):
    res, width = 26, 6
    Nx = (res + 2 * width, res + 2 * width - 3)
    dx = (1.0 / (res - 3),) * 2
    return scalar.ScalarWave(Nx, dx, N, 0.85 * wave.stable_dt(dx, 1.0, order),
                             precision="float64", space_order=order, wavespeed=1.0,
                             density=1.3, pml=boundary.Pml(width, 1.0))


def _case(sim):
    t = np.arange(sim.N) * sim.dt
    signal = ricker(t, 1.0, 6.0) / np.prod(sim.dx)
    source = wave.Source(np.array([[9, 12], [10, 25]]),
                         np.stack((signal, -0.6 * signal), axis=1))
    sensors = np.array([[8, 20, 29, 4], [7, 15, 28, 3]])
    rng = np.random.default_rng(0)
    indicator = 1.0 + 0.3 * rng.random(sim.Nx_padded)
    return source, sensors, indicator


def _combine(sim, grads, indicator):
    jm, js = sim.parametrization_jacobian(indicator)
    return jm * host(grads["mass"]) + js * host(grads["stiff"])


def test_it_matches_sensitivity_through_an_open_domain():
    sim = _matched(4800)
    source, sensors, indicator = _case(sim)
    print("case spectral open domain 38x35 order 4, 4800 steps", flush=True)
    objective = utils.intensity(sim, FREQS)
    c0, g0, _, _ = S.sensitivity(sim, source, indicator, sensors, objective)
    c1, g1, traces, info = S.spectral_sensitivity(sim, source, indicator, sensors, objective,
                                                  FREQS)
    assert c0 == c1
    a, b = _combine(sim, g1, indicator), _combine(sim, g0, indicator)
    rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"relative difference {rel:.3e}, decay {info['decay']:.3e}, fit {info['fit']:.1e}",
          flush=True)
    assert info["decay"] < 1e-3
    assert info["fit"] < S.FIT_LIMIT["float64"]
    assert rel < 1e-3
    np.testing.assert_array_equal(host(traces), host(wave.simulate(sim, source, indicator,
                                                                   sensors)[1]))


def test_a_region_keeps_the_gradient_inside_it():
    sim = _matched(1200)
    source, sensors, indicator = _case(sim)
    print("case spectral region", flush=True)
    objective = utils.intensity(sim, FREQS)
    region = np.zeros(sim.Nx_padded, dtype=bool)
    region[12:22, 10:18] = True
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)  # the decay is beside the point here
        _, whole, _, _ = S.spectral_sensitivity(sim, source, indicator, sensors, objective, FREQS)
        _, part, _, _ = S.spectral_sensitivity(sim, source, indicator, sensors, objective, FREQS,
                                               region)
    a, b = _combine(sim, part, indicator), _combine(sim, whole, indicator)
    rel = float(np.linalg.norm(a[region] - b[region]) / np.linalg.norm(b[region]))
    assert rel < 1e-12
    assert float(np.abs(a[~region]).max()) == 0.0


def test_a_run_that_has_not_decayed_warns_and_a_non_spectral_cost_is_refused():
    sim = _matched(300)
    source, sensors, indicator = _case(sim)
    print("case spectral short run", flush=True)
    with pytest.warns(RuntimeWarning, match="last period"):
        S.spectral_sensitivity(sim, source, indicator, sensors, utils.intensity(sim, FREQS),
                               FREQS)
    observed = np.random.default_rng(1).standard_normal((sim.N, sensors.shape[1]))
    with pytest.raises(ValueError, match="not spectral"):
        S.spectral_sensitivity(sim, source, indicator, sensors, S.l2_misfit(observed), FREQS)


def test_the_field_gradient_is_the_fused_adjoints_densities():
    # one adjoint level of the stored-history sweep adds, per node,
    # -mf u^n (lam^n - 2 lam^(n+1) + lam^(n+2)) and the stiffness density of
    # (u^n, lam^(n+1)); the field kernel pairs a triplet with one adjoint field
    # instead. On the same u and lam the stiffness densities agree (the fused
    # one carries mf = 1/dt^2 on the step factors, the field one the factors
    # over dt^2), and each mass density is its closed form.
    from cuwave_amd.engine import Propagator

    for order, bnd in ((2, boundary.Neumann), (4, boundary.Dirichlet), (8, boundary.Neumann)):
        sim = scalar.AcousticWave((37, 45), (0.1, 0.1), 3, 0.3 * wave.stable_dt((0.1, 0.1), 1.0,
                                                                                  order),
                                  precision="float64", space_order=order, boundary=bnd,
                                  rho1=1.0, rho2=3.0, kappa1=1.0, kappa2=5.0)
        print(f"case field gradient order {order}", flush=True)
        rng = np.random.default_rng(order)
        mat = sim.build_materials(0.2 + rng.random(sim.Nx_padded))
        prop = Propagator(sim, mat, 1)
        dev = prop.device
        mask = np.zeros(sim.Nx_padded)
        mask[tuple(slice(1, n - 1) for n in sim.Nx)] = 1.0
        shape = tuple(sim.Nx_padded)
        u1 = torch.as_tensor(rng.standard_normal(shape) * mask, device=dev)
        l1 = torch.as_tensor(rng.standard_normal(shape) * mask, device=dev)
        for f in (u1, l1):
            S._refill(sim, f)
        # the stored-history adjoint at level n = 0 with lam^1 = lam^2 = l1 and
        # u^0 = u1: its densities are -mf u1 (lam^0 - l1) and the stiffness one
        U = torch.zeros((3, *shape, 1), dtype=torch.float64, device=dev)
        U[2, ..., 0] = u1  # level 0 at slot (0 + 2) % 3
        P = torch.zeros((2, *shape, 1), dtype=torch.float64, device=dev)
        P[1, ..., 0] = l1
        P[0, ..., 0] = l1
        acc = {k: prop.accumulator() for k in ("mass", "stiff")}
        prop.backward(P, U, 0, 1, u_offset=2, eta=prop.eta_memory(), g_mass=acc["mass"],
                      g_stiff=acc["stiff"])
        lam0 = P[0, ..., 0]
        mf = sim.np_dtype(sim.adjoint_mass_factor())
        # the field kernel on (u0, u1, u2) = (0, u1, 0) against l1 gives mass
        # -inv_dt2 l1 (-2 u1) and the same stiffness density (factors over dt^2,
        # the fused one carries mf = 1 / dt^2 on the step factors)
        g = {k: prop.accumulator() for k in ("mass", "stiff")}
        zero = torch.zeros(shape, dtype=torch.float64, device=dev)
        prop.field_gradient(zero, u1, zero, l1, g["mass"], g["stiff"])
        inner = tuple(slice(1, n - 1) for n in sim.Nx)
        a, b = host(g["stiff"])[inner], host(acc["stiff"])[inner]
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b)), order
        fused_mass = host(acc["mass"])[inner]
        expect = -mf * host(u1)[inner] * (host(lam0)[inner] - 2 * host(l1)[inner]
                                           + host(l1)[inner])
        np.testing.assert_allclose(fused_mass, expect, rtol=1e-9, atol=1e-12)
        field_mass = host(g["mass"])[inner]
        np.testing.assert_allclose(field_mass, 2.0 / sim.dt**2 * host(l1)[inner]
                                   * host(u1)[inner], rtol=1e-12, atol=1e-12)
