"""Damping through the HIP kernels, against the oracle and the mathematics.

A damping field d makes the step (2 u1 - u0 (1 - beta) + minv L u1) / (1 + beta),
beta = minv d dt / 2 (cuwave/kernels/scalar.cu:455-459), the adjoint likewise
(scalar_sensitivity.cu:444-515); the reverse march of a reconstruction steps
the reversible damped nodes with the damping negated and records the rest
(cuwave/sensitivity.py:139-195, 592-598).

- the damped forward is bit for bit the oracle's: a quadratic sponge (float64,
  Neumann), the same in float32 with Dirichlet faces, a uniform damping with a
  CPML, 1D and 3D;
- the stored-history adjoint matches the oracle's within the stated tolerance;
- the reconstruction, with a patch too lossy to run back (its nodes recorded),
  with and without checkpoints, matches the oracle's gradient, strip and valid
  set, and the product's own gradient matches central differences in float64;
- a damped closed domain is where `spectral_sensitivity` decays: it matches the
  stored-history gradient to 1e-5 (tests/sensitivity_test.py:1146-1171);
- superposition refuses a damping field.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from helpers import host, rel_max  # noqa: E402

from oracle import driver as D  # noqa: E402

from cuwave_amd import boundary, scalar, utils, wave  # noqa: E402
from cuwave_amd import sensitivity as S  # noqa: E402
from cuwave_amd.signals import ricker  # noqa: E402


def _pair(Nx=(34, 31), N=90, precision="float64", order=4, bnd=None, pml=None, beta=0.05,
          sponge=True, patch=None, seed=4):
    base = scalar.AcousticWave(Nx, (1.0,) * len(Nx), N, 0.3, precision=precision,
                               space_order=order, rho1=1.0, rho2=2.5, kappa1=1.0, kappa2=4.0,
                               boundary=bnd, pml=pml)
    rng = np.random.default_rng(seed)
    ind = (0.2 + 0.6 * rng.random(base.Nx_padded)).astype(base.np_dtype)
    minv = base.inverse_inertia(ind)
    if sponge:
        damping = boundary.sponge(base, ind, 6, beta)
    else:
        inner = np.zeros(base.Nx_padded, dtype=bool)
        inner[tuple(slice(1, n - 1) for n in base.Nx)] = True
        damping = np.where(inner, 2.0 * beta / (base.dt * minv), 0.0).astype(base.np_dtype)
    if patch is not None:
        damping[patch] = (4.0 / (base.dt * minv[patch])).astype(base.np_dtype)  # beta 2
    sim = scalar.AcousticWave(Nx, base.dx, N, base.dt, precision=precision, space_order=order,
                              rho1=1.0, rho2=2.5, kappa1=1.0, kappa2=4.0, boundary=bnd,
                              pml=pml, damping=damping)
    faces = tuple(tuple(D.DIRICHLET if c is boundary.Dirichlet else D.NEUMANN for c in pair)
                  for pair in sim.boundary)
    opml = None if pml is None else D.Pml(pml.width, pml.wavespeed, pml.reflection)
    osim = D.Sim("AcousticWave", Nx, base.dx, N, base.dt, precision, order,
                 boundary=faces, materials=(1.0, 2.5, 1.0, 4.0), damping=damping, pml=opml)
    mid = [n // 2 for n in Nx]
    src = np.array([[m] for m in mid])
    src[0, 0] = 12
    t = np.arange(N) * base.dt
    sig = (np.exp(-((t - 3.0) ** 2))[:, None]).astype(base.np_dtype)
    sensors = np.array([[5 + d, n // 2 + 3, n - 5] for d, n in enumerate(Nx)])
    return sim, osim, ind, src, sig, sensors


CASES = {
    "sponge f64": dict(),
    "sponge f32 dirichlet": dict(precision="float32",
                                 bnd=((boundary.Dirichlet, boundary.Neumann),
                                      (boundary.Neumann, boundary.Dirichlet))),
    "uniform with pml": dict(sponge=False, beta=0.02, pml=boundary.Pml(6, 2.0)),
    "1d": dict(Nx=(60,), order=2),
    "3d f32": dict(Nx=(22, 21, 20), precision="float32", N=50),
}


@pytest.mark.parametrize("name", list(CASES))
def test_the_damped_forward_is_the_oracles_bit_for_bit(name):
    sim, osim, ind, src, sig, sensors = _pair(**CASES[name])
    print(f"case damped forward {name}", flush=True)
    field, traces = wave.simulate(sim, wave.Source(src, sig), ind, sensors)
    ofield, otraces = D.simulate(osim, src, sig, ind, sensors)
    np.testing.assert_array_equal(host(field), ofield)
    np.testing.assert_array_equal(host(traces), otraces)


@pytest.mark.parametrize("name", ["sponge f64", "sponge f32 dirichlet", "uniform with pml"])
def test_the_damped_adjoint_matches_the_oracle(name):
    sim, osim, ind, src, sig, sensors = _pair(**CASES[name])
    print(f"case damped adjoint {name}", flush=True)
    tol = 1e-10 if sim.precision == "float64" else 1e-4
    obs = D.simulate(osim, src, sig, ind, sensors)[1] * 0.7
    cost, g, _, _ = S.sensitivity(sim, wave.Source(src, sig), ind, sensors, S.l2_misfit(obs))
    ocost, og, _, _ = D.sensitivity(osim, src, sig, ind, sensors, D.l2_misfit(obs))
    assert abs(cost - ocost) <= tol * abs(ocost)
    assert rel_max(g["mass"], og["mass"]) < tol
    assert rel_max(g["stiff"], og["stiff"]) < tol


@pytest.mark.parametrize("checkpoints", [None, 2])
def test_a_damped_reconstruction_matches_the_oracle_and_finite_differences(checkpoints):
    patch = (slice(10, 17), slice(10, 17))
    sim, osim, ind, src, sig, sensors = _pair(patch=patch)
    print(f"case damped reconstruction checkpoints={checkpoints}", flush=True)
    obs = D.simulate(osim, src, sig, ind, sensors)[1] * 0.7
    objective = S.l2_misfit(obs)
    cost, g, _, info = S.reconstruction_sensitivity(sim, wave.Source(src, sig), ind, sensors,
                                                    objective, checkpoints=checkpoints)
    ocost, og, _, oinfo = D.reconstruction_sensitivity(osim, src, sig, ind, sensors,
                                                       D.l2_misfit(obs), checkpoints=checkpoints)
    assert info["strip"] == oinfo["strip"] > 0
    assert info["segment"] == oinfo["segment"]
    assert info["drift"] < 1e-10
    assert abs(cost - ocost) <= 1e-10 * abs(ocost)
    assert rel_max(g["mass"], og["mass"]) < 1e-9
    assert rel_max(g["stiff"], og["stiff"]) < 1e-9
    jm, js = sim.parametrization_jacobian()
    gradient = jm * host(g["mass"]) + js * host(g["stiff"])
    for node in [(21, 9), (17, 22), (24, 15)]:
        h = 1e-6

        def cost_at(delta):
            f = ind.copy()
            f[node] += delta
            return objective(wave.simulate(sim, wave.Source(src, sig), f, sensors)[1])[0]

        fd = (cost_at(h) - cost_at(-h)) / (2 * h)
        assert abs(gradient[node] - fd) <= 1e-5 * abs(fd), (node, gradient[node], fd)


def test_spectral_matches_sensitivity_once_a_damped_domain_decays():
    res = 30
    dx = (1.0 / (res - 3),) * 2
    base = scalar.ScalarWave((res, res - 2), dx, 900, 0.85 * wave.stable_dt(dx, 1.0, 2),
                             precision="float64", wavespeed=1.0, density=1.3)
    x, y = wave.grid_coords(base.Nx, base.dx)
    ind = 1.0 + 0.25 * np.sin(2 * np.pi * x) * np.cos(3 * np.pi * y)
    minv = base.inverse_inertia(ind)
    inner = np.zeros(base.Nx_padded, dtype=bool)
    inner[tuple(slice(1, n - 1) for n in base.Nx)] = True
    damping = np.where(inner, 2.0 * 0.02 / (base.dt * minv), 0.0)
    sim = scalar.ScalarWave(base.Nx, dx, 900, base.dt, precision="float64", wavespeed=1.0,
                            density=1.3, damping=damping)
    print("case damped spectral 30x28", flush=True)
    t = np.arange(sim.N) * sim.dt
    source = wave.Source(np.array([[1], [sim.Nx[1] // 2]]),
                         (ricker(t, 1.0, 5.0) / np.prod(sim.dx))[:, None])
    sensors = np.array([[1, sim.Nx[0] - 2, sim.Nx[0] // 3], [5, sim.Nx[1] - 2, sim.Nx[1] // 3]])
    freqs = (5.0, 7.0)
    objective = utils.intensity(sim, freqs)
    c0, g0, _, _ = S.sensitivity(sim, source, ind, sensors, objective)
    c1, g1, _, info = S.spectral_sensitivity(sim, source, ind, sensors, objective, freqs)
    jm, js = sim.parametrization_jacobian()
    a = jm * host(g1["mass"]) + js * host(g1["stiff"])
    b = jm * host(g0["mass"]) + js * host(g0["stiff"])
    rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"relative difference {rel:.2e}, decay {info['decay']:.2e}", flush=True)
    assert c0 == c1
    assert info["decay"] < 1e-5
    assert rel < 1e-5


def test_superposition_refuses_a_damping_field():
    sim, _, ind, src, sig, sensors = _pair()
    with pytest.raises(NotImplementedError, match="lossless"):
        S.superposition_sensitivity(sim, wave.Source(src, sig), ind, sensors,
                                    S.l2_misfit(np.zeros((sim.N, 3))))
