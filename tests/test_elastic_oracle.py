"""The CPU oracle of the elastic path, pinned before it is trusted.

- The golden vectors the reference's own CPU backend wrote
  (tests/golden/el_*.npz, by tests/golden/make_golden_elastic.py): forward
  fields and traces in 1D, 2D and 3D at orders 2-8, traction and clamped
  walls, the folded PML, plane stress; the stored-history adjoint, the
  strip-replayed reconstruction with and without checkpoints, its strip and
  valid region; and the shot-summed misfit gradient. All bit for bit.
- The 1D reduction: on a uniform medium the staggered scheme is the scalar
  stencil on the offset grid (the reference's own OneDimensionTest,
  tests/elastic_test.py:412-449), here against the scalar oracle, which its
  own golden vectors pin.
- Central finite differences of the cost against the reconstruction
  gradient in float64, through the PML.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import golden, golden_names

from oracle import driver as D
from oracle import elastic as E


@pytest.mark.parametrize("name", golden_names("el_fwd_"))
def test_forward_matches_the_reference_bit_for_bit(name):
    z = golden(name)
    sim = E.sim_from_npz(z)
    field, traces = E.simulate(sim, z["src_pos"], z["src_sig"], z["indicator"], z["sensors"])
    np.testing.assert_array_equal(traces, z["traces"])
    np.testing.assert_array_equal(field, z["field"])


@pytest.mark.parametrize("name", golden_names("el_adj_"))
def test_adjoints_match_the_reference_bit_for_bit(name):
    z = golden(name)
    sim = E.sim_from_npz(z)
    objective = D.l2_misfit(z["observed"])
    args = (sim, z["src_pos"], z["src_sig"], z["indicator"], z["sensors"], objective)
    cost, g, _, _ = E.sensitivity(*args)
    assert cost == float(z["s_cost"])
    np.testing.assert_array_equal(g["mass"], z["s_mass"])
    np.testing.assert_array_equal(g["stiff"], z["s_stiff"])
    cost, g, _, info = E.reconstruction_sensitivity(*args)
    assert cost == float(z["r_cost"])
    np.testing.assert_array_equal(g["mass"], z["r_mass"])
    np.testing.assert_array_equal(g["stiff"], z["r_stiff"])
    assert info["strip"] == int(z["r_strip"])
    assert info["drift"] == float(z["r_drift"])
    if "c_cost" in z:
        cost, g, _, info = E.reconstruction_sensitivity(*args, checkpoints=2)
        assert cost == float(z["c_cost"])
        np.testing.assert_array_equal(g["mass"], z["c_mass"])
        np.testing.assert_array_equal(g["stiff"], z["c_stiff"])
        assert info["segment"] == int(z["c_segment"])
        assert info["sweeps"] == float(z["c_sweeps"])
        strip, valid = E.reconstruction_nodes(sim)
        np.testing.assert_array_equal(strip, z["strip_nodes"])
        np.testing.assert_array_equal(valid, z["valid"])


@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_misfit_gradient_matches_the_reference(tag):
    z = golden(f"el_misfit_o4_{tag}")
    sim = E.sim_from_npz(z)
    cost, g = E.misfit_gradient(sim, list(zip(z["src_pos"], z["src_sig"])), z["indicator"],
                                z["sensors"], list(z["observed"]))
    assert cost == float(z["cost"])
    np.testing.assert_array_equal(g, z["gradient"])


def test_one_dimension_is_the_scalar_stencil_on_the_offset_grid():
    res, N = 400, 150
    dx = (1.0 / (res - 3),)
    cp, rho = 2.0, 1.3
    dt = 0.7 * D.stable_dt(dx, cp, 2, "taylor")
    scalar = D.Sim("ScalarWave", (res,), dx, N, dt, "float64", 2, stencil="taylor",
                   wavespeed=cp, density=rho)
    elastic = E.ElSim((res,), dx, N, dt, "float64", 2, "taylor", density=rho, wavespeed_p=cp,
                      wavespeed_s=0.8 * cp)
    t = np.arange(N) * dt
    sig = (np.exp(-((t - 0.02) / 0.005) ** 2))[:, None]
    gamma = np.full(scalar.Nx_padded, 0.8)
    # node 140 of the scalar grid is unknown 140 (at 140 + 1/2) of the staggered one
    a = D.simulate(scalar, np.array([[140]]), sig / dx[0], gamma)
    b, _ = E.simulate(elastic, np.array([[140]]), sig, gamma)
    a, b = np.asarray(a)[1:-1], np.asarray(b)[1:-2]
    assert np.abs(a).max() > 0
    assert np.abs(a[:-1] - b).max() < 1e-12 * np.abs(a).max()


def test_the_reconstruction_gradient_matches_finite_differences():
    from scipy.ndimage import gaussian_filter

    z = golden("el_adj_2d_o4_pml_f64")
    sim = E.sim_from_npz(z)
    sim.N = 60
    sig = z["src_sig"][:60]
    _, observed = E.simulate(sim, z["src_pos"], sig, z["indicator"], z["sensors"])
    objective = D.l2_misfit(0.8 * observed)
    gamma = np.array(z["indicator"])
    _, g, _, info = E.reconstruction_sensitivity(sim, z["src_pos"], sig, gamma, z["sensors"],
                                                 objective, checkpoints=2)
    gradient = g["mass"] + g["stiff"]  # gamma scales inertia and stiffness alike
    rng = np.random.default_rng(3)
    direction = gaussian_filter(rng.standard_normal(sim.Nx_padded), 2) * info["valid"]
    h = 1e-5

    def cost(step):
        return objective(E.simulate(sim, z["src_pos"], sig, gamma + step * direction,
                                    z["sensors"])[1])[0]

    fd = (cost(h) - cost(-h)) / (2 * h)
    adj = float(np.sum(gradient * direction))
    assert abs(adj - fd) <= 1e-6 * abs(fd), (adj, fd)
