"""Host logic of the product against the oracle and the golden vectors (no GPU).

What the product computes on the host before any kernel runs: stencil rows, PML
profiles, excitation and adjoint weights, the strip, the treeverse schedule, the
multilinear interpolation, and the refusals the reference makes."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import golden, golden_names, oracle_sim, product_sim

from cuwave_amd import boundary, scalar, stencils, utils
from cuwave_amd import sensitivity as S  # noqa: F401  (module, not the function)
import importlib

S = importlib.import_module("cuwave_amd.sensitivity")
from oracle import driver as D  # noqa: E402


def test_stencils_match_the_reference():
    z = golden("stencils")
    for R in range(1, 5):
        for kind in ("taylor", "optimized"):
            np.testing.assert_array_equal(stencils.cell_coefficients(R, kind), z[f"cell_{kind}_{R}"])
            np.testing.assert_array_equal(stencils.staggered_weights(R, kind), z[f"stag_{kind}_{R}"])
    for order in (2, 4, 6, 8):
        for kind in ("taylor", "optimized"):
            assert stencils.stable_dt((1e-3, 2e-3), 1500.0, order, kind) == float(
                z[f"stable_dt_{kind}_{order}"])


@pytest.mark.parametrize("name", golden_names("adj_pml"))
def test_pml_profiles_and_layer_match_the_oracle(name):
    z = golden(name)
    sim, osim = product_sim(z), oracle_sim(z)
    for mine, theirs in zip(boundary.pml_profiles(sim), osim.pml_profiles()):
        if theirs is None:
            assert mine is None
        else:
            np.testing.assert_array_equal(mine, theirs)
    np.testing.assert_array_equal(boundary.pml_nodes(sim), osim.pml_nodes())


@pytest.mark.parametrize("name", golden_names("adj_"))
def test_strip_and_valid_match_the_reference(name):
    z = golden(name)
    sim = product_sim(z)
    strip, valid = S.reconstruction_nodes(sim)
    if "strip_nodes" in z:
        np.testing.assert_array_equal(strip, z["strip_nodes"])
        np.testing.assert_array_equal(valid, z["valid"])


@pytest.mark.parametrize("name", golden_names("fwd_") + golden_names("adj_"))
def test_material_and_weights_match_the_oracle(name):
    z = golden(name)
    sim, osim = product_sim(z), oracle_sim(z)
    stiff, minv = sim.parametrization(np.asarray(z["indicator"], dtype=sim.np_dtype))
    ost, omi = osim.parametrization(np.asarray(z["indicator"], dtype=osim.dtype))
    np.testing.assert_array_equal(stiff, ost)
    if omi is not None:
        np.testing.assert_array_equal(minv, omi)
    mat = {"stiff_host": scalar.mirror_ghosts(sim, np.array(stiff)),
           "minv_host": None if minv is None else scalar.mirror_ghosts(sim, np.array(minv))}
    omat = osim.build_materials(z["indicator"])
    lin = sim.flatten_indices(z["src_pos"])
    np.testing.assert_array_equal(sim.excitation_weights(mat, lin), osim.excitation_weights(omat, lin))
    np.testing.assert_array_equal(sim.sensor_cell_weights(z["sensors"]),
                                  osim.sensor_cell_weights(z["sensors"]))
    assert [float(f) for f in sim.step_factors()] == [float(f) for f in osim.step_factors()]


@pytest.mark.parametrize("segments,slots", [(1, 0), (5, 0), (7, 2), (12, 3), (30, 4)])
def test_checkpoint_schedule_matches_the_oracle(segments, slots):
    assert S.checkpoint_schedule(segments, slots) == D.checkpoint_schedule(segments, slots)


def test_checkpoint_schedule_matches_the_oracle_over_a_range():
    # the product's stack formulation against the oracle's recursive restatement
    for segments in range(1, 41):
        for slots in range(0, 6):
            assert S.checkpoint_schedule(segments, slots) == D.checkpoint_schedule(segments, slots)


def test_every_segment_is_reversed_once():
    for segments, slots in ((9, 2), (20, 3)):
        leaves = [a[1] for a in S.checkpoint_schedule(segments, slots) if a[0] == "leaf"]
        assert leaves == list(range(segments - 1, -1, -1))


def test_sensors_and_sources_match_the_reference():
    z = golden("misfit_acoustic_o4_f32")
    sim = product_sim(z)
    sens = utils.Sensors(sim, z["rec_coords"])
    np.testing.assert_array_equal(sens.nodes, z["sensor_nodes"])
    np.testing.assert_array_equal(sens.weights, z["sensor_weights"])
    t = np.arange(sim.N) * sim.dt
    from cuwave_amd.signals import ricker

    for c, pos, sig in zip(z["coords"], z["src_pos"], z["src_sig"]):
        src = utils.point_source(sim, c, ricker(t, 1.0, 6.0))
        np.testing.assert_array_equal(src.position, pos)
        np.testing.assert_array_equal(src.signal, sig)


def test_refusals():
    sim = scalar.ScalarWave((30, 28), (0.1, 0.1), 10, 0.01, wavespeed=1.0, density=1.0)
    with pytest.raises(ValueError, match="ghost"):
        sim.require_interior(np.array([[0], [5]]), "sensor")
    with pytest.raises(ValueError, match="ghost"):
        sim.require_interior(np.array([[5], [27]]), "source")
    with pytest.raises(ValueError, match="space_order"):
        scalar.ScalarWave((30, 28), (0.1, 0.1), 10, 0.01, space_order=3, wavespeed=1.0, density=1.0)
    with pytest.raises(ValueError, match="requires"):
        scalar.ScalarWave((30, 28), (0.1, 0.1), 10, 0.01)
    with pytest.raises(ValueError, match="leave no interior"):
        scalar.ScalarWave((20, 28), (0.1, 0.1), 10, 0.01, wavespeed=1.0, density=1.0,
                          pml=boundary.Pml(8, 1.0))
    with pytest.raises(NotImplementedError):
        scalar.ScalarWave((30, 28), (0.1, 0.1), 10, 0.01, wavespeed=1.0, density=1.0,
                          devices=(0, 1))
    dirichlet = scalar.ScalarWave((30, 28), (0.1, 0.1), 10, 0.01, wavespeed=1.0, density=1.0,
                                  boundary=boundary.Dirichlet)
    with pytest.raises(ValueError, match="held at zero"):
        boundary.require_off_held_walls(dirichlet, np.array([[1], [5]]))
