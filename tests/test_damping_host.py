"""The host side of damping (cuwave/boundary.py:162-269, cuwave/scalar.py:857-872,
cuwave/sensitivity.py:139-195), against closed forms and the oracle.

- `sponge` is 2 beta ramp^2 / (dt minv): beta at a lined wall node, zero past
  `width` nodes in and on the ghosts, a corner graded once (the larger ramp);
  it refuses a negative beta and a layer that leaves no interior;
- `pad_for_sponge` grows the grid by the width on the lined faces and maps the
  region of interest back;
- the damped source weights share the update's 1 / (1 + beta), as the oracle
  has them;
- the strip and reversible set of a damped reconstruction are the oracle's,
  with and without a PML, for the whole run and for a shorter span.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import driver as D

from cuwave_amd import boundary, scalar, sensitivity
from cuwave_amd.elastic import ElasticWave
from cuwave_amd.stencils import stable_dt


def _sim(damping=None, pml=None, precision="float64"):
    return scalar.AcousticWave((34, 31), (1.0, 1.0), 90, 0.3, precision=precision,
                               space_order=4, rho1=1.0, rho2=2.5, kappa1=1.0, kappa2=4.0,
                               damping=damping, pml=pml)


def _indicator(sim, seed=4):
    return (0.2 + 0.6 * np.random.default_rng(seed).random(sim.Nx_padded)).astype(sim.np_dtype)


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_the_sponge_is_its_closed_form(precision):
    sim = _sim(precision=precision)
    ind = _indicator(sim)
    beta, width = 0.05, 6
    field = boundary.sponge(sim, ind, width, beta)
    assert field.dtype == sim.np_dtype and field.shape == tuple(sim.Nx_padded)
    minv = sim.inverse_inertia(ind)
    # a lined wall node: the whole beta; the node `width` in: nothing
    assert field[1, 15] == pytest.approx(2 * beta / (sim.dt * minv[1, 15]), rel=1e-6)
    assert field[1 + width, 15] == 0.0 and field[15, 15] == 0.0
    # two in from a wall, far from the others: ramp (width - 2) / width
    r = (width - 2) / width
    assert field[3, 15] == pytest.approx(2 * beta * r * r / (sim.dt * minv[3, 15]), rel=1e-6)
    # a corner takes the larger ramp, not a product or a sum
    r0, r1 = 1 - 1 / width, 1 - 3 / width
    assert field[2, 4] == pytest.approx(2 * beta * max(r0, r1) ** 2 / (sim.dt * minv[2, 4]),
                                        rel=1e-6)
    assert np.all(field[0] == 0) and np.all(field[:, sim.Nx[1]:] == 0)  # ghosts, padding
    only = boundary.sponge(sim, ind, width, beta, faces=(0,))
    assert only[1, 15] > 0 and only[sim.Nx[0] - 2, 15] == 0 and only[15, 1] == 0
    with pytest.raises(ValueError):
        boundary.sponge(sim, ind, width, -0.1)
    with pytest.raises(ValueError):
        boundary.sponge(sim, ind, 16, beta)
    with pytest.raises(ValueError):
        boundary.sponge(sim, ind, 0, beta)


def test_pad_for_sponge_grows_the_lined_faces():
    grown, width, origin, region = boundary.pad_for_sponge((40, 30), (0.5, 0.25), 2.0,
                                                           faces=(0, 2, 3))
    assert width == 8  # off the finest axis
    assert grown == (48, 46) and origin == (4.0, 2.0)
    assert region == (slice(9, 47), slice(9, 37))
    with pytest.raises(ValueError):
        boundary.pad_for_sponge((40, 30), (0.5, 0.25), 0.1)


def test_a_damping_field_is_checked_and_refused_where_not_built():
    sim = _sim()
    with pytest.raises(ValueError, match="shape"):
        _sim(damping=np.zeros((3, 3)))
    with pytest.raises(ValueError, match="finite"):
        _sim(damping=np.full(sim.Nx_padded, np.nan))
    dx = (1e-3, 1e-3)
    with pytest.raises(NotImplementedError):
        ElasticWave((20, 20), dx, 10, 0.5 * stable_dt(dx, 3000.0, 4), density=1.0,
                    wavespeed_p=3000.0, wavespeed_s=1500.0, space_order=4,
                    damping=np.zeros((20, 32)))


def _oracle(sim, damping):
    pml = None if sim.pml is None else D.Pml(sim.pml.width, sim.pml.wavespeed,
                                             sim.pml.reflection)
    return D.Sim("AcousticWave", sim.Nx, sim.dx, sim.N, sim.dt, sim.precision,
                 sim.space_order, materials=(1.0, 2.5, 1.0, 4.0), damping=damping, pml=pml)


def test_damped_source_weights_share_the_divisor():
    sim0 = _sim()
    ind = _indicator(sim0)
    damping = boundary.sponge(sim0, ind, 6, 0.05)
    sim = _sim(damping=damping)
    stiff, minv = sim.parametrization(ind)
    mat = {"stiff_host": scalar.mirror_ghosts(sim, np.array(stiff)),
           "minv_host": scalar.mirror_ghosts(sim, np.array(minv))}
    osim = _oracle(sim, damping)
    lin = sim.flatten_indices(np.array([[1, 2, 12, 30], [5, 15, 12, 28]]))
    np.testing.assert_array_equal(sim.excitation_weights(mat, lin),
                                  osim.excitation_weights(osim.build_materials(ind), lin))
    plain = _sim().excitation_weights(mat, lin)
    beta = 0.5 * mat["minv_host"].ravel()[lin] * damping.ravel()[lin] * sim.dt
    np.testing.assert_allclose(sim.excitation_weights(mat, lin), plain / (1 + beta), rtol=1e-15)
    back = -damping
    np.testing.assert_array_equal(sim.excitation_weights(mat, lin, damping=back),
                                  osim.excitation_weights(osim.build_materials(ind), lin,
                                                          damping=back))


@pytest.mark.parametrize("with_pml", [False, True])
def test_the_damped_strip_is_the_oracles(with_pml):
    sim0 = _sim()
    ind = _indicator(sim0)
    damping = boundary.sponge(sim0, ind, 6, 0.05)
    damping[10:17, 10:17] = 4.0 / (sim0.dt * sim0.inverse_inertia(ind)[10:17, 10:17])
    pml = boundary.Pml(5, 2.0) if with_pml else None
    sim = _sim(damping=damping, pml=pml)
    osim = _oracle(sim, damping)
    for span in (None, 20):
        strip, valid = sensitivity.reconstruction_nodes(sim, ind, span)
        ostrip, ovalid = D.reconstruction_nodes(osim, ind, span)
        np.testing.assert_array_equal(strip, ostrip)
        np.testing.assert_array_equal(valid, ovalid)
        assert not valid[10:17, 10:17].any() and valid[20, 20]
    np.testing.assert_array_equal(sensitivity.reversible_damping(sim, ind, 90),
                                  D.reversible_damping(osim, ind, 90))
