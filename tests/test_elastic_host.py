"""The elastic family's host side, without a GPU: what the kernels will read.

- The staggered closure rows, their counts, the staggered weights and the
  norms are the reference's to the bit (tests/golden/el_stencils.npz, which
  the reference wrote), for every order and both stencils.
- The excitation weights, the PML profiles, the strip of irreversible points
  and the chain rule of the gradient densities equal the oracle's on every
  elastic golden case (the oracle is itself pinned bit for bit by those cases,
  tests/test_elastic_oracle.py).
- The C ABI's grid struct carries the simulation's constants.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import elastic_sim, golden, golden_names

from cuwave_amd import boundary, elastic, sbp
from cuwave_amd.stencils import staggered_weights
from oracle import elastic as E


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("stencil", ["taylor", "optimized"])
def test_closure_tables_are_the_reference_s(order, stencil):
    z = golden("el_stencils")
    key = f"o{order}_{stencil}"
    t = sbp.closure_tables(order, stencil)
    np.testing.assert_array_equal(t["node"], z[f"{key}_node_rows"])
    np.testing.assert_array_equal(t["half"], z[f"{key}_half_rows"])
    assert t["counts"] == tuple(int(c) for c in z[f"{key}_counts"])
    np.testing.assert_array_equal(staggered_weights(order // 2, stencil), z[f"{key}_stag"])
    node, half = sbp.closure_norms(order, 40)
    np.testing.assert_array_equal(node, z[f"o{order}_node_norm"])
    np.testing.assert_array_equal(half, z[f"o{order}_half_norm"])


def test_the_pair_sums_by_parts():
    # H+ D+ + (H- D-)^T = B on the wall points: the property the closures are for
    for order in (2, 4, 6, 8):
        n = 40
        pair = sbp.staggered_pair(order, n)
        hp, hm = np.diag(pair.node_norm), np.diag(pair.half_norm)
        lhs = hp @ pair.node + (hm @ pair.half).T
        # dropped wall points leave only the two wall nodes' boundary terms
        assert np.abs(lhs).max() < 1e-12, order


CASES = golden_names("el_fwd_") + golden_names("el_adj_") + golden_names("el_misfit")


@pytest.mark.parametrize("name", CASES)
def test_host_side_material_matches_the_oracle(name):
    z = golden(name)
    sim, osim = elastic_sim(z), E.sim_from_npz(z)
    gamma = np.ascontiguousarray(z["indicator"], dtype=sim.np_dtype)
    pos = z["src_pos"] if z["src_pos"].ndim == 2 else z["src_pos"][0]
    lin = sim.flatten_indices(pos)
    np.testing.assert_array_equal(lin, osim.flatten(pos))
    mat = {"gamma_host": gamma}
    np.testing.assert_array_equal(sim.excitation_weights(mat, lin), osim.excitation_weights(gamma, lin))
    rec = sim.flatten_indices(z["sensors"])
    np.testing.assert_array_equal(sim.excitation_weights(mat, rec), osim.excitation_weights(gamma, rec))
    np.testing.assert_array_equal(sim.adjoint_weights(z["sensors"]), osim.adjoint_weights(z["sensors"]))
    for half in ((), (0,), tuple(range(sim.ndim))):
        np.testing.assert_array_equal(sim.point_weights(half), osim.point_weights(half))
    if sim.pml is not None:
        ours = [p for p in boundary.pml_profiles(sim) if p is not None]
        theirs = [s[3] for s in osim.pml_sim().pml_slabs()]
        for a, b in zip(ours, theirs):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(boundary.pml_nodes(sim), osim.pml_nodes())
    strip, valid = elastic.reconstruction_nodes(sim)
    ostrip, ovalid = E.reconstruction_nodes(osim)
    np.testing.assert_array_equal(strip, ostrip)
    np.testing.assert_array_equal(valid, ovalid)
    if "strip_nodes" in z:
        np.testing.assert_array_equal(strip, z["strip_nodes"])


@pytest.mark.parametrize("name", ["el_adj_2d_o4_pml_f32", "el_adj_3d_o2_pml_f64", "el_adj_1d_o4_f64"])
def test_the_chain_rule_matches_the_oracle(name):
    z = golden(name)
    sim, osim = elastic_sim(z), E.sim_from_npz(z)
    rng = np.random.default_rng(7)
    gamma = np.ascontiguousarray(z["indicator"], dtype=sim.np_dtype)
    host = {"mass": rng.standard_normal((sim.ncomp, *sim.Nx_padded)).astype(sim.np_dtype),
            "normal": rng.standard_normal(sim.Nx_padded).astype(sim.np_dtype)}
    if sim.ndim > 1:
        host["shear"] = rng.standard_normal((sim.npairs, *sim.Nx_padded)).astype(sim.np_dtype)
    ours = elastic.finalize_gradients(sim, {k: torch.as_tensor(v) for k, v in host.items()}, gamma)
    theirs = E.finalize_gradients(osim, {k: v.copy() for k, v in host.items()}, gamma)
    for k in ("mass", "stiff"):
        np.testing.assert_array_equal(ours[k], theirs[k])


def test_the_abi_grid_carries_the_constants():
    z = golden("el_adj_2d_o4_pml_f32")
    sim = elastic_sim(z)
    g = sim.abi_grid(5)
    assert (g.ndim, g.order, g.nbatch, g.lanes, g.dtype) == (2, 4, 5, 8, 0)
    assert tuple(g.closure) == sbp.closure_tables(4, "taylor")["counts"]
    assert g.lame == float(np.float32(sim.lame)) and g.shear == float(np.float32(sim.shear))
    assert g.clamped == boundary.face_mask(sim, boundary.Clamped)
    assert g.step[0] == float(np.float32(sim.dt**2 / sim.dx[0]))
    assert g.inv_dx[1] == float(np.float32(1.0 / sim.dx[1]))


def test_elastic_faces_and_layers_are_validated():
    kw = dict(density=1.0, wavespeed_p=2.0, wavespeed_s=1.0, space_order=4, stencil="taylor")
    with pytest.raises(ValueError, match="traction|clamped"):
        elastic.ElasticWave((40, 40), (1.0, 1.0), 10, 0.1, boundary=boundary.Neumann, **kw)
    with pytest.raises(ValueError, match="unshifted"):
        elastic.ElasticWave((40, 40), (1.0, 1.0), 10, 0.1, boundary=boundary.Clamped,
                            pml=boundary.Pml(6, 2.0), **kw)
    with pytest.raises(ValueError, match="traps"):
        elastic.ElasticWave((40, 40), (1.0, 1.0), 10, 0.1, pml=boundary.Pml(6, 2.0, frequency=1.0),
                            **kw)
    with pytest.raises(ValueError, match="below"):
        elastic.ElasticWave((40, 40), (1.0, 1.0), 10, 0.1, density=1.0, wavespeed_p=1.0,
                            wavespeed_s=1.5)
    with pytest.raises(NotImplementedError):
        elastic.ElasticWave((40, 40), (1.0, 1.0), 10, 0.1, relaxation=object(), **kw)
