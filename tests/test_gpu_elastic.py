"""Parity of the elastic HIP path, on the GPU, through the C ABI.

- Forward fields and traces against the golden vectors the reference wrote:
  1D, 2D, 3D, orders 2-8, traction and clamped faces, the folded PML, plane
  stress. The kernels round as the reference's IEEE path does, so the bits
  are expected; the bar stated is 1e-5 relative L2 (float32), 1e-10 (float64).
- The stored-history adjoint, the strip-replayed reconstruction with and
  without checkpoints, its strip and segment, against the golden vectors:
  1e-4 relative for float32 gradients, 1e-10 for float64.
- A batch of shots steps each shot bitwise as alone; the shot-summed misfit
  gradient of three shots against the reference's, in one batch and in
  batches of two with checkpoints.
- On a grid past the golden sizes, sixteen shots in one batch (four threads a
  node, their densities summed across the threads) against the oracle shot
  by shot, and the gradient against central finite differences in float64.
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest
import torch

from helpers import elastic_sim, golden, golden_names, host, rel_l2, rel_max

pytestmark = pytest.mark.gpu

S = importlib.import_module("cuwave_amd.sensitivity")
from cuwave_amd import boundary, elastic, utils, wave  # noqa: E402

F32_FIELD, F32_GRAD, F64 = 1e-5, 1e-4, 1e-10


def tol(sim, grad=False):
    if sim.precision == "float64":
        return F64
    return F32_GRAD if grad else F32_FIELD


@pytest.mark.parametrize("name", golden_names("el_fwd_"))
def test_elastic_forward_matches_the_reference(name):
    z = golden(name)
    sim = elastic_sim(z)
    print(f"case {name}", flush=True)
    field, traces = wave.simulate(sim, wave.Source(z["src_pos"], z["src_sig"]), z["indicator"],
                                  z["sensors"])
    assert rel_l2(field, z["field"]) < tol(sim)
    assert rel_l2(traces, z["traces"]) < tol(sim)
    np.testing.assert_array_equal(host(traces), z["traces"])
    np.testing.assert_array_equal(host(field), z["field"])


@pytest.mark.parametrize("name", golden_names("el_adj_"))
def test_elastic_adjoints_match_the_reference(name):
    z = golden(name)
    sim = elastic_sim(z)
    print(f"case {name}", flush=True)
    src = wave.Source(z["src_pos"], z["src_sig"])
    args = (sim, src, z["indicator"], z["sensors"], S.l2_misfit(z["observed"]))
    cost, g, traces, _ = S.sensitivity(*args)
    assert abs(cost - float(z["s_cost"])) <= tol(sim) * abs(float(z["s_cost"]))
    np.testing.assert_array_equal(host(traces), z["s_traces"])
    assert rel_max(g["mass"], z["s_mass"]) < tol(sim, True)
    assert rel_max(g["stiff"], z["s_stiff"]) < tol(sim, True)
    cost, g, _, info = S.reconstruction_sensitivity(*args)
    assert abs(cost - float(z["r_cost"])) <= tol(sim) * abs(float(z["r_cost"]))
    assert rel_max(g["mass"], z["r_mass"]) < tol(sim, True)
    assert rel_max(g["stiff"], z["r_stiff"]) < tol(sim, True)
    assert info["strip"] == int(z["r_strip"])
    # the rebuilt initial state's norm: the reference's, up to the order of the sums
    assert info["drift"] == pytest.approx(float(z["r_drift"]), rel=1e-3,
                                          abs=1e-13 if sim.precision == "float64" else 1e-9)
    if "c_cost" in z:
        cost, g, _, info = S.reconstruction_sensitivity(*args, checkpoints=2)
        assert rel_max(g["mass"], z["c_mass"]) < tol(sim, True)
        assert rel_max(g["stiff"], z["c_stiff"]) < tol(sim, True)
        assert info["segment"] == int(z["c_segment"])
        assert info["sweeps"] == float(z["c_sweeps"])
        np.testing.assert_array_equal(info["valid"], z["valid"])


def test_an_elastic_batch_steps_every_shot_as_alone():
    z = golden("el_fwd_2d_o4_pml_f32")
    sim = elastic_sim(z)
    rng = np.random.default_rng(4)
    sources = []
    for k in range(5):
        pos = np.array([[k % 2], [rng.integers(3, sim.Nx[0] - 3)], [rng.integers(3, sim.Nx[1] - 3)]])
        sources.append(wave.Source(pos, z["src_sig"][:, :1] * (1 + k)))
    fields, records = wave.simulate_shots(sim, sources, z["indicator"], z["sensors"])
    for k, s in enumerate(sources):
        f, r = wave.simulate(sim, s, z["indicator"], z["sensors"])
        assert torch.equal(fields[k], f) and torch.equal(records[k], r), k


class _Receivers:
    """Receivers on unknowns: the record is the traces."""

    def __init__(self, nodes):
        self.nodes = nodes

    def objective(self, objective):
        return objective

    def traces(self, record):
        return record


@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_elastic_misfit_gradient_matches_the_reference(tag):
    z = golden(f"el_misfit_o4_{tag}")
    sim = elastic_sim(z)
    sources = [wave.Source(p, s) for p, s in zip(z["src_pos"], z["src_sig"])]
    for kw in ({}, {"batch": 2, "checkpoints": 1}):
        print(f"case el_misfit_o4_{tag} {kw}", flush=True)
        cost, gradient = utils.misfit_gradient(sim, sources, z["indicator"],
                                               _Receivers(z["sensors"]), list(z["observed"]),
                                               adjoint=S.reconstruction_sensitivity, **kw)
        assert abs(cost - float(z["cost"])) <= tol(sim) * float(z["cost"])
        assert rel_max(gradient, z["gradient"]) < tol(sim, True)


def _wide_case(precision, shots, N, seed):
    """A 2D order-4 elastic case past the golden sizes: free top, clamped and lined
    elsewhere, a smooth random design, shots along the top."""
    sim = elastic.ElasticWave((70, 150), (1.0, 1.0), N, 0.15, precision=precision,
                              space_order=4, stencil="taylor", density=1.3, wavespeed_p=2.0,
                              wavespeed_s=1.0,
                              boundary=((boundary.Traction, boundary.Clamped),
                                        (boundary.Clamped, boundary.Clamped)),
                              pml=boundary.Pml(8, 2.0, 1e-6, 2.0, (1, 2, 3)))
    rng = np.random.default_rng(seed)
    smooth = np.cumsum(np.cumsum(rng.standard_normal(sim.Nx_padded), 0), 1)
    smooth = (smooth - smooth.min()) / (smooth.max() - smooth.min())
    gamma = (0.8 + 0.4 * smooth).astype(sim.np_dtype)
    t = np.arange(N) * sim.dt
    sig = (np.exp(-((t - 3.0) / 0.8) ** 2) * np.sin(2.0 * t))[:, None].astype(sim.np_dtype)
    cols = np.linspace(20, 130, shots).astype(int)
    # vertical forces on the free surface, the vertical displacement recorded there
    sources = [wave.Source(np.array([[0], [1], [c]]), sig * (1 + 0.1 * k))
               for k, c in enumerate(cols)]
    sensors = np.stack([np.zeros(24, int), np.ones(24, int), np.linspace(10, 140, 24).astype(int)])
    return sim, gamma, sources, sensors


def test_sixteen_shots_in_one_batch_match_the_oracle_shot_by_shot():
    from oracle import driver as D
    from oracle import elastic as E

    sim, gamma, sources, sensors = _wide_case("float32", 16, 90, 3)
    print("case elastic 70x150 o4 f32 16 shots", flush=True)
    fields, records = wave.simulate_shots(sim, sources, gamma, sensors)
    osim = E.ElSim(sim.Nx, sim.dx, sim.N, sim.dt, sim.precision, sim.space_order, sim.stencil,
                   ((0, 1), (1, 1)), sim.density, sim.wavespeed_p, sim.wavespeed_s, "strain",
                   D.Pml(8, 2.0, 1e-6, 2.0, (1, 2, 3)))
    observed = []
    for k in (0, 7, 15):
        f, r = E.simulate(osim, sources[k].position, sources[k].signal, gamma, sensors)
        np.testing.assert_array_equal(host(fields[k]), f, err_msg=f"shot {k}")
        np.testing.assert_array_equal(host(records[k]), r, err_msg=f"shot {k}")
    observed = [host(r) * 0.9 for r in records]
    cost, gradient = utils.misfit_gradient(sim, sources, gamma, _Receivers(sensors), observed,
                                           adjoint=S.reconstruction_sensitivity)
    ocost, ograd = E.misfit_gradient(osim, [(s.position, s.signal) for s in sources], gamma,
                                     sensors, observed)
    assert abs(cost - ocost) <= 1e-5 * ocost
    assert rel_max(gradient, ograd) < F32_GRAD


def test_the_elastic_gradient_matches_finite_differences_in_float64():
    sim, gamma, sources, sensors = _wide_case("float64", 2, 60, 5)
    print("case elastic fd f64", flush=True)
    observed = [np.zeros((sim.N, sensors.shape[1]))] * 2
    _, gradient = utils.misfit_gradient(sim, sources, gamma, _Receivers(sensors), observed,
                                        adjoint=S.reconstruction_sensitivity)
    g = host(gradient)
    valid = elastic.reconstruction_nodes(sim)[1]
    for node in [(12, 40), (20, 75), (6, 100)]:
        assert valid[node]
        h = 1e-5
        costs = []
        for sign in (1, -1):
            pert = gamma.copy()
            pert[node] += sign * h
            costs.append(utils.misfit(sim, sources, pert, _Receivers(sensors), observed))
        fd = (costs[0] - costs[1]) / (2 * h)
        assert abs(fd - g[node]) <= 1e-6 * max(abs(fd), np.abs(g).max()), (node, fd, g[node])
