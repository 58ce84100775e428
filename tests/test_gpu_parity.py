"""Parity of the HIP path with the reference, on the GPU, through the C ABI.

The golden vectors are the reference's own outputs (its CPU backend, IEEE
arithmetic). The kernels round as that backend does, so the forward field and
traces are expected bit for bit; the bar the tests hold every result to is the
stated tolerance: 1e-5 relative L2 for float32 fields and traces, 1e-4
relative for float32 gradients, 1e-10 for float64.
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest
import torch

from helpers import golden, golden_names, host, oracle_sim, product_sim, rel_l2, rel_max

pytestmark = pytest.mark.gpu

S = importlib.import_module("cuwave_amd.sensitivity")
from cuwave_amd import utils, wave  # noqa: E402

F32_FIELD, F32_GRAD, F64 = 1e-5, 1e-4, 1e-10


def tol(sim, grad=False):
    if sim.precision == "float64":
        return F64
    return F32_GRAD if grad else F32_FIELD


@pytest.mark.parametrize("name", golden_names("fwd_") + ["disk_298x298_o8_forward"])
def test_forward_matches_the_reference(name):
    z = golden(name)
    sim = product_sim(z)
    print(f"case {name}", flush=True)
    field, traces = wave.simulate(sim, wave.Source(z["src_pos"], z["src_sig"]), z["indicator"],
                                  z["sensors"])
    assert rel_l2(field, z["field"]) < tol(sim)
    assert rel_l2(traces, z["traces"]) < tol(sim)
    # the kernels round as the reference's IEEE path does: the same bits
    np.testing.assert_array_equal(host(field), z["field"])
    np.testing.assert_array_equal(host(traces), z["traces"])


def _objective(z, sim):
    obs = z["observed"] if "observed" in z else np.zeros((sim.N, z["sensors"].shape[1]), sim.np_dtype)
    return S.l2_misfit(obs)


@pytest.mark.parametrize("name", golden_names("adj_"))
def test_adjoint_variants_match_the_reference(name):
    z = golden(name)
    sim = product_sim(z)
    print(f"case {name}", flush=True)
    src = wave.Source(z["src_pos"], z["src_sig"])
    args = (sim, src, z["indicator"], z["sensors"], _objective(z, sim))
    cost, g, traces, _ = S.sensitivity(*args)
    assert abs(cost - float(z["cost"])) <= tol(sim) * abs(float(z["cost"]))
    assert rel_l2(traces, z["traces"]) < tol(sim)
    assert rel_max(g["mass"], z["g_mass"]) < tol(sim, True)
    assert rel_max(g["stiff"], z["g_stiff"]) < tol(sim, True)
    _, g, _, info = S.reconstruction_sensitivity(*args)
    assert rel_max(g["mass"], z["r_g_mass"]) < tol(sim, True)
    assert rel_max(g["stiff"], z["r_g_stiff"]) < tol(sim, True)
    if "r_strip" in z:
        assert info["strip"] == int(z["r_strip"])
    assert info["drift"] < (1e-4 if sim.precision == "float32" else 1e-12)
    if "c_g_mass" in z:
        _, g, _, info = S.reconstruction_sensitivity(*args, checkpoints=2)
        assert rel_max(g["mass"], z["c_g_mass"]) < tol(sim, True)
        assert rel_max(g["stiff"], z["c_g_stiff"]) < tol(sim, True)
        assert info["segment"] == int(z["c_segment"])
        assert info["sweeps"] == float(z["c_sweeps"])
    if "b_g_mass" in z:
        # a bfloat16 strip trades exactness for memory: held to the exact gradient
        _, g, _, _ = S.reconstruction_sensitivity(*args, strip_dtype="bfloat16")
        assert rel_l2(g["mass"], z["r_g_mass"]) < 1e-2
    if "s_g_mass" in z:
        _, g, _, info = S.superposition_sensitivity(*args, scale=float(z["s_scale"]))
        assert rel_max(g["mass"], z["s_g_mass"]) < tol(sim, True)
        assert rel_max(g["stiff"], z["s_g_stiff"]) < tol(sim, True)


@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_batched_misfit_gradient_matches_the_reference(tag):
    z = golden(f"misfit_acoustic_o4_{tag}")
    sim = product_sim(z)
    sensors = utils.Sensors(sim, z["rec_coords"])
    sources = [wave.Source(p, s) for p, s in zip(z["src_pos"], z["src_sig"])]
    for kw in ({}, {"batch": 2, "checkpoints": 1}):
        cost, gradient = utils.misfit_gradient(sim, sources, z["indicator"], sensors,
                                               list(z["observed"]), **kw)
        assert abs(cost - float(z["cost"])) <= tol(sim) * float(z["cost"])
        assert rel_max(gradient, z["r_gradient"]) < tol(sim, True)


def test_the_gradient_sweep_gathers_the_illumination_and_stops_above_a_bound():
    # illum= is the illumination map of the same shots, whatever the checkpoints;
    # below= runs the adjoint half only under the bound, bitwise as without it
    z = golden("misfit_acoustic_o4_f64")
    sim = product_sim(z)
    sensors = utils.Sensors(sim, z["rec_coords"])
    sources = [wave.Source(p, s) for p, s in zip(z["src_pos"], z["src_sig"])]
    observed = list(z["observed"])
    print("case illum and below misfit_acoustic_o4_f64", flush=True)
    expect = wave.illumination(sim, sources, z["indicator"])
    for kw in ({}, {"batch": 2, "checkpoints": 1}):
        illum = torch.zeros_like(expect)
        cost, gradient = utils.misfit_gradient(sim, sources, z["indicator"], sensors,
                                               observed, illum=illum, **kw)
        assert rel_l2(host(illum), host(expect)) < 1e-12, kw
        c, g = utils.misfit_gradient(sim, sources, z["indicator"], sensors, observed,
                                     below=cost * (1 + 1e-9), **kw)
        assert c == cost and torch.equal(g, gradient), kw
        c, g = utils.misfit_gradient(sim, sources, z["indicator"], sensors, observed,
                                     below=cost, **kw)
        assert c == cost and g is None, kw


def test_a_batch_steps_every_shot_as_alone():
    # shots in one launch: each bitwise its single-shot run
    z = golden("fwd_pml_acoustic_o8_f32")
    sim = product_sim(z)
    rng = np.random.default_rng(3)
    sources = []
    for k in range(5):
        pos = np.array([[rng.integers(2, sim.Nx[0] - 2)], [rng.integers(2, sim.Nx[1] - 2)]])
        sources.append(wave.Source(pos, z["src_sig"][:, :1] * (1 + k)))
    fields, records = wave.simulate_shots(sim, sources, z["indicator"], z["sensors"])
    for k, s in enumerate(sources):
        f, r = wave.simulate(sim, s, z["indicator"], z["sensors"])
        assert torch.equal(fields[k], f) and torch.equal(records[k], r)


def test_one_adjoint_level_is_the_transpose_of_one_forward_step():
    # <W m A^T p, u> = <W m p, A u> in float64, through the kernels themselves,
    # for every shot of a batch of three (four lanes, one of them padding)
    from cuwave_amd import boundary, scalar
    from cuwave_amd.engine import Propagator

    for order, bnd in ((2, boundary.Neumann), (4, boundary.Dirichlet), (8, ((boundary.Dirichlet, boundary.Neumann), (boundary.Neumann, boundary.Dirichlet)))):
        sim = scalar.AcousticWave((37, 45), (0.1, 0.1), 1, 0.3 * wave.stable_dt((0.1, 0.1), 1.0, order),
                                  precision="float64", space_order=order, boundary=bnd,
                                  rho1=1.0, rho2=3.0, kappa1=1.0, kappa2=5.0)
        rng = np.random.default_rng(0)
        mat = sim.build_materials(0.2 + rng.random(sim.Nx_padded))
        prop = Propagator(sim, mat, 3)
        assert prop.lanes == 4
        dev = prop.device
        mask = np.zeros(sim.Nx_padded)
        mask[tuple(slice(1, n - 1) for n in sim.Nx)] = 1.0
        U, L = prop.levels_memory(2), prop.levels_memory(2)
        # level -1 (slot 1) is u with its ghosts refilled, level -2 (slot 0) zero;
        # lambda^1 (slot 1) is p, lambda^2 zero
        for T in (U, L):
            for b in range(3):
                level = torch.as_tensor(rng.standard_normal(sim.Nx_padded) * mask, device=dev)
                S._refill(sim, level)
                T[1, ..., b] = level
        u, p = host(prop.shots(U[1])) * mask, host(prop.shots(L[1])) * mask
        prop.forward(U, 0, 1)
        Au = host(prop.shots(U[0])) * mask - 2 * u
        g = prop.accumulator()
        prop.backward(L, U, 0, 1, eta=prop.eta_memory(), g_mass=g, g_stiff=g.clone())
        Ap = host(prop.shots(L[0])) * mask - 2 * p
        Wm = sim.cell_weights() * mask / mat["minv_host"]
        for b in range(3):
            scale = np.sqrt(np.sum(Wm * Ap[b] ** 2) * np.sum(Wm * u[b] ** 2))
            assert abs(np.sum(Wm * Ap[b] * u[b]) - np.sum(Wm * p[b] * Au[b])) / scale < 1e-12, (order, b)
        assert not host(U[..., 3]).any() and not host(L[..., 3]).any()  # the padding lane stays zero


def test_the_gradient_matches_finite_differences_in_float64():
    z = golden("adj_pml_o8_dirichlet_f64")
    sim = product_sim(z, )
    sim.N = 90
    src = wave.Source(z["src_pos"], z["src_sig"][: sim.N])
    obs = host(wave.simulate(sim, src, z["indicator"], z["sensors"])[1]) * 0.7
    objective = S.l2_misfit(obs)
    indicator = np.array(z["indicator"])
    _, grads, _, _ = S.reconstruction_sensitivity(sim, src, indicator, z["sensors"], objective)
    jm, js = sim.parametrization_jacobian()
    gradient = host(jm * grads["mass"] + js * grads["stiff"])

    def cost():
        return objective(wave.simulate(sim, src, indicator, z["sensors"])[1])[0]

    for node in [(15, 15), (20, 12), (12, 22), (25, 18)]:
        base, h = indicator[node], 1e-6
        indicator[node] = base + h
        plus = cost()
        indicator[node] = base - h
        minus = cost()
        indicator[node] = base
        fd = (plus - minus) / (2 * h)
        assert abs(gradient[node] - fd) <= 1e-6 * abs(fd), (node, gradient[node], fd)


def test_the_illumination_map_matches_the_oracle():
    # sum over shots and steps of u^2, three shots in one batch and in two
    from oracle import driver

    z = golden("fwd_pml_acoustic_o8_f32")
    sim, osim = product_sim(z), oracle_sim(z)
    print("case illumination fwd_pml_acoustic_o8_f32", flush=True)
    rng = np.random.default_rng(7)
    sources = []
    for k in range(3):
        pos = np.array([[rng.integers(3, sim.Nx[0] - 3)], [rng.integers(3, sim.Nx[1] - 3)]])
        sources.append(wave.Source(pos, z["src_sig"][:, :1] * (1 + k)))
    expect = driver.illumination(osim, [(s.position, s.signal) for s in sources], z["indicator"])
    for batch in (None, 2):
        illum = wave.illumination(sim, sources, z["indicator"], batch=batch)
        assert rel_l2(illum, expect) < 1e-5, batch


def test_a_3d_order_8_float32_adjoint_matches_the_oracle():
    # the 3D one-lane float adjoint at order 8 (its own register budget)
    from oracle import driver

    base = golden("adj_pml_3d_o4_f64")
    z = {k: base[k] for k in base.files}
    z.update(space_order=np.array(8), precision=np.array("float32"), N=np.array(40))
    z["indicator"] = base["indicator"].astype(np.float32)
    z["src_sig"] = base["src_sig"][:40].astype(np.float32)
    sim, osim = product_sim(z), oracle_sim(z)
    print("case 3d order 8 float32 adjoint", flush=True)
    src = wave.Source(z["src_pos"], z["src_sig"])
    observed = np.zeros((sim.N, z["sensors"].shape[1]), np.float32)
    cost, g, traces, _ = S.sensitivity(sim, src, z["indicator"], z["sensors"],
                                       S.l2_misfit(observed))
    ocost, og, otraces, _ = driver.sensitivity(osim, z["src_pos"], z["src_sig"], z["indicator"],
                                               z["sensors"], driver.l2_misfit(observed))
    np.testing.assert_array_equal(host(traces), otraces)
    assert abs(cost - ocost) <= F32_FIELD * abs(ocost)
    assert rel_max(g["mass"], og["mass"]) < F32_GRAD
    assert rel_max(g["stiff"], og["stiff"]) < F32_GRAD


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_3d_order_8_bulk_density_reconstruction_with_checkpoints_matches_the_oracle(precision):
    # the shape of the 3D acoustic benchmark at a size the oracle finishes in
    # seconds: independent bulk modulus and density (smoothed noise plus a low
    # density void), order 8, a CPML on every face, the reverse march with its
    # strip replay and two treeverse checkpoints
    from scipy.ndimage import gaussian_filter

    from cuwave_amd import boundary, scalar
    from oracle import driver

    n, width, N, dt = 40, 7, 120, 2.0e-7
    dx = 1e-3
    rng = np.random.default_rng(11)
    npad = (n, n, 64)

    def field(lo, hi):
        f = gaussian_filter(rng.standard_normal(npad), 3)
        return lo + (hi - lo) * (f - f.min()) / (f.max() - f.min())

    rho, vp = field(900.0, 1300.0), field(1400.0, 1800.0)
    kappa = rho * vp**2
    c = (np.arange(n) - n / 2) ** 2
    void = (c[:, None, None] + c[None, :, None] + c[None, None, :] <= 16)
    void = np.pad(void, ((0, 0), (0, 0), (0, npad[2] - n)))
    rho[void], kappa[void] = 1.204, 1.419e5
    dtype = np.float32 if precision == "float32" else np.float64
    stack = np.stack([kappa, rho]).astype(dtype)
    pml = boundary.Pml(width, 1800.0, 1e-6)
    sim = scalar.BulkDensityWave((n, n, n), (dx,) * 3, N, dt, precision=precision,
                                 space_order=8, pml=pml)
    osim = driver.Sim("BulkDensityWave", (n, n, n), (dx,) * 3, N, dt, precision, 8,
                      pml=driver.Pml(width, 1800.0, 1e-6))
    t = np.arange(N) * dt
    from cuwave_amd.signals import ricker

    sig = (ricker(t, 1.0, 1e5) / dx**3)[:, None].astype(dtype)
    src_pos = np.array([[width + 2], [n // 2], [n // 2]])
    g = np.arange(width + 2, n - width - 2, 4)
    sensors = np.stack([np.full(len(g) ** 2, width + 2), np.repeat(g, len(g)), np.tile(g, len(g))])
    print(f"case 3d order 8 bulk-density reconstruction checkpoints=2 {precision}", flush=True)
    _, observed = driver.simulate(osim, src_pos, sig, stack, sensors)
    observed = (0.9 * observed).astype(dtype)
    cost, grads, traces, info = S.reconstruction_sensitivity(
        sim, wave.Source(src_pos, sig), stack, sensors, S.l2_misfit(observed), checkpoints=2)
    ocost, ograds, otraces, oinfo = driver.reconstruction_sensitivity(
        osim, src_pos, sig, stack, sensors, driver.l2_misfit(observed), checkpoints=2)
    assert info["segment"] < N and info["sweeps"] > 1.0  # the checkpointed schedule ran
    assert info["segment"] == oinfo["segment"] and info["sweeps"] == oinfo["sweeps"]
    assert info["strip"] == oinfo["strip"]
    bar = F64 if precision == "float64" else F32_GRAD
    assert rel_l2(traces, otraces) < (F64 if precision == "float64" else F32_FIELD)
    assert abs(cost - ocost) <= bar * abs(ocost)
    for key in ("mass", "stiff"):
        assert rel_max(grads[key], ograds[key]) < bar, key
