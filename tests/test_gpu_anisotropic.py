"""The cell-assembled elastic kernels against the oracle on the GPU.

- the forward march, several shots in one batch, bit for bit against the
  oracle's one-shot march: 1D, 2D and 3D, traction, clamped and mixed faces,
  isotropic and orthotropic C, plane stress, float64 and float32;
- the gradient kernel's two densities bit for bit against the oracle's;
- `sensitivity` against central finite differences of the misfit
  (cuwave tests/anisotropic_test.py:315-363);
- `reconstruction_sensitivity` (the field rebuilt by the reverse march)
  against the stored-history gradient;
- the top-level drivers reaching the family.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cuwave_amd import _abi
from cuwave_amd import anisotropic as AN
from cuwave_amd.boundary import Clamped, Traction
from cuwave_amd.wave import Source
from oracle import anisotropic as O

pytestmark = pytest.mark.gpu

FAMILY = "anisotropic"
FACES = {"traction": None, "clamped": Clamped,
         "mixed": lambda n: ((Traction, Clamped),) + ((Clamped, Traction),) * (n - 1)}
ORTHO = np.array([[30.0, 6.0, 0.0], [6.0, 12.0, 0.0], [0.0, 0.0, 4.0]])


@pytest.fixture(scope="module", autouse=True)
def library():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    _abi.load()


def aniso_sim(Nx, faces="traction", precision="float64", C=None, plane="strain", N=30):
    boundary = FACES[faces]
    if callable(boundary):
        boundary = boundary(len(Nx))
    dx = tuple(0.01 * (1.0 + 0.1 * d) for d in range(len(Nx)))
    return AN.AnisotropicElasticWave(Nx, dx, N, 1e-3, precision=precision, density=2.0,
                                     wavespeed_p=3.0, wavespeed_s=1.5, plane=plane, C=C,
                                     boundary=boundary)


def misfit_case():
    sim = aniso_sim((18, 20), "mixed", N=40, C=ORTHO)
    return sim, 0.5 + np.random.default_rng(30).random(sim.Nx_padded)


def logical(sim, a):
    return a[(..., *(slice(0, n) for n in sim.Nx))]


def oracle_march(sim, gamma, source, sensors):
    mat = sim.build_materials(gamma)
    minv, cell = logical(sim, mat["minv"]), logical(sim, mat["cell"])
    pos = np.asarray(source.position)
    if sim.ncomp == 1:
        pos = np.concatenate([np.zeros((1, pos.shape[1]), dtype=pos.dtype), pos])
        sensors = np.concatenate([np.zeros((1, sensors.shape[1]), dtype=sensors.dtype), sensors])
    weight = (sim.np_dtype(sim.dt**2) * minv[tuple(pos[1:])]).astype(sim.np_dtype)
    return O.simulate(minv, cell, mat["table"], sim.np_dtype(sim.dt**2), sim.N, tuple(pos),
                      weight, np.asarray(source.signal, dtype=sim.np_dtype), tuple(sensors))


def shots(sim, rng, count):
    out = []
    for _ in range(count):
        pos = [rng.integers(0, sim.ncomp, 2)] if sim.ncomp > 1 else []
        pos += [rng.integers(2, n - 2, 2) for n in sim.Nx]
        out.append(Source(np.array(pos), rng.standard_normal((sim.N, 2)).astype(sim.np_dtype)))
    return out


def sensor_points(sim, rng, count=5):
    pos = [rng.integers(0, sim.ncomp, count)] if sim.ncomp > 1 else []
    pos += [rng.integers(1, n - 1, count) for n in sim.Nx]
    return np.array(pos)


CASES = [((40,), "traction", None, "strain"), ((40,), "mixed", None, "strain"),
         ((22, 26), "traction", None, "strain"), ((22, 26), "clamped", None, "strain"),
         ((22, 26), "mixed", ORTHO, "strain"), ((22, 26), "traction", None, "stress"),
         ((12, 13, 14), "traction", None, "strain"), ((12, 13, 14), "mixed", None, "strain")]


@pytest.mark.parametrize("Nx, faces, C, plane", CASES)
def test_the_forward_march_is_the_oracles_bit_for_bit(Nx, faces, C, plane):
    print(f"anisotropic forward {Nx} {faces} C={C is not None} {plane}", flush=True)
    sim = aniso_sim(Nx, faces, C=C, plane=plane)
    rng = np.random.default_rng(len(Nx))
    gamma = 0.5 + rng.random(sim.Nx_padded)
    sources = shots(sim, rng, 3)
    sensors = sensor_points(sim, rng)
    field, record = AN.simulate_shots(sim, sources, gamma, sensors)
    for b, source in enumerate(sources):
        last, traces = oracle_march(sim, gamma, source, sensors)
        assert np.array_equal(field[b].cpu().numpy(), last), f"shot {b}"
        assert np.array_equal(record[b].cpu().numpy(), traces), f"shot {b}"


@pytest.mark.parametrize("Nx", [(22, 26), (12, 13, 14)])
def test_the_float32_march_is_the_oracles_bit_for_bit(Nx):
    print(f"anisotropic float32 {Nx}", flush=True)
    sim = aniso_sim(Nx, "mixed", precision="float32")
    rng = np.random.default_rng(3)
    gamma = (0.5 + rng.random(sim.Nx_padded)).astype(np.float32)
    sources = shots(sim, rng, 5)
    sensors = sensor_points(sim, rng)
    field, record = AN.simulate_shots(sim, sources, gamma, sensors)
    for b, source in enumerate(sources):
        last, traces = oracle_march(sim, gamma, source, sensors)
        assert np.array_equal(field[b].cpu().numpy(), last), f"shot {b}"
        assert np.array_equal(record[b].cpu().numpy(), traces), f"shot {b}"


@pytest.mark.parametrize("Nx", [(20, 23), (11, 12, 13)])
def test_the_gradient_kernel_is_the_oracles(Nx):
    print(f"anisotropic gradient kernel {Nx}", flush=True)
    sim = aniso_sim(Nx, "mixed")
    rng = np.random.default_rng(4)
    mat = sim.build_materials(0.5 + rng.random(sim.Nx_padded))
    eng = AN.AnisoEngine(sim, mat, 1)
    free = np.zeros((sim.ncomp, *sim.Nx_padded, 1), dtype=bool)
    free[(slice(None), *(slice(1, n - 1) for n in sim.Nx))] = True
    fields = [np.where(free, rng.standard_normal(free.shape), 0.0) for _ in range(4)]
    u0, u1, u2, l1 = (torch.as_tensor(f, device="cuda") for f in fields)
    g0 = [rng.standard_normal(sim.Nx_padded) for _ in range(2)]
    grads = {"mass": torch.as_tensor(g0[0], device="cuda"),
             "cell": torch.as_tensor(g0[1], device="cuda")}
    eng.gradient(grads, u0, u1, u2, l1)
    gm, gc = logical(sim, g0[0]).copy(), logical(sim, g0[1]).copy()
    O.gradient(gm, gc, *(logical(sim, f[..., 0]) for f in fields), mat["table"],
               sim.mass_factor())
    assert np.array_equal(logical(sim, grads["mass"].cpu().numpy()), gm)
    assert np.array_equal(logical(sim, grads["cell"].cpu().numpy()), gc)


@pytest.mark.parametrize("Nx, faces", [((18, 20), "mixed"), ((9, 10, 11), "traction")])
def test_the_gradient_matches_finite_differences(Nx, faces):
    print(f"anisotropic FD {Nx} {faces}", flush=True)
    sim = aniso_sim(Nx, faces, N=80)
    rng = np.random.default_rng(6)
    gamma = 0.7 + 0.6 * rng.random(sim.Nx_padded)
    t = np.arange(sim.N) * sim.dt
    f0 = 4.0 / (sim.N * sim.dt)
    arg = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    pulse = (1 - 2 * arg) * np.exp(-arg)
    centre = [n // 2 for n in Nx]
    source = Source(np.array([[1], *([c] for c in centre)]), pulse[:, None])
    sensors = np.array([[0, 1, sim.ncomp - 1], *([c - 2, c + 2, c + 1] for c in centre)])
    observed = AN.simulate(sim, source, np.ones(sim.Nx_padded), sensors)[1].cpu().numpy()

    def objective(um):
        r = um.cpu().numpy() - observed
        return 0.5 * float(np.sum(r * r)), torch.as_tensor(r)

    cost, grads, _, _ = AN.sensitivity(sim, source, gamma, sensors, objective)
    g = (grads["mass"] + grads["stiff"]).cpu().numpy()
    h = 1e-6
    for node in [tuple(c + 1 for c in centre), tuple(c - 1 for c in centre),
                 tuple(min(c + 2, n - 2) for c, n in zip(centre, Nx))]:
        costs = []
        for s in (1, -1):
            probe = gamma.copy()
            probe[node] += s * h
            costs.append(objective(AN.simulate(sim, source, probe, sensors)[1])[0])
        fd = (costs[0] - costs[1]) / (2 * h)
        assert abs(g[node] - fd) <= 1e-5 * abs(fd) + 1e-10 * abs(cost), (node, g[node], fd)


def test_the_top_level_drivers_reach_the_family():
    from cuwave_amd import sensitivity as S
    from cuwave_amd import wave as W

    print("anisotropic top-level drivers", flush=True)
    sim = aniso_sim((16, 18), "mixed", N=25)
    rng = np.random.default_rng(8)
    gamma = 0.5 + rng.random(sim.Nx_padded)
    source = shots(sim, rng, 1)[0]
    sensors = sensor_points(sim, rng, 3)
    mine = AN.simulate(sim, source, gamma, sensors)
    top = W.simulate(sim, source, gamma, sensors)
    assert all(torch.equal(a, b) for a, b in zip(mine, top))

    def objective(um):
        return 0.5 * float(torch.sum(um * um)), um

    _, g1, _, _ = AN.sensitivity(sim, source, gamma, sensors, objective)
    _, g2, _, _ = S.sensitivity(sim, source, gamma, sensors, objective)
    assert torch.equal(g1["stiff"], g2["stiff"]) and torch.equal(g1["mass"], g2["mass"])


@pytest.mark.parametrize("Nx, faces, precision, bound",
                         [((18, 20), "mixed", "float64", 1e-9), ((9, 10, 11), "clamped", "float64", 1e-9),
                          ((20, 22), "traction", "float32", 2e-3)])
def test_the_reverse_march_reproduces_the_stored_gradient(Nx, faces, precision, bound):
    # no pml: the leapfrog runs backwards exactly but for rounding
    print(f"anisotropic reconstruction {Nx} {faces} {precision}", flush=True)
    from cuwave_amd import sensitivity as S

    sim = aniso_sim(Nx, faces, precision=precision, N=60)
    rng = np.random.default_rng(13)
    gamma = (0.5 + rng.random(sim.Nx_padded)).astype(sim.np_dtype)
    source = shots(sim, rng, 1)[0]
    sensors = sensor_points(sim, rng, 4)

    def objective(um):
        return 0.5 * float(torch.sum(um * um)), um

    c1, g1, u1, _ = AN.sensitivity(sim, source, gamma, sensors, objective)
    c2, g2, u2, info = S.reconstruction_sensitivity(sim, source, gamma, sensors, objective)
    assert c1 == c2 and torch.equal(u1, u2)
    for k in ("mass", "stiff"):
        gap = float(torch.linalg.vector_norm(g2[k] - g1[k]) / torch.linalg.vector_norm(g1[k]))
        assert gap < bound, (k, gap)
    assert info["strip"] == 0 and info["drift"] < (1e-10 if precision == "float64" else 1e-3)


@pytest.mark.parametrize("batch", [None, 2])
def test_the_misfit_gradient_sums_the_shots(batch):
    # one sweep of three shots against the three single-shot gradients, chained
    from helpers import Receivers

    from cuwave_amd import sensitivity as S
    from cuwave_amd import utils as U

    print(f"{FAMILY} misfit gradient batch={batch}", flush=True)
    sim, gamma = misfit_case()
    rng = np.random.default_rng(31)
    sources = shots(sim, rng, 3)
    sensors = sensor_points(sim, rng, 4)
    observed = [0.5 * rng.standard_normal((sim.N, sensors.shape[1])) for _ in sources]
    cost, gradient = U.misfit_gradient(sim, sources, gamma, Receivers(sensors), observed,
                                       batch=batch)
    want_cost, want = 0.0, 0.0
    jm, js = sim.parametrization_jacobian(gamma)
    for source, obs in zip(sources, observed):
        c, g, _, _ = S.reconstruction_sensitivity(sim, source, gamma, sensors,
                                                  S.l2_misfit(obs))
        want_cost += c
        want = want + jm * g["mass"].cpu().numpy() + js * g["stiff"].cpu().numpy()
    assert abs(cost - want_cost) <= 1e-12 * want_cost
    got = gradient.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)


# ---------------------------------------------------------------- damping
def damped(sim, gamma, width=4, beta=0.2):
    import dataclasses

    from cuwave_amd import boundary

    return dataclasses.replace(sim, damping=boundary.sponge(sim, gamma, width, beta))


@pytest.mark.parametrize("Nx, faces, precision", [((22, 26), "mixed", "float64"),
                                                  ((12, 13, 14), "traction", "float64"),
                                                  ((22, 26), "clamped", "float32")])
def test_a_damped_march_is_the_oracles_bit_for_bit(Nx, faces, precision):
    print(f"anisotropic damped forward {Nx} {faces} {precision}", flush=True)
    rng = np.random.default_rng(43)
    sim = aniso_sim(Nx, faces, precision=precision)
    gamma = (0.5 + rng.random(sim.Nx_padded)).astype(sim.np_dtype)
    sim = damped(sim, gamma)
    sources = shots(sim, rng, 2)
    sensors = sensor_points(sim, rng)
    field, record = AN.simulate_shots(sim, sources, gamma, sensors)
    mat = sim.build_materials(gamma)
    for b, source in enumerate(sources):
        pos = np.asarray(source.position)
        weight = sim.excitation_weights(mat, sim.flatten_indices(pos))
        last, traces = O.simulate(logical(sim, mat["minv"]), logical(sim, mat["cell"]),
                                  mat["table"], sim.np_dtype(sim.dt**2), sim.N, tuple(pos), weight,
                                  np.asarray(source.signal, dtype=sim.np_dtype), tuple(sensors),
                                  logical(sim, sim.damping), sim.dt)
        assert np.array_equal(field[b].cpu().numpy(), last), f"shot {b}"
        assert np.array_equal(record[b].cpu().numpy(), traces), f"shot {b}"


def test_a_damped_gradient_matches_finite_differences():
    print("anisotropic damped FD", flush=True)
    Nx = (18, 20)
    sim = damped(aniso_sim(Nx, "mixed", N=80), np.ones((18, 32)), width=3, beta=0.3)
    rng = np.random.default_rng(44)
    gamma = 0.7 + 0.6 * rng.random(sim.Nx_padded)
    t = np.arange(sim.N) * sim.dt
    f0 = 4.0 / (sim.N * sim.dt)
    arg = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    pulse = (1 - 2 * arg) * np.exp(-arg)
    source = Source(np.array([[1], [9], [10]]), pulse[:, None])
    sensors = np.array([[0, 1, 1], [7, 11, 10], [8, 12, 11]])
    observed = AN.simulate(sim, source, np.ones(sim.Nx_padded), sensors)[1].cpu().numpy()

    def objective(um):
        r = um.cpu().numpy() - observed
        return 0.5 * float(np.sum(r * r)), torch.as_tensor(r)

    cost, grads, _, _ = AN.sensitivity(sim, source, gamma, sensors, objective)
    g = (grads["mass"] + grads["stiff"]).cpu().numpy()
    h = 1e-6
    for node in [(10, 11), (8, 9), (2, 2)]:
        costs = []
        for s in (1, -1):
            probe = gamma.copy()
            probe[node] += s * h
            costs.append(objective(AN.simulate(sim, source, probe, sensors)[1])[0])
        fd = (costs[0] - costs[1]) / (2 * h)
        assert abs(g[node] - fd) <= 1e-5 * abs(fd) + 1e-10 * abs(cost), (node, g[node], fd)


def test_the_source_gradient_matches_finite_differences():
    # cuwave tests/sensitivity_test.py source_sensitivity: the adjoint read at the source
    print(f"{FAMILY} source sensitivity", flush=True)
    from cuwave_amd import sensitivity as S
    from cuwave_amd.wave import family_drivers

    sim, gamma = misfit_case()
    rng = np.random.default_rng(51)
    source = shots(sim, rng, 1)[0]
    sensors = sensor_points(sim, rng, 4)
    observed = 0.5 * rng.standard_normal((sim.N, sensors.shape[1]))
    objective = S.l2_misfit(observed)
    cost, grad, _, _ = S.source_sensitivity(sim, source, gamma, sensors, objective)
    g = grad.cpu().numpy()
    h = 1e-4
    for t, k in [(3, 0), (sim.N // 2, 1), (sim.N - 5, 0)]:
        costs = []
        for sgn in (1, -1):
            sig = np.array(source.signal, dtype=np.float64)
            sig[t, k] += sgn * h
            um = family_drivers(sim).simulate(sim, Source(source.position, sig), gamma,
                                              sensors)[1]
            costs.append(objective(um)[0])
        fd = (costs[0] - costs[1]) / (2 * h)
        assert abs(g[t, k] - fd) <= 1e-6 * abs(fd) + 1e-9 * abs(cost), (t, k, g[t, k], fd)


def test_a_damped_misfit_gradient_sums_the_stored_gradients():
    # a damped field keeps its history: the batch against the single-shot sensitivity
    from helpers import Receivers

    from cuwave_amd import sensitivity as S
    from cuwave_amd import utils as U

    print(f"{FAMILY} damped misfit gradient", flush=True)
    sim, gamma = misfit_case()
    sim = damped(sim, gamma, width=3, beta=0.2)
    rng = np.random.default_rng(61)
    sources = shots(sim, rng, 3)
    sensors = sensor_points(sim, rng, 4)
    observed = [0.5 * rng.standard_normal((sim.N, sensors.shape[1])) for _ in sources]
    cost, gradient = U.misfit_gradient(sim, sources, gamma, Receivers(sensors), observed)
    want_cost, want = 0.0, 0.0
    jm, js = sim.parametrization_jacobian(gamma)
    for source, obs in zip(sources, observed):
        c, g, _, _ = S.sensitivity(sim, source, gamma, sensors, S.l2_misfit(obs))
        want_cost += c
        want = want + jm * g["mass"].cpu().numpy() + js * g["stiff"].cpu().numpy()
    assert abs(cost - want_cost) <= 1e-12 * want_cost
    got = gradient.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)
