"""The vector Maxwell kernels against the oracle on the GPU.

- the forward march, several shots in one batch, bit for bit against the
  oracle's one-shot march: 2D and 3D, orders 2 to 8, conductor, magnetic and
  mixed walls, float64 and float32;
- the fdtdx dipole cases through the kernels, at the gaps the reference's
  tests hold (cuwave tests/maxwell_test.py:968-978);
- one adjoint step with its inertia density, bit for bit against the
  oracle's GRADIENT step;
- `sensitivity` against central finite differences of the misfit
  (cuwave tests/maxwell_test.py:839-846), `reconstruction_sensitivity` (the
  field rebuilt by the reverse march) against it.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import ROOT, magnetodielectric_class

from cuwave_amd import _abi
from cuwave_amd import maxwell as MX
from cuwave_amd.boundary import Conductor, Magnetic
from cuwave_amd.stencils import staggered_weights
from cuwave_amd.wave import Source
from oracle import maxwell as M

pytestmark = pytest.mark.gpu

FAMILY = "maxwell"
WALLS = {"conductor": None, "magnetic": Magnetic,
         "mixed": lambda n: ((Conductor, Magnetic),) + ((Magnetic, Conductor),) * (n - 1)}


@pytest.fixture(scope="module", autouse=True)
def library():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    _abi.load()


def vector_sim(Nx, order, walls="conductor", precision="float64", N=30, magnetic=False):
    boundary = WALLS[walls]
    if callable(boundary):
        boundary = boundary(len(Nx))
    dx = tuple(1.0 / (n - 3) for n in Nx)
    cls = magnetodielectric_class() if magnetic else MX.DielectricWave
    return cls(Nx, dx, N, 0.25 * min(dx), precision=precision, space_order=order,
               permittivity1=1.0, permittivity2=4.0, boundary=boundary)


def misfit_case():
    sim = vector_sim((16, 17, 18), 4, "mixed", N=40, magnetic=True)
    return sim, np.random.default_rng(30).random(sim.Nx_padded)


def logical(sim, a):
    return a[(..., *(slice(0, n) for n in sim.Nx))]


def oracle_march(sim, gamma, source, sensors):
    """The oracle's one-shot march of `source` with the product's weights' arithmetic."""
    dt = sim.np_dtype
    eps, nu = sim.parametrization(gamma)
    eps = logical(sim, eps)
    nu_pair = M.pair_inverse_permeability(logical(sim, nu)) if sim.magnetic else None
    grid = sim.abi_grid(1)
    minv = M.point_inverse_inertia(eps, grid.conductor)
    pos = np.asarray(source.position)
    volume = dt(1.0 / float(np.prod(sim.dx)))
    weight = (dt(sim.dt**2) * (minv[tuple(pos)] * volume)).astype(dt)
    sens = np.asarray(sensors)
    return M.simulate(minv, staggered_weights(sim.radius, sim.stencil).astype(dt),
                      [dt(1.0 / d) for d in sim.dx], sim.step_factors(), dt(1.0), grid.conductor,
                      sim.N, tuple(pos), weight, np.asarray(source.signal, dtype=dt), tuple(sens),
                      nu_pair=nu_pair)


def shots(sim, rng, count):
    """`count` shots of two sources each on random component points of the interior."""
    out = []
    for _ in range(count):
        pos = [rng.integers(0, sim.ncomp, 2)]
        for n in sim.Nx:
            pos.append(rng.integers(2, n - 3, 2))
        out.append(Source(np.array(pos), rng.standard_normal((sim.N, 2)).astype(sim.np_dtype)))
    return out


def sensor_points(sim, rng, count=5):
    pos = [rng.integers(0, sim.ncomp, count)] + [rng.integers(1, n - 2, count) for n in sim.Nx]
    return np.array(pos)


@pytest.mark.parametrize("Nx", [(22, 26), (14, 16, 18)])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("walls", sorted(WALLS))
def test_the_forward_march_is_the_oracles_bit_for_bit(Nx, order, walls):
    print(f"maxwell forward {Nx} order {order} {walls}", flush=True)
    sim = vector_sim(Nx, order, walls)
    rng = np.random.default_rng(order + len(Nx))
    gamma = rng.random(sim.Nx_padded)
    sources = shots(sim, rng, 3)
    sensors = sensor_points(sim, rng)
    field, record = MX.simulate_shots(sim, sources, gamma, sensors)
    for b, source in enumerate(sources):
        last, traces = oracle_march(sim, gamma, source, sensors)
        assert np.array_equal(field[b].cpu().numpy(), last), f"shot {b}"
        assert np.array_equal(record[b].cpu().numpy(), traces), f"shot {b}"


@pytest.mark.parametrize("Nx", [(22, 26), (14, 16, 18)])
def test_the_float32_march_is_the_oracles_bit_for_bit(Nx):
    print(f"maxwell float32 {Nx}", flush=True)
    sim = vector_sim(Nx, 4, "mixed", precision="float32")
    rng = np.random.default_rng(7)
    gamma = rng.random(sim.Nx_padded).astype(np.float32)
    sources = shots(sim, rng, 5)
    sensors = sensor_points(sim, rng)
    field, record = MX.simulate_shots(sim, sources, gamma, sensors)
    for b, source in enumerate(sources):
        last, traces = oracle_march(sim, gamma, source, sensors)
        assert np.array_equal(field[b].cpu().numpy(), last), f"shot {b}"
        assert np.array_equal(record[b].cpu().numpy(), traces), f"shot {b}"


@pytest.mark.parametrize("name, bound", [("vector3D_vacuum", 1e-12), ("vector3D_block", 1e-3)])
def test_the_kernels_reproduce_fdtdx(name, bound):
    print(f"maxwell fdtdx {name}", flush=True)
    ref = np.load(ROOT / "tests" / "golden" / f"fdtdx_{name}.npz")
    shape = tuple(int(n) + 2 for n in ref["shape"])
    S = float(ref["courant"])
    sim = MX.DielectricWave(shape, (1.0,) * 3, len(ref["amplitude"]), S, precision="float64",
                            permittivity1=1.0, permittivity2=float(ref["permittivity"]),
                            boundary=Conductor)
    gamma = np.zeros(sim.Nx_padded)
    if ref["block_low"].size:
        low, size = ref["block_low"] + 1, ref["block_size"]
        gamma[tuple(slice(lo, lo + n) for lo, n in zip(low, size))] = 1.0
    source = np.array([[2], *([s + 1] for s in ref["source"])])
    mat = sim.build_materials(gamma)
    weight = sim.excitation_weights(mat, sim.flatten_indices(source))
    loads = np.diff(-S * ref["amplitude"], prepend=0.0)
    probes = ref["probes"].T + 1
    sensors = np.concatenate([np.full((1, probes.shape[1]), 2), probes])
    _, traces = MX.simulate(sim, Source(source, (loads / weight)[:, None]), gamma, sensors)
    theirs = ref["traces"]
    assert np.linalg.norm(traces.cpu().numpy() - theirs) / np.linalg.norm(theirs) < bound


@pytest.mark.parametrize("Nx, order", [((20, 23), 6), ((11, 12, 14), 4)])
def test_the_adjoint_step_is_the_oracles_gradient_step(Nx, order):
    print(f"maxwell adjoint step {Nx} order {order}", flush=True)
    sim = vector_sim(Nx, order, "mixed")
    rng = np.random.default_rng(11)
    mat = sim.build_materials(rng.random(sim.Nx_padded))
    eng = MX.MaxwellEngine(sim, mat, 1)
    levels = [rng.standard_normal((sim.ncomp, *sim.Nx_padded, 1)) for _ in range(3)]
    free = np.zeros(levels[0].shape, dtype=bool)
    free[(slice(None), *(slice(1, n - 1) for n in sim.Nx))] = True
    l0, l1, v1 = (torch.as_tensor(np.where(free, x, 0.0), device="cuda") for x in levels)
    g0 = rng.standard_normal((sim.ncomp, *sim.Nx_padded))
    g = torch.as_tensor(g0, device="cuda")
    mf = 1.0 / sim.dt**2
    out = eng.step(l0.clone(), l1, l0.clone(), g, v1, mf)
    d = sim.np_dtype
    g_want = logical(sim, g0).copy()
    want = M.step(logical(sim, l0.cpu().numpy()[..., 0]), logical(sim, l1.cpu().numpy()[..., 0]),
                  logical(sim, mat["minv"]), staggered_weights(sim.radius, sim.stencil).astype(d),
                  [d(1.0 / x) for x in sim.dx], sim.step_factors(), 1.0,
                  sim.abi_grid(1).conductor, v1=logical(sim, v1.cpu().numpy()[..., 0]),
                  g_mass=g_want, mf=mf)
    got = logical(sim, out.cpu().numpy()[..., 0])
    valid = np.zeros(want.shape, dtype=bool)
    for c in range(sim.ncomp):
        valid[(c, *(slice(1, n - 1 - (k == c)) for k, n in enumerate(sim.Nx)))] = True
    assert np.array_equal(np.where(valid, got, 0.0), want)
    assert np.array_equal(logical(sim, g.cpu().numpy()), g_want)


@pytest.mark.parametrize("Nx, order", [((16, 18), 4), ((10, 11, 12), 2)])
def test_the_gradient_matches_finite_differences(Nx, order):
    print(f"maxwell gradient {Nx} order {order}", flush=True)
    sim = vector_sim(Nx, order, "mixed", N=60)
    rng = np.random.default_rng(5)
    gamma = 0.3 + 0.4 * rng.random(sim.Nx_padded)
    t = np.arange(sim.N) * sim.dt
    f0 = 4.0 / (sim.N * sim.dt)
    arg = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    pulse = (1 - 2 * arg) * np.exp(-arg)
    centre = [n // 2 for n in Nx]
    source = Source(np.array([[1], *([c] for c in centre)]), pulse[:, None])
    sensors = np.array([[0, 1, sim.ncomp - 1], *([c - 2, c + 2, c + 1] for c in centre)])
    observed = MX.simulate(sim, source, 0.5 * np.ones(sim.Nx_padded), sensors)[1].cpu().numpy()

    def objective(um):
        r = um.cpu().numpy() - observed
        return 0.5 * float(np.sum(r * r)), torch.as_tensor(r)

    cost, grads, _, _ = MX.sensitivity(sim, source, gamma, sensors, objective)
    d_mass, _ = sim.parametrization_jacobian()
    g = (d_mass * grads["mass"]).cpu().numpy()
    h = 1e-5
    for node in [tuple(c + 1 for c in centre), tuple(c - 1 for c in centre),
                 tuple(min(c + 3, n - 2) for c, n in zip(centre, Nx))]:
        costs = []
        for s in (1, -1):
            probe = gamma.copy()
            probe[node] += s * h
            um = MX.simulate(sim, source, probe, sensors)[1]
            costs.append(objective(um)[0])
        fd = (costs[0] - costs[1]) / (2 * h)
        assert abs(g[node] - fd) <= 1e-5 * abs(fd) + 1e-10 * abs(cost), (node, g[node], fd)


def test_the_top_level_drivers_reach_the_family():
    from cuwave_amd import sensitivity as S
    from cuwave_amd import wave as W

    print("maxwell top-level drivers", flush=True)
    sim = vector_sim((12, 13, 14), 4, "mixed", N=25)
    rng = np.random.default_rng(9)
    gamma = rng.random(sim.Nx_padded)
    source = shots(sim, rng, 1)[0]
    sensors = sensor_points(sim, rng, 3)
    mine = MX.simulate(sim, source, gamma, sensors)
    top = W.simulate(sim, source, gamma, sensors)
    assert all(torch.equal(a, b) for a, b in zip(mine, top))

    def objective(um):
        return 0.5 * float(torch.sum(um * um)), um

    _, g1, _, _ = MX.sensitivity(sim, source, gamma, sensors, objective)
    _, g2, _, _ = S.sensitivity(sim, source, gamma, sensors, objective)
    assert torch.equal(g1["mass"], g2["mass"])


def test_the_measured_timestep_sits_at_the_yee_limit():
    # a uniform vacuum at order 2: c dt <= dx / sqrt(ndim) on the infinite lattice,
    # which a closed box approaches from above
    print("maxwell stable timestep", flush=True)
    from cuwave_amd.wave import stable_timestep

    n = 24
    sim = MX.DielectricWave((n,) * 3, (1.0 / n,) * 3, 10, 0.1 / n, precision="float64",
                            permittivity1=1.0, permittivity2=4.0, boundary=Magnetic)
    dt = stable_timestep(sim, np.zeros(sim.Nx_padded), iterations=200, safety=1.0)
    limit = (1.0 / n) / np.sqrt(3.0)
    assert limit <= dt <= 1.1 * limit, (dt, limit)


@pytest.mark.parametrize("Nx, order, precision, bound",
                         [((16, 18), 4, "float64", 1e-9), ((10, 11, 12), 6, "float64", 1e-9),
                          ((14, 15, 16), 2, "float32", 2e-3)])
def test_the_reverse_march_reproduces_the_stored_gradient(Nx, order, precision, bound):
    # no pml: the leapfrog runs backwards exactly but for rounding
    # (cuwave tests/maxwell_test.py:844-846)
    print(f"maxwell reconstruction {Nx} order {order} {precision}", flush=True)
    from cuwave_amd import sensitivity as S

    sim = vector_sim(Nx, order, "mixed", precision=precision, N=60)
    rng = np.random.default_rng(12)
    gamma = rng.random(sim.Nx_padded).astype(sim.np_dtype)
    source = shots(sim, rng, 1)[0]
    sensors = sensor_points(sim, rng, 4)

    def objective(um):
        return 0.5 * float(torch.sum(um * um)), um

    c1, g1, u1, _ = MX.sensitivity(sim, source, gamma, sensors, objective)
    c2, g2, u2, info = S.reconstruction_sensitivity(sim, source, gamma, sensors, objective)
    assert c1 == c2 and torch.equal(u1, u2)
    gap = float(torch.linalg.vector_norm(g2["mass"] - g1["mass"]) /
                torch.linalg.vector_norm(g1["mass"]))
    assert gap < bound, gap
    assert info["strip"] == 0 and info["drift"] < (1e-10 if precision == "float64" else 1e-3)


# ------------------------------------------------------ designed permeability
@pytest.mark.parametrize("Nx, order, walls", [((22, 26), 4, "mixed"), ((14, 16, 18), 2, "magnetic")])
def test_a_designed_permeability_march_is_the_oracles_bit_for_bit(Nx, order, walls):
    print(f"maxwell magnetic forward {Nx} order {order} {walls}", flush=True)
    sim = vector_sim(Nx, order, walls, magnetic=True)
    rng = np.random.default_rng(21)
    gamma = rng.random(sim.Nx_padded)
    sources = shots(sim, rng, 2)
    sensors = sensor_points(sim, rng)
    field, record = MX.simulate_shots(sim, sources, gamma, sensors)
    for b, source in enumerate(sources):
        last, traces = oracle_march(sim, gamma, source, sensors)
        assert np.array_equal(field[b].cpu().numpy(), last), f"shot {b}"
        assert np.array_equal(record[b].cpu().numpy(), traces), f"shot {b}"


@pytest.mark.parametrize("Nx, order", [((20, 23), 4), ((11, 12, 14), 2)])
def test_the_permeability_density_is_the_oracles(Nx, order):
    print(f"maxwell magnetic adjoint step {Nx} order {order}", flush=True)
    sim = vector_sim(Nx, order, "mixed", magnetic=True)
    rng = np.random.default_rng(22)
    gamma = rng.random(sim.Nx_padded)
    mat = sim.build_materials(gamma)
    eng = MX.MaxwellEngine(sim, mat, 1)
    free = np.zeros((sim.ncomp, *sim.Nx_padded, 1), dtype=bool)
    free[(slice(None), *(slice(1, n - 1) for n in sim.Nx))] = True
    l0, l1, v1 = (torch.as_tensor(np.where(free, rng.standard_normal(free.shape), 0.0),
                                  device="cuda") for _ in range(3))
    grads = eng.gradient_fields()
    g0 = rng.standard_normal(grads["nu"].shape)
    grads["nu"].copy_(torch.as_tensor(g0))
    mf = 1.0 / sim.dt**2
    out = eng.adjoint_step(l0.clone(), l1, grads, v1, mf)
    d = sim.np_dtype
    g_mass = np.zeros((sim.ncomp, *sim.Nx))
    g_nu = logical(sim, g0).copy()
    want = M.step(logical(sim, l0.cpu().numpy()[..., 0]), logical(sim, l1.cpu().numpy()[..., 0]),
                  logical(sim, mat["minv"]), staggered_weights(sim.radius, sim.stencil).astype(d),
                  [d(1.0 / x) for x in sim.dx], sim.step_factors(), 1.0,
                  sim.abi_grid(1).conductor, v1=logical(sim, v1.cpu().numpy()[..., 0]),
                  g_mass=g_mass, mf=mf, nu_pair=logical(sim, mat["nu_pair"]), g_nu=g_nu)
    valid = np.zeros(want.shape, dtype=bool)
    for c in range(sim.ncomp):
        valid[(c, *(slice(1, n - 1 - (k == c)) for k, n in enumerate(sim.Nx)))] = True
    assert np.array_equal(np.where(valid, logical(sim, out.cpu().numpy()[..., 0]), 0.0), want)
    assert np.array_equal(logical(sim, grads["mass"].cpu().numpy()), g_mass)
    assert np.array_equal(logical(sim, grads["nu"].cpu().numpy()), g_nu)


@pytest.mark.parametrize("Nx, order", [((16, 18), 4), ((10, 11, 12), 2)])
def test_a_designed_permeability_matches_finite_differences(Nx, order):
    # cuwave tests/maxwell_test.py:847-851
    print(f"maxwell magnetic FD {Nx} order {order}", flush=True)
    from cuwave_amd import sensitivity as S

    sim = vector_sim(Nx, order, "mixed", N=60, magnetic=True)
    rng = np.random.default_rng(23)
    gamma = 0.3 + 0.4 * rng.random(sim.Nx_padded)
    t = np.arange(sim.N) * sim.dt
    f0 = 4.0 / (sim.N * sim.dt)
    arg = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    pulse = (1 - 2 * arg) * np.exp(-arg)
    centre = [n // 2 for n in Nx]
    source = Source(np.array([[1], *([c] for c in centre)]), pulse[:, None])
    sensors = np.array([[0, 1, sim.ncomp - 1], *([c - 2, c + 2, c + 1] for c in centre)])
    observed = MX.simulate(sim, source, 0.5 * np.ones(sim.Nx_padded), sensors)[1].cpu().numpy()

    def objective(um):
        r = um.cpu().numpy() - observed
        return 0.5 * float(np.sum(r * r)), torch.as_tensor(r)

    cost, grads, _, _ = MX.sensitivity(sim, source, gamma, sensors, objective)
    d_mass, d_stiff = sim.parametrization_jacobian(gamma)
    g = d_mass * grads["mass"].cpu().numpy() + d_stiff * grads["stiff"].cpu().numpy()
    assert np.abs(grads["stiff"].cpu().numpy()).max() > 0
    h = 1e-5
    for node in [tuple(c + 1 for c in centre), tuple(c - 1 for c in centre),
                 tuple(min(c + 3, n - 2) for c, n in zip(centre, Nx))]:
        costs = []
        for sgn in (1, -1):
            probe = gamma.copy()
            probe[node] += sgn * h
            costs.append(objective(MX.simulate(sim, source, probe, sensors)[1])[0])
        fd = (costs[0] - costs[1]) / (2 * h)
        assert abs(g[node] - fd) <= 1e-5 * abs(fd) + 1e-10 * abs(cost), (node, g[node], fd)
    _, g2, _, info = S.reconstruction_sensitivity(sim, source, gamma, sensors, objective)
    for k in ("mass", "stiff"):
        gap = float(torch.linalg.vector_norm(g2[k] - grads[k]) /
                    torch.linalg.vector_norm(grads[k]))
        assert gap < 1e-9, (k, gap)


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


@pytest.mark.parametrize("Nx, order, precision", [((22, 26), 4, "float64"),
                                                  ((14, 16, 18), 2, "float64"),
                                                  ((22, 26), 6, "float32")])
def test_a_damped_march_is_the_oracles_bit_for_bit(Nx, order, precision):
    print(f"maxwell damped forward {Nx} order {order} {precision}", flush=True)
    rng = np.random.default_rng(41)
    sim = vector_sim(Nx, order, "mixed", precision=precision)
    gamma = rng.random(sim.Nx_padded).astype(sim.np_dtype)
    sim = damped(sim, gamma)
    sources = shots(sim, rng, 2)
    sensors = sensor_points(sim, rng)
    field, record = MX.simulate_shots(sim, sources, gamma, sensors)
    mat = sim.build_materials(gamma)
    for b, source in enumerate(sources):
        pos = np.asarray(source.position)
        weight = sim.excitation_weights(mat, sim.flatten_indices(pos))
        d = sim.np_dtype
        last, traces = M.simulate(
            logical(sim, mat["minv"]), staggered_weights(sim.radius, sim.stencil).astype(d),
            [d(1.0 / x) for x in sim.dx], sim.step_factors(), d(1.0), sim.abi_grid(1).conductor,
            sim.N, tuple(pos), weight, np.asarray(source.signal, dtype=d), tuple(sensors),
            damping=logical(sim, sim.damping), dt_=sim.dt)
        assert np.array_equal(field[b].cpu().numpy(), last), f"shot {b}"
        assert np.array_equal(record[b].cpu().numpy(), traces), f"shot {b}"


@pytest.mark.parametrize("Nx, order", [((16, 18), 4), ((10, 11, 12), 2)])
def test_a_damped_gradient_matches_finite_differences(Nx, order):
    print(f"maxwell damped FD {Nx} order {order}", flush=True)
    sim = vector_sim(Nx, order, "mixed", N=60)
    rng = np.random.default_rng(42)
    gamma = 0.3 + 0.4 * rng.random(sim.Nx_padded)
    sim = damped(sim, 0.5 * np.ones(sim.Nx_padded), width=3, beta=0.3)
    t = np.arange(sim.N) * sim.dt
    f0 = 4.0 / (sim.N * sim.dt)
    arg = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    pulse = (1 - 2 * arg) * np.exp(-arg)
    centre = [n // 2 for n in Nx]
    source = Source(np.array([[1], *([c] for c in centre)]), pulse[:, None])
    sensors = np.array([[0, 1, sim.ncomp - 1], *([c - 2, c + 2, c + 1] for c in centre)])
    observed = MX.simulate(sim, source, 0.5 * np.ones(sim.Nx_padded), sensors)[1].cpu().numpy()

    def objective(um):
        r = um.cpu().numpy() - observed
        return 0.5 * float(np.sum(r * r)), torch.as_tensor(r)

    cost, grads, _, _ = MX.sensitivity(sim, source, gamma, sensors, objective)
    d_mass, _ = sim.parametrization_jacobian()
    g = d_mass * grads["mass"].cpu().numpy()
    h = 1e-5
    for node in [tuple(c + 1 for c in centre), tuple(c - 1 for c in centre), (2,) * len(Nx)]:
        costs = []
        for sgn in (1, -1):
            probe = gamma.copy()
            probe[node] += sgn * h
            costs.append(objective(MX.simulate(sim, source, probe, sensors)[1])[0])
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
