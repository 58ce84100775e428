"""Domain masks through the HIP kernels, against the oracle and the mathematics.

A `domain` mask takes nodes out of the stepped set and `dirichlet` holds nodes at
zero (cuwave/wave.py:661-726); the masked step grades the cells at their walls
(cuwave/kernels/scalar.cu:85-158, 434-452). Every masked sweep runs the general
kernels with the per-node cell codes.

- the masked forward is bit for bit the oracle's: a disk in 2D at orders 4 and 8
  with a held obstacle, a held row, a ball in 3D, a 1D segment with a held node,
  a sponge-damped disk with Dirichlet faces, three shots in one batch, a larger
  2D grid;
- a held row is the method of images (tests/wave_test.py:222-247 of the reference);
- the illumination of a masked sweep is the oracle's, zero off the stepped set;
- the material off the stepped set is never read: voided or filled there, the
  bits are the same (wave_test.py:248-254);
- a domain covering everything is the plain sweep bit for bit, with the plain
  sweep's core blocks marched by the stream kernels, and its measured timestep
  is the plain one;
- at order 2 `source_sensitivity` through a disk matches central differences:
  its recursion is the forward step itself (cuwave/scalar.py:321-323;
  wave_test.py:262-280);
- a graphed march of a masked grid is bitwise one plain call;
- the adjoint entry points refuse a mask at the library.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import host, rel_l2

pytestmark = pytest.mark.gpu

from oracle import driver as D  # noqa: E402

from cuwave_amd import _abi, boundary, scalar, wave  # noqa: E402
from cuwave_amd import sensitivity as S  # noqa: E402
from cuwave_amd.engine import Propagator, sensor_points, source_points  # noqa: E402


def _ball(Nx, dx, r=0.42, centre=0.5):
    X = wave.grid_coords(Nx, dx)
    return sum((x - centre) ** 2 for x in X) <= r**2


def _obstacle(Nx, dx):
    X = wave.grid_coords(Nx, dx)
    return (X[0] - 0.62) ** 2 + (X[1] - 0.35) ** 2 <= 0.1**2


def _faces(sim):
    return tuple(tuple(D.DIRICHLET if c is boundary.Dirichlet else D.NEUMANN for c in pair)
                 for pair in sim.boundary)


def _pair(Nx, order=4, precision="float32", N=200, domain="ball", held=None, bnd=None,
          damped=False, seed=0):
    dx = tuple(1.0 / (n - 3) for n in Nx)
    dt = 0.5 * wave.stable_dt(dx, 1.0, order)
    shape = wave.padded_shape(Nx)
    dom = _ball(Nx, dx) if isinstance(domain, str) else domain
    common = dict(precision=precision, space_order=order, rho1=1.0, rho2=2.5, kappa1=1.0,
                  kappa2=4.0, boundary=bnd)
    base = scalar.AcousticWave(Nx, dx, N, dt, **common)
    rng = np.random.default_rng(seed)
    ind = (0.2 + 0.6 * rng.random(shape)).astype(base.np_dtype)
    damping = boundary.sponge(base, ind, 6, 0.05) if damped else None
    sim = scalar.AcousticWave(Nx, dx, N, dt, damping=damping, domain=dom, dirichlet=held,
                              **common)
    osim = D.Sim("AcousticWave", Nx, dx, N, dt, precision, order, boundary=_faces(sim),
                 materials=(1.0, 2.5, 1.0, 4.0), damping=damping, domain=dom, dirichlet=held)
    return sim, osim, ind


def _inside_points(sim, count, seed):
    """`count` stepped interior nodes, as (ndim, count) rows."""
    stepped = wave.stepped_nodes(sim)
    inner = np.zeros_like(stepped)
    inner[tuple(slice(3, n - 3) for n in sim.Nx)] = True
    nodes = np.argwhere(stepped & inner)
    pick = np.random.default_rng(seed).choice(len(nodes), count, replace=False)
    return nodes[pick].T.copy()


def _signal(sim, columns=1):
    t = np.arange(sim.N)
    return np.stack([np.sin(0.3 * (k + 1) * t) * np.exp(-t / 60.0) for k in range(columns)],
                    1).astype(sim.np_dtype)


def _held_row(Nx, row=20):
    held = np.zeros(wave.padded_shape(Nx), dtype=bool)
    held[row] = True
    return held


CASES = {
    "disk o4 f32": dict(Nx=(40, 48)),
    "disk and obstacle o8 f32": dict(Nx=(40, 48), order=8, held="obstacle"),
    "held row o2 f64": dict(Nx=(40, 36), order=2, precision="float64", domain=None,
                            held="row"),
    "ball o6 f64 3d": dict(Nx=(22, 21, 20), order=6, precision="float64", N=80),
    "segment and held node o8 f32 1d": dict(Nx=(60,), order=8, domain="segment", held="node"),
    "damped disk, dirichlet faces, o4 f64": dict(
        Nx=(40, 48), precision="float64", damped=True,
        bnd=((boundary.Dirichlet, boundary.Neumann), (boundary.Neumann, boundary.Dirichlet))),
    "larger disk o8 f32": dict(Nx=(200, 230), order=8, N=120, held="obstacle"),
}


def _case(name):
    spec = dict(CASES[name])
    Nx = spec["Nx"]
    dx = tuple(1.0 / (n - 3) for n in Nx)
    shape = wave.padded_shape(Nx)
    if spec.get("domain") == "segment":
        dom = np.zeros(shape, dtype=bool)
        dom[5:50] = True
        spec["domain"] = dom
    held = spec.pop("held", None)
    if held == "obstacle":
        spec["held"] = _obstacle(Nx, dx)
    elif held == "row":
        spec["held"] = _held_row(Nx)
    elif held == "node":
        spec["held"] = np.zeros(shape, dtype=bool)
        spec["held"][30] = True
    return _pair(**spec)


@pytest.mark.parametrize("name", list(CASES))
def test_the_masked_forward_is_the_oracles_bit_for_bit(name):
    sim, osim, ind = _case(name)
    print(f"case masked forward {name}", flush=True)
    src = _inside_points(sim, 1, 0)
    sensors = _inside_points(sim, 4, 1)
    sig = _signal(sim)
    field, traces = wave.simulate(sim, wave.Source(src, sig), ind, sensors)
    ofield, otraces = D.simulate(osim, src, sig, ind, sensors)
    assert np.abs(otraces).max() > 0
    np.testing.assert_array_equal(host(field), ofield)
    np.testing.assert_array_equal(host(traces), otraces)


def test_three_masked_shots_in_one_batch_are_the_oracles():
    sim, osim, ind = _case("disk and obstacle o8 f32")
    print("case masked forward three shots", flush=True)
    sensors = _inside_points(sim, 4, 1)
    sources = [wave.Source(_inside_points(sim, 1 + k % 2, 10 + k), _signal(sim, 1 + k % 2))
               for k in range(3)]
    fields, traces = wave.simulate_shots(sim, sources, ind, sensors)
    for k, source in enumerate(sources):
        ofield, otraces = D.simulate(osim, source.position, source.signal, ind, sensors)
        np.testing.assert_array_equal(host(fields[k]), ofield)
        np.testing.assert_array_equal(host(traces[k]), otraces)


def test_a_held_row_is_the_method_of_images():
    # the box mirrors about row 19.5, the midpoint the held row 20 puts its zero at
    dx = (0.02, 0.02)
    common = dict(precision="float64", wavespeed=1.0, density=1.0)
    sim = scalar.ScalarWave((40, 36), dx, 150, 0.9 * wave.stable_dt(dx, 1.0), **common)
    print("case method of images", flush=True)
    signal = np.sin(0.3 * np.arange(sim.N))
    ones = np.ones(sim.Nx_padded)
    imaged = host(wave.simulate(sim, wave.Source(np.array([[12, 27], [15, 15]]),
                                                 np.stack([signal, -signal], 1)), ones))
    held = scalar.ScalarWave((40, 36), dx, 150, sim.dt, dirichlet=_held_row((40, 36)), **common)
    one = host(wave.simulate(held, wave.Source(np.array([[12], [15]]), signal[:, None]), ones))
    below = (slice(1, 20), slice(1, 35))
    scale = float(np.abs(imaged[below]).max())
    assert scale > 0
    np.testing.assert_allclose(one[below], imaged[below], atol=1e-12 * scale)


def test_the_illumination_of_a_masked_sweep_is_the_oracles():
    sim, osim, ind = _case("disk and obstacle o8 f32")
    print("case masked illumination", flush=True)
    shots = [(_inside_points(sim, 1, k), _signal(sim) * (1 + k)) for k in range(2)]
    expect = D.illumination(osim, shots, ind)
    got = wave.illumination(sim, [wave.Source(p, s) for p, s in shots], ind)
    assert rel_l2(got, expect) < 1e-5
    assert (host(got)[~wave.stepped_nodes(sim)] == 0).all()


def test_the_material_off_the_domain_is_never_read():
    sim, _, ind = _case("disk and obstacle o8 f32")
    print("case material outside", flush=True)
    src, sensors, sig = _inside_points(sim, 1, 0), _inside_points(sim, 4, 1), _signal(sim)
    stepped = wave.stepped_nodes(sim)
    voided = np.where(stepped, ind, sim.np_dtype(1e-4))
    filled = np.where(stepped, ind, sim.np_dtype(1.0))
    a = wave.simulate(sim, wave.Source(src, sig), voided, sensors)
    b = wave.simulate(sim, wave.Source(src, sig), filled, sensors)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.fixture
def _march_every_box():
    previous = _abi.call_value("cw_set_stream_threshold", 0)
    yield
    _abi.call_value("cw_set_stream_threshold", previous)


def test_a_domain_covering_everything_is_the_plain_sweep(_march_every_box):
    Nx = (90, 100)
    dx = tuple(1.0 / (n - 3) for n in Nx)
    common = dict(precision="float32", space_order=8, wavespeed=1.0, density=1.0)
    dt = 0.5 * wave.stable_dt(dx, 1.0, 8)
    plain = scalar.ScalarWave(Nx, dx, 120, dt, **common)
    covered = scalar.ScalarWave(Nx, dx, 120, dt, domain=np.ones(plain.Nx_padded, dtype=bool),
                                **common)
    print("case covering domain", flush=True)
    ind = (1.0 + 0.2 * np.random.default_rng(0).random(plain.Nx_padded)).astype(np.float32)
    src, sensors, sig = np.array([[45], [50]]), np.array([[10, 80], [12, 90]]), _signal(plain)
    a = wave.simulate(plain, wave.Source(src, sig), ind, sensors)
    b = wave.simulate(covered, wave.Source(src, sig), ind, sensors)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    assert wave.stable_timestep(covered, ind, iterations=20) == \
        wave.stable_timestep(plain, ind, iterations=20)


def test_source_sensitivity_through_a_disk_matches_finite_differences():
    Nx = (40, 48)
    dx = tuple(1.0 / (n - 3) for n in Nx)
    sim = scalar.ScalarWave(Nx, dx, 160, 0.5 * wave.stable_dt(dx, 1.0, 2), precision="float64",
                            wavespeed=1.0, density=1.0, domain=_ball(Nx, dx),
                            dirichlet=_obstacle(Nx, dx))
    print("case masked source sensitivity", flush=True)
    ind = 1.0 + 0.2 * np.random.default_rng(0).random(sim.Nx_padded)
    source = wave.Source(_inside_points(sim, 2, 0), _signal(sim, 2))
    sensors = _inside_points(sim, 4, 1)
    objective = S.l2_misfit(np.zeros((sim.N, 4)))
    cost, gradient, _, _ = S.source_sensitivity(sim, source, ind, sensors, objective)
    gradient = host(gradient)
    direction = np.random.default_rng(1).standard_normal(source.signal.shape)
    h = 1e-3

    def cost_at(signal):
        traces = wave.simulate(sim, wave.Source(source.position, signal), ind, sensors)[1]
        return objective(traces)[0]

    fd = (cost_at(source.signal + h * direction) - cost_at(source.signal - h * direction)) / (2 * h)
    assert abs(float(np.sum(gradient * direction)) / fd - 1) < 1e-8


def test_a_graphed_masked_march_is_one_plain_call():
    sim, _, ind = _case("disk and obstacle o8 f32")
    print("case graphed masked march", flush=True)
    mat = sim.build_materials(ind)
    runs = []
    for window in (0, 24, 0):
        prop = Propagator(sim, mat, 1)
        src, signal = source_points(prop, mat, [wave.Source(_inside_points(sim, 1, 0),
                                                            _signal(sim))])
        rec = sensor_points(prop, _inside_points(sim, 4, 1))
        record = prop.record_memory(rec)
        U = prop.levels_memory(2)
        kw = dict(sources=src, signal=signal, sensors=rec, record=record)
        if window:
            prop.march(U, 0, sim.N, window, **kw)
        else:
            prop.forward(U, 0, sim.N, **kw)
        torch.cuda.synchronize()
        runs.append((U.clone(), record.clone()))
    for a, b in zip(runs[0], runs[1]):
        assert torch.equal(a, b)
    for a, b in zip(runs[0], runs[2]):
        assert torch.equal(a, b)


def test_the_adjoint_entry_points_refuse_a_mask():
    sim, _, ind = _case("disk o4 f32")
    mat = sim.build_materials(ind)
    prop = Propagator(sim, mat, 1)
    P, U = prop.levels_memory(2), prop.levels_memory(2)
    with pytest.raises(RuntimeError, match="forward only"):
        prop.backward(P, U, 0, 1, eta=prop.eta_memory(), g_mass=prop.accumulator(),
                      combine=(0.0, 0.0))
