"""3D forward sweeps that run the 3D core kernel, against the oracle.

A 3D forward sweep of one shot lane marches the box of nodes off the walls
and the PML with the core kernel (planes staged in LDS, a register queue down
the slow axis) and leaves the shell around it to the general kernels. Boxes
this small hold fewer blocks than the library marches by default, so the tests
lower that threshold to 0, and every case is also run with the threshold out of
reach (the general kernels alone): both must give the oracle's bits.

- the two 3D golden forwards (orders 4 and 8, a CPML on every face);
- a larger acoustic grid at order 8 in float32 with shots and sensors inside
  the box and in the shell, a plane of sensors across the box, and the
  illumination map;
- order 2 in float64 on extents that are no multiple of the block's;
- a graphed march of a grid with a core is bitwise one plain call;
- the adjoint levels with a core (lambda and the reverse march marched the same
  way) give the general kernels' bits: the stored-history and the
  reconstruction gradients of the 3D golden case, a larger bulk-density
  reconstruction with checkpoints and points in the box (both accumulators),
  and a misfit gradient of two one-lane shots (one combined accumulator);
  the larger one also against the oracle; a float64 order-8 reconstruction
  with Dirichlet faces on one axis behind the layer, also against the oracle.
"""

from __future__ import annotations

import contextlib

import numpy as np
import pytest
import torch

from helpers import golden, host, product_sim, rel_l2, rel_max

pytestmark = pytest.mark.gpu

from oracle import driver as D  # noqa: E402

from cuwave_amd import _abi, boundary, scalar, utils, wave  # noqa: E402
from cuwave_amd import sensitivity as S  # noqa: E402
from cuwave_amd.engine import Propagator, sensor_points, source_points  # noqa: E402


@contextlib.contextmanager
def _threshold(blocks):
    previous = _abi.call_value("cw_set_stream_threshold", blocks)
    try:
        yield
    finally:
        _abi.call_value("cw_set_stream_threshold", previous)


def _both(run):
    """`run()` with the core marching every box, then with the general kernels alone."""
    with _threshold(0):
        cored = run()
    with _threshold(1 << 30):
        plain = run()
    return cored, plain


@pytest.mark.parametrize("name", ["fwd_pml_3d_o4_f32", "fwd_pml_3d_o8_f64"])
def test_the_3d_golden_forwards_with_a_core(name):
    z = golden(name)
    sim, osim = product_sim(z), D.sim_from_npz(z)
    print(f"case 3d core golden {name}", flush=True)
    src = wave.Source(z["src_pos"], z["src_sig"])
    ofield, otraces = D.simulate(osim, z["src_pos"], z["src_sig"], z["indicator"], z["sensors"])
    for field, traces in _both(lambda: wave.simulate(sim, src, z["indicator"], z["sensors"])):
        np.testing.assert_array_equal(host(field), ofield)
        np.testing.assert_array_equal(host(traces), otraces)


def _large(Nx, order, precision, pml, N):
    dx = tuple(1.0 / (n - 3) for n in Nx)
    dt = 0.4 * wave.stable_dt(dx, 1.0, order)
    common = dict(precision=precision, space_order=order, rho1=1.0, rho2=2.5, kappa1=1.0,
                  kappa2=4.0)
    sim = scalar.AcousticWave(Nx, dx, N, dt, pml=None if pml is None else boundary.Pml(pml, 2.0),
                              **common)
    osim = D.Sim("AcousticWave", Nx, dx, N, dt, precision, order,
                 materials=(1.0, 2.5, 1.0, 4.0),
                 pml=None if pml is None else D.Pml(pml, 2.0))
    rng = np.random.default_rng(order)
    ind = (0.2 + 0.6 * rng.random(sim.Nx_padded)).astype(sim.np_dtype)
    return sim, osim, ind, rng


@pytest.mark.parametrize("Nx,order,precision,pml", [((70, 66, 72), 8, "float32", 8),
                                                    ((45, 51, 43), 2, "float64", None)])
def test_a_larger_3d_forward_with_points_in_the_core(Nx, order, precision, pml):
    sim, osim, ind, rng = _large(Nx, order, precision, pml, 60)
    print(f"case 3d core forward {Nx} order {order} {precision}", flush=True)
    mid = [n // 2 for n in Nx]
    # a shot in the box, one in the shell; a plane of sensors across the box
    positions = [np.array([[m] for m in mid]), np.array([[3], [mid[1]], [4]])]
    t = np.arange(sim.N)
    sig = (np.sin(0.4 * t) * np.exp(-t / 30.0))[:, None].astype(sim.np_dtype)
    ys, xs = np.meshgrid(np.arange(mid[1] - 12, mid[1] + 12, 5),
                         np.arange(mid[2] - 12, mid[2] + 12, 5), indexing="ij")
    plane = np.stack([np.full(ys.size, mid[0] + 3), ys.ravel(), xs.ravel()])
    sensors = np.concatenate([plane, [[2, Nx[0] - 3], [mid[1], 5], [6, Nx[2] - 4]]], 1)
    for pos in positions:
        source = wave.Source(pos, sig)
        ofield, otraces = D.simulate(osim, pos, sig, ind, sensors)
        for field, traces in _both(lambda: wave.simulate(sim, source, ind, sensors)):
            np.testing.assert_array_equal(host(field), ofield)
            np.testing.assert_array_equal(host(traces), otraces)
    shots = [(pos, sig * (1 + k)) for k, pos in enumerate(positions)]
    oillum = D.illumination(osim, shots, ind)
    sources = [wave.Source(p, s) for p, s in shots]
    for illum in _both(lambda: wave.illumination(sim, sources, ind, batch=1)):
        assert rel_l2(illum, oillum) < (1e-5 if precision == "float32" else 1e-12)


def test_a_graphed_3d_march_with_a_core_is_one_plain_call():
    sim, _, ind, _ = _large((70, 66, 72), 8, "float32", 8, 77)
    print("case graphed 3d march with a core", flush=True)
    mat = sim.build_materials(ind)
    mid = [[n // 2] for n in sim.Nx]
    sig = np.sin(0.4 * np.arange(sim.N))[:, None].astype(sim.np_dtype)
    runs = []
    with _threshold(0):
        for window in (0, 24, 0):
            prop = Propagator(sim, mat, 1)
            src, signal = source_points(prop, mat, [wave.Source(np.array(mid), sig)])
            rec = sensor_points(prop, np.array([[30, 40], [33, 20], [36, 50]]))
            record = prop.record_memory(rec)
            U = prop.levels_memory(2)
            kw = dict(aux=prop.aux_memory(), sources=src, signal=signal, sensors=rec,
                      record=record)
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


def _same(cored, plain):
    """Two results of a gradient run, bit for bit: tensors, arrays, dicts of them, numbers."""
    if isinstance(cored, dict):
        assert cored.keys() == plain.keys()
        for k in cored:
            _same(cored[k], plain[k])
    elif isinstance(cored, (tuple, list)):
        for a, b in zip(cored, plain):
            _same(a, b)
    elif isinstance(cored, torch.Tensor):
        assert torch.equal(cored, plain)
    elif isinstance(cored, np.ndarray):
        np.testing.assert_array_equal(cored, plain)
    else:
        assert cored == plain


def test_the_3d_golden_adjoints_with_a_core():
    z = golden("adj_pml_3d_o4_f64")
    sim = product_sim(z)
    print("case 3d core adjoint golden adj_pml_3d_o4_f64", flush=True)
    src = wave.Source(z["src_pos"], z["src_sig"])
    objective = S.l2_misfit(z["observed"])
    for run in (S.sensitivity, S.reconstruction_sensitivity):
        cored, plain = _both(lambda: run(sim, src, z["indicator"], z["sensors"], objective))
        _same(cored[:3], plain[:3])


def test_a_larger_3d_bulk_density_reconstruction_with_a_core():
    n, width, N = 60, 8, 70
    dx = 1e-3
    dt = 0.4 * wave.stable_dt((dx,) * 3, 1800.0, 8)
    rng = np.random.default_rng(5)
    npad = wave.padded_shape((n, n, n))
    rho = 900.0 + 400.0 * rng.random(npad)
    kappa = rho * (1400.0 + 400.0 * rng.random(npad)) ** 2
    stack = np.stack([kappa, rho]).astype(np.float32)
    sim = scalar.BulkDensityWave((n, n, n), (dx,) * 3, N, dt, precision="float32",
                                 space_order=8, pml=boundary.Pml(width, 1800.0, 1e-6))
    osim = D.Sim("BulkDensityWave", (n, n, n), (dx,) * 3, N, dt, "float32", 8,
                 pml=D.Pml(width, 1800.0, 1e-6))
    t = np.arange(N) * dt
    from cuwave_amd.signals import ricker

    sig = (ricker(t, 1.0, 1e5) / dx**3)[:, None].astype(np.float32)
    src = np.array([[n // 2], [n // 2 - 3], [n // 2 + 2]])
    g = np.arange(width + 3, n - width - 3, 5)
    sensors = np.stack([np.full(len(g) ** 2, n // 2 + 4), np.repeat(g, len(g)), np.tile(g, len(g))])
    print("case 3d core bulk-density reconstruction checkpoints=2", flush=True)
    _, observed = D.simulate(osim, src, sig, stack, sensors)
    objective = S.l2_misfit((0.9 * observed).astype(np.float32))

    def run():
        return S.reconstruction_sensitivity(sim, wave.Source(src, sig), stack, sensors,
                                            objective, checkpoints=2)

    cored, plain = _both(run)
    _same(cored[:3], plain[:3])
    ocost, ograds, _, _ = D.reconstruction_sensitivity(
        osim, src, sig, stack, sensors, D.l2_misfit((0.9 * observed).astype(np.float32)),
        checkpoints=2)
    assert abs(cored[0] - ocost) <= 1e-4 * abs(ocost)
    for key in ("mass", "stiff"):
        assert rel_max(cored[1][key], ograds[key]) < 1e-4, key


def test_a_3d_misfit_gradient_of_one_lane_shots_with_a_core():
    sim, _, ind, _ = _large((64, 60, 62), 8, "float32", 8, 60)
    print("case 3d core misfit gradient", flush=True)
    t = np.arange(sim.N)
    sig = (np.sin(0.4 * t) * np.exp(-t / 30.0))[:, None].astype(np.float32)
    sources = [wave.Source(np.array([[32], [30 + k], [31]]), sig * (1 + k)) for k in range(2)]
    coords = np.stack([np.full(9, 0.5), np.repeat([0.35, 0.5, 0.65], 3),
                       np.tile([0.35, 0.5, 0.65], 3)], 1)  # (sensors, axes)
    sensors = utils.Sensors(sim, coords)
    observed = [np.zeros((sim.N, 9), dtype=np.float32) for _ in sources]

    def run():
        return utils.misfit_gradient(sim, sources, ind, sensors, observed, batch=1)

    cored, plain = _both(run)
    _same(cored, plain)


def test_a_float64_order_8_reconstruction_with_held_faces_and_a_core():
    # the float64 adjoint box at the widest stencil (its own register budget)
    # beside Dirichlet faces on one axis behind the layer
    Nx, N = (40, 38, 41), 50
    dx = tuple(1.0 / (n - 3) for n in Nx)
    dt = 0.4 * wave.stable_dt(dx, 1.0, 8)
    faces = ((boundary.Neumann, boundary.Neumann), (boundary.Dirichlet, boundary.Dirichlet),
             (boundary.Neumann, boundary.Neumann))
    sim = scalar.AcousticWave(Nx, dx, N, dt, precision="float64", space_order=8, rho1=1.0,
                              rho2=2.5, kappa1=1.0, kappa2=4.0, boundary=faces,
                              pml=boundary.Pml(6, 2.0))
    osim = D.Sim("AcousticWave", Nx, dx, N, dt, "float64", 8, materials=(1.0, 2.5, 1.0, 4.0),
                 boundary=((D.NEUMANN, D.NEUMANN), (D.DIRICHLET, D.DIRICHLET),
                           (D.NEUMANN, D.NEUMANN)),
                 pml=D.Pml(6, 2.0))
    rng = np.random.default_rng(9)
    ind = 0.2 + 0.6 * rng.random(sim.Nx_padded)
    src = np.array([[20], [19], [21]])
    t = np.arange(N)
    sig = (np.sin(0.5 * t) * np.exp(-t / 20.0))[:, None]
    sensors = np.array([[12, 25, 20], [15, 20, 26], [14, 22, 18]])
    print("case 3d core float64 order 8 reconstruction, dirichlet faces", flush=True)
    observed = 0.9 * D.simulate(osim, src, sig, ind, sensors)[1]

    def run():
        return S.reconstruction_sensitivity(sim, wave.Source(src, sig), ind, sensors,
                                            S.l2_misfit(observed))

    cored, plain = _both(run)
    _same(cored[:3], plain[:3])
    ocost, ograds, _, _ = D.reconstruction_sensitivity(osim, src, sig, ind, sensors,
                                                       D.l2_misfit(observed))
    assert abs(cored[0] - ocost) <= 1e-10 * abs(ocost)
    for key in ("mass", "stiff"):
        assert rel_max(cored[1][key], ograds[key]) < 1e-10, key
