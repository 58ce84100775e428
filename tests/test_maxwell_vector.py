"""The vector Maxwell oracle pinned, and the family's host logic (no GPU).

The oracle (`oracle/maxwell.py`) is checked against what fixes it
independently of its own formulas:

- the fdtdx traces the reference's tests compare with
  (tests/reference/fdtdx_vector3D_*.npz, cuwave tests/maxwell_test.py:
  912-978): a dipole in a conducting box, vacuum and a dielectric block, at
  round-off where the placement is shared;
- the exact symmetry of the mass-weighted operator at every order and wall
  (maxwell_test.py:416-443), which a wrong image parity or index breaks;
- the null space a discrete gradient spans (maxwell_test.py:445-463), which a
  curl sign or a swapped factor breaks where symmetry cannot see it;
- a cavity mode's convergence at the interior order up to the walls
  (maxwell_test.py:465-510), which a wrong image offset breaks;
- a conductor holding the tangential field at zero (maxwell_test.py:512-526).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from helpers import ROOT

from cuwave_amd import maxwell as MX
from cuwave_amd.boundary import Conductor, Magnetic, Traction
from cuwave_amd.stencils import staggered_weights
from oracle import maxwell as M

GOLDEN = ROOT / "tests" / "golden"


def fdtdx_traces(name: str, harmonic: bool = True):
    """The oracle run on fdtdx case `name` as the reference's test places it: fdtdx
    cell i is node i + 1, the ghost ring wraps the conductor faces, a z dipole fed
    the differenced current (maxwell_test.py:912-957)."""
    ref = np.load(GOLDEN / f"fdtdx_{name}.npz")
    shape = tuple(int(n) + 2 for n in ref["shape"])
    S = float(ref["courant"])
    eps = np.ones(shape)
    if ref["block_low"].size:
        low, size = ref["block_low"] + 1, ref["block_size"]
        eps[tuple(slice(lo, lo + n) for lo, n in zip(low, size))] = float(ref["permittivity"])
    conductor = (1 << 6) - 1
    if harmonic:
        minv = M.point_inverse_inertia(eps, conductor)
    else:  # fdtdx's staircase: a component takes the node below it
        minv = np.zeros((3, *shape))
        for c in range(3):
            minv[c] = 1.0 / (M.wall_weights(shape, (c,)) * eps)
        for face in range(6):
            wall = [slice(None)] * 3
            wall[face // 2] = 1 if face % 2 == 0 else shape[face // 2] - 2
            for c in range(3):
                if c != face // 2:
                    minv[(c, *wall)] = 0.0
    src = np.array(ref["source"]) + 1
    loads = np.diff(-S * ref["amplitude"], prepend=0.0)
    weight = np.array([S**2 * minv[(2, *src)]])
    signal = (loads / weight)[:, None]
    probes = ref["probes"].T + 1
    _, traces = M.simulate(minv, staggered_weights(1), [1.0] * 3, [S**2] * 3, 1.0, conductor,
                           len(loads), (np.array([2]), *[np.array([s]) for s in src]), weight,
                           signal, (np.full(probes.shape[1], 2), *probes))
    return traces, ref["traces"]


def gap(ours, theirs) -> float:
    return float(np.linalg.norm(ours - theirs) / np.linalg.norm(theirs))


def test_the_oracle_reproduces_fdtdx_at_round_off():
    assert gap(*fdtdx_traces("vector3D_vacuum")) < 1e-12
    assert gap(*fdtdx_traces("vector3D_block", harmonic=False)) < 1e-12


def test_the_harmonic_permittivity_closes_the_block_gap():
    assert gap(*fdtdx_traces("vector3D_block")) < 1e-3


def test_a_planted_force_sign_breaks_the_fdtdx_match(monkeypatch):
    # the force negated: the operator stays symmetric with the same null space, so
    # only the dynamics can see it (at order 2 inside a conductor the walls read
    # nothing that matters: every image lands on a tangential point it holds)
    real = M.force
    monkeypatch.setattr(M, "force", lambda *a: -real(*a))
    assert gap(*fdtdx_traces("vector3D_vacuum")) > 1e-2


# ------------------------------------------------------------ operator checks
def conductor_mask(shape, walls) -> int:
    mask = 0
    for d, pair in enumerate(walls):
        for side, c in enumerate(pair):
            if c == "c":
                mask |= 1 << (2 * d + side)
    return mask


WALLS = {"conductor": lambda n: (("c", "c"),) * n,
         "magnetic": lambda n: (("m", "m"),) * n,
         "mixed": lambda n: (("c", "m"),) + (("m", "m"),) * (n - 1)}


def operator(shape, order, conductor, rng, stencil="optimized"):
    """x -> (2 x - step(0, x)) M on the free unknowns, M the point mass
    (maxwell_test.py:365-385)."""
    ndim = len(shape)
    dx = tuple(1.0 / (n - 3) for n in shape)
    dt = 0.3 * min(dx)
    eps = 1.0 + 3.0 * rng.random(shape)
    minv = M.point_inverse_inertia(eps, conductor)
    mass = np.where(minv > 0, 1.0 / np.maximum(minv, 1e-300), 0.0)
    w = staggered_weights(order // 2, stencil)
    inv_dx = [1.0 / d for d in dx]
    f = [dt**2 / d for d in dx]
    free = np.zeros((ndim, *shape), dtype=bool)
    for c in range(ndim):
        free[(c, *(slice(1, n - 1 - (d == c)) for d, n in enumerate(shape)))] = True
    free &= minv > 0

    def apply(x):
        out = M.step(np.zeros_like(x), x, minv, w, inv_dx, f, 1.0, conductor)
        return np.where(free, (2.0 * x - out) * mass, 0.0)

    return apply, free, dx


@pytest.mark.parametrize("shape", [(22, 26), (14, 16, 18)])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("walls", sorted(WALLS))
def test_the_operator_is_symmetric(shape, order, walls):
    rng = np.random.default_rng(order)
    conductor = conductor_mask(shape, WALLS[walls](len(shape)))
    apply, free, _ = operator(shape, order, conductor, rng)
    x = np.where(free, rng.standard_normal(free.shape), 0.0)
    y = np.where(free, rng.standard_normal(free.shape), 0.0)
    ax, ay = apply(x), apply(y)
    scale = np.linalg.norm(ax) * np.linalg.norm(y)
    assert abs(np.sum(y * ax) - np.sum(x * ay)) / scale < 1e-12


def staggered_gradient(phi, order, dx):
    """The half-point gradient of phi at the field points (maxwell_test.py:388-409)."""
    R = order // 2
    w = staggered_weights(R, "optimized")  # the operator's own weights
    out = np.zeros((phi.ndim, *phi.shape))
    inner = tuple(slice(R, n - R) for n in phi.shape)
    for c in range(phi.ndim):
        for j, weight in enumerate(w, start=1):
            def window(offset):
                return tuple(slice(R + offset, n - R + offset) if d == c else slice(R, n - R)
                             for d, n in enumerate(phi.shape))
            out[c][inner] += weight / dx[c] * (phi[window(j)] - phi[window(-(j - 1))])
    return out


@pytest.mark.parametrize("shape", [(30, 34), (18, 20, 22)])
@pytest.mark.parametrize("order", [2, 4, 6])
def test_a_discrete_gradient_is_a_null_mode(shape, order):
    rng = np.random.default_rng(1)
    apply, _, dx = operator(shape, order, 0, rng)
    grad = staggered_gradient(rng.standard_normal(shape), order, dx)
    margin = order + 2
    inner = (slice(None), *(slice(margin, n - margin) for n in shape))
    assert np.linalg.norm(apply(grad)[inner]) < 1e-10 * np.linalg.norm(grad[inner])


def test_a_planted_curl_factor_breaks_the_null_mode(monkeypatch):
    # the curl's two axis factors swapped (the axes' spacings differ)
    real = M.curl
    monkeypatch.setattr(M, "curl",
                        lambda u, w, inv_dx, nu, c, *rest: real(u, w, inv_dx[::-1], nu, c, *rest))
    shape, order = (30, 34), 4
    rng = np.random.default_rng(1)
    apply, _, dx = operator(shape, order, 0, rng)
    grad = staggered_gradient(rng.standard_normal(shape), order, dx)
    inner = (slice(None), *(slice(order + 2, n - order - 2) for n in shape))
    assert np.linalg.norm(apply(grad)[inner]) > 1e-3 * np.linalg.norm(grad[inner])


def cavity_error(n, order, wall):
    """One step on a cavity mode the images keep an eigenvector (maxwell_test.py:465-501)."""
    shape = (n + 3, n + 3)
    dx, dt = 1.0 / n, 0.2 / n
    conductor = 0xF if wall == "c" else 0
    minv = M.point_inverse_inertia(np.ones(shape), conductor)
    kx, ky = 3 * np.pi, 2 * np.pi
    x = (np.arange(n + 3) - 1.0) / n
    tangent, normal = (np.cos, np.sin) if wall == "c" else (np.sin, np.cos)
    mode = np.stack([ky * tangent(kx * (x + 0.5 / n))[:, None] * normal(ky * x)[None, :],
                     -kx * normal(kx * x)[:, None] * tangent(ky * (x + 0.5 / n))[None, :]])
    free = np.zeros(mode.shape, dtype=bool)
    free[0, 1:n + 1, 1:n + 2] = free[1, 1:n + 2, 1:n + 1] = True
    free &= minv > 0
    mode = np.where(free, mode, 0.0)
    w = staggered_weights(order // 2, "taylor")
    out = M.step(np.zeros_like(mode), mode, minv, w, [1 / dx] * 2, [dt**2 / dx] * 2, 1.0,
                 conductor)
    applied = (2 * mode - out) / dt**2
    exact = (kx**2 + ky**2) * mode
    return float(np.where(free, np.abs(applied - exact), 0.0).max() / np.abs(exact).max())


@pytest.mark.parametrize("order", [4, 6, 8])
@pytest.mark.parametrize("wall", ["c", "m"])
def test_the_walls_keep_the_interior_order(order, wall):
    fine = cavity_error(64, order, wall)
    assert math.log2(cavity_error(32, order, wall) / fine) > order - 0.3


def test_a_conductor_holds_the_tangential_field_at_zero():
    rng = np.random.default_rng(2)
    shape = (24, 26)
    minv = M.point_inverse_inertia(1.0 + rng.random(shape), 0xF)
    w = staggered_weights(1)
    dt = 0.3 / 23
    t = np.arange(80) * dt
    sig = (1 - 2 * (np.pi * 8 * (t - 0.15)) ** 2) * np.exp(-(np.pi * 8 * (t - 0.15)) ** 2)
    field, _ = M.simulate(minv, w, [23.0, 25.0], [dt**2 * 23, dt**2 * 25], 1.0, 0xF, 80,
                          (np.array([1]), np.array([12]), np.array([13])), np.array([1.0]),
                          sig[:, None], (np.array([0]), np.array([5]), np.array([5])))
    assert np.abs(field).max() > 0
    for d in range(2):
        for index in (1, shape[d] - 2):
            wall = [slice(None)] * 2
            wall[d] = index
            assert np.abs(field[(1 - d, *wall)]).max() == 0.0


# ------------------------------------------------------------------ host side
def vector_sim(Nx, order=2, precision="float64", **kw):
    dx = tuple(1.0 / (n - 3) for n in Nx)
    return MX.DielectricWave(Nx, dx, 20, 0.2 * min(dx), precision=precision, space_order=order,
                             permittivity1=1.0, permittivity2=4.0, **kw)


@pytest.mark.parametrize("boundary", [None, Magnetic, ((Conductor, Magnetic), (Magnetic, Conductor),
                                                       (Conductor, Conductor))])
def test_the_point_inertia_is_the_oracles(boundary):
    sim = vector_sim((9, 11, 13), boundary=boundary)
    rng = np.random.default_rng(3)
    gamma = rng.random(sim.Nx_padded)
    mat = sim.build_materials(gamma)
    eps = sim.parametrization(gamma)[0][tuple(slice(0, n) for n in sim.Nx)]
    want = M.point_inverse_inertia(eps, sim.abi_grid(1).conductor)
    got = mat["minv"][(slice(None), *(slice(0, n) for n in sim.Nx))]
    assert np.array_equal(got, want)


def test_the_family_refuses_what_it_cannot_step():
    with pytest.raises(ValueError, match="no curl"):
        MX.DielectricWave((12,), (0.1,), 5, 0.01, permittivity1=1.0, permittivity2=2.0)
    with pytest.raises(ValueError):
        vector_sim((12, 14), boundary=Traction)
    with pytest.raises(ValueError, match="permittivity1"):
        MX.DielectricWave((12, 14), (0.1, 0.1), 5, 0.01)
    with pytest.raises(ValueError, match="positive"):
        vector_sim((12, 14), permeability=-1.0)


def test_indices_carry_the_component_row():
    sim = vector_sim((9, 11, 13))
    lin = sim.flatten_indices(np.array([[2, 0], [3, 1], [4, 2], [5, 3]]))
    assert lin.tolist() == [2 * sim.comp_stride + 3 * sim.strides[0] + 4 * sim.strides[1] + 5,
                            sim.strides[0] + 2 * sim.strides[1] + 3]
    with pytest.raises(ValueError, match="component"):
        sim.flatten_indices(np.array([[3], [3], [4], [5]]))


def test_the_drivers_refuse_what_the_family_does_not_take():
    from cuwave_amd import sensitivity as S
    from cuwave_amd import utils as U
    from cuwave_amd.wave import Source

    sim = vector_sim((9, 11, 13))
    source = Source(np.array([[0], [4], [5], [6]]), np.zeros((sim.N, 1)))
    sensors = np.array([[1], [4], [5], [7]])
    with pytest.raises(NotImplementedError, match="no strip"):
        S.reconstruction_sensitivity(sim, source, np.zeros(sim.Nx_padded), sensors, None,
                                     checkpoints=2)
    with pytest.raises(NotImplementedError, match="checkpoints"):
        U.misfit_gradient(sim, [source], np.zeros(sim.Nx_padded), sensors,
                          np.zeros((1, sim.N, 1)), checkpoints=2)
    with pytest.raises(NotImplementedError, match="scalar"):
        U.misfit_gradient(sim, [source], np.zeros(sim.Nx_padded), sensors,
                          np.zeros((1, sim.N, 1)), below=1.0)


def test_the_designed_permeability_is_the_oracles():
    from helpers import magnetodielectric_class

    cls = magnetodielectric_class()
    for Nx in ((10, 12), (9, 11, 13)):
        dx = tuple(1.0 / (n - 3) for n in Nx)
        sim = cls(Nx, dx, 10, 0.1 * min(dx), precision="float64", permittivity1=1.0,
                  permittivity2=4.0, boundary=Magnetic)
        gamma = np.random.default_rng(5).random(sim.Nx_padded)
        mat = sim.build_materials(gamma)
        nu = sim.parametrization(gamma)[1][tuple(slice(0, n) for n in Nx)]
        got = mat["nu_pair"][(slice(None), *(slice(0, n) for n in Nx))]
        assert np.array_equal(got, M.pair_inverse_permeability(nu))


def test_a_designed_permeability_keeps_the_operator_symmetric():
    rng = np.random.default_rng(9)
    shape = (12, 13, 14)
    conductor = conductor_mask(shape, WALLS["mixed"](3))
    eps = 1.0 + rng.random(shape)
    nu_pair = M.pair_inverse_permeability(0.5 + rng.random(shape))
    minv = M.point_inverse_inertia(eps, conductor)
    mass = np.where(minv > 0, 1.0 / np.maximum(minv, 1e-300), 0.0)
    w = staggered_weights(2, "optimized")
    dx = tuple(1.0 / (n - 3) for n in shape)
    free = np.zeros((3, *shape), dtype=bool)
    for c in range(3):
        free[(c, *(slice(1, n - 1 - (d == c)) for d, n in enumerate(shape)))] = True
    free &= minv > 0

    def apply(x):
        out = M.step(np.zeros_like(x), x, minv, w, [1 / d for d in dx], [1e-4 / d for d in dx],
                     1.0, conductor, nu_pair=nu_pair)
        return np.where(free, (2.0 * x - out) * mass, 0.0)

    x = np.where(free, rng.standard_normal(free.shape), 0.0)
    y = np.where(free, rng.standard_normal(free.shape), 0.0)
    ax, ay = apply(x), apply(y)
    assert abs(np.sum(y * ax) - np.sum(x * ay)) / (np.linalg.norm(ax) * np.linalg.norm(y)) < 1e-12


def test_the_damped_update_is_the_damped_particle():
    # magnetic walls keep a uniform field curl-free, so each point follows
    # x2 = (2 x1 - (1 - beta) x0) / (1 + beta), beta = minv d dt / 2
    # (cuwave tests/maxwell_test.py:340-345, 560-622)
    rng = np.random.default_rng(11)
    shape = (12, 13, 14)
    minv = M.point_inverse_inertia(1.0 + rng.random(shape), 0)
    damping = rng.random(shape)
    dt = 0.01
    a, b = rng.standard_normal(3), rng.standard_normal(3)
    u0 = np.broadcast_to(a[:, None, None, None], (3, *shape)).copy()
    u1 = np.broadcast_to(b[:, None, None, None], (3, *shape)).copy()
    u2 = M.step(u0, u1, minv, staggered_weights(2), [10.0] * 3, [1e-3] * 3, 1.0, 0,
                damping=damping, dt_=dt)
    for c in range(3):
        valid = (slice(1, -1),) * 3
        valid = tuple(slice(1, n - 2) if d == c else slice(1, n - 1) for d, n in enumerate(shape))
        beta = 0.5 * minv[c][valid] * damping[valid] * dt
        want = (2 * b[c] - (1 - beta) * a[c]) / (1 + beta)
        assert np.allclose(u2[c][valid], want, rtol=1e-14, atol=1e-14)


def test_the_reverse_march_refuses_a_damped_field():
    import dataclasses

    from cuwave_amd import boundary
    from cuwave_amd import sensitivity as S
    from cuwave_amd.wave import Source

    sim = vector_sim((12, 13))
    sim = dataclasses.replace(sim, damping=boundary.sponge(sim, np.zeros(sim.Nx_padded), 3, 0.1))
    source = Source(np.array([[0], [5], [6]]), np.zeros((sim.N, 1)))
    with pytest.raises(NotImplementedError, match="damped"):
        S.reconstruction_sensitivity(sim, source, np.zeros(sim.Nx_padded),
                                     np.array([[1], [5], [7]]), lambda u: (0.0, u))
