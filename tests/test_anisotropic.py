"""The cell-assembled elastic oracle and the family's host logic (no GPU).

The one-cell stencil K (`cuwave_amd.anisotropic.cell_stencil`) and the
oracle's step (`oracle/anisotropic.py`) are pinned by what fixes them
independently of their own formulas (cuwave tests/anisotropic_test.py:
126-300):

- K is symmetric and its null space is exactly the rigid-body motions;
- a rigid translation or rotation carries no force through the assembled step,
  walls included, and the checkerboard is not a null mode;
- a plane wave's acceleration far from the walls is the continuum's:
  -c^2 |k|^2 u with the Lame speeds for isotropic C, and with the Christoffel
  moduli C11 and C66 of an orthotropic C;
- the mass-weighted operator is exactly symmetric under a varying material,
  traction and clamped faces;
- a planted mistake (the corners mirrored along one axis) fails the plane-wave
  check.
"""

from __future__ import annotations

import numpy as np
import pytest

from cuwave_amd import anisotropic as AN
from cuwave_amd.boundary import Clamped, Traction
from oracle import anisotropic as O


def aniso_sim(Nx, dx=None, precision="float64", C=None, plane="strain", boundary=None, N=20):
    dx = dx or tuple(0.01 * (1.0 + 0.1 * d) for d in range(len(Nx)))
    return AN.AnisotropicElasticWave(Nx, dx, N, 1e-4, precision=precision, density=2.0,
                                     wavespeed_p=3.0, wavespeed_s=1.5, plane=plane, C=C,
                                     boundary=boundary)


def logical(sim, a):
    return a[(..., *(slice(0, n) for n in sim.Nx))]


def materials(sim, gamma):
    mat = sim.build_materials(gamma)
    return (logical(sim, mat["minv"]), logical(sim, mat["cell"]), mat["table"])


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_the_stencil_is_symmetric_with_the_rigid_null_space(ndim):
    sim = aniso_sim((8,) * ndim)
    K = AN.cell_stencil(ndim, sim.dx, sim.stiffness_matrix())
    assert np.allclose(K, K.T, atol=0, rtol=1e-13)
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-10 * np.abs(w).max()) == ndim * (ndim + 1) // 2


@pytest.mark.parametrize("ndim", [2, 3])
def test_rigid_motions_carry_no_force(ndim):
    Nx = (12, 13, 14)[:ndim]
    sim = aniso_sim(Nx)
    rng = np.random.default_rng(0)
    minv, cell, K = materials(sim, 0.5 + rng.random(sim.Nx_padded))
    x = np.meshgrid(*(np.arange(n) * d for n, d in zip(Nx, sim.dx)), indexing="ij")
    motions = [np.stack([np.full(Nx, 1.0 if c == k else 0.0) for c in range(ndim)])
               for k in range(ndim)]
    rot = np.zeros((ndim, *Nx))
    rot[0], rot[1] = -x[1], x[0]  # a rotation about the last axis (2D) or axis 2 (3D)
    motions.append(rot)
    for u in motions:
        F = O.force(u, cell, K)
        scale = np.abs(K).max() * np.abs(u).max() * np.abs(cell).max()
        assert np.abs(F).max() < 1e-12 * scale


def test_the_checkerboard_is_not_a_null_mode():
    sim = aniso_sim((12, 14))
    minv, cell, K = materials(sim, np.ones(sim.Nx_padded))
    a = np.add.outer(np.arange(12), np.arange(14))
    u = np.stack([(-1.0) ** a, (-1.0) ** (a + 1)])
    F = O.force(u, cell, K)
    assert np.abs(F[:, 3:-3, 3:-3]).min() > 1e-3 * np.abs(K).max()


def plane_wave_gap(sim, k, a, modulus):
    """||minv F + (modulus / rho0) |k|^2 u|| / ||...|| far from the walls, u = a cos(k.x)."""
    Nx = sim.Nx
    minv, cell, K = materials(sim, np.ones(sim.Nx_padded))
    x = np.meshgrid(*(np.arange(n) * d for n, d in zip(Nx, sim.dx)), indexing="ij")
    phase = sum(kd * xd for kd, xd in zip(k, x))
    u = np.stack([ad * np.cos(phase) for ad in a])
    accel = minv * O.force(u, cell, K)
    want = -(modulus / sim.density) * float(np.dot(k, k)) * u
    inner = (slice(None), *(slice(4, n - 4) for n in Nx))
    return np.linalg.norm(accel[inner] - want[inner]) / np.linalg.norm(want[inner])


def test_a_plane_wave_runs_at_the_lame_speeds():
    sim = aniso_sim((60, 64))
    k = np.array([3.0, 4.0]) * 0.05 / 0.01 / 5.0  # |k| dx = 0.05
    p, s = k / np.linalg.norm(k), np.array([-k[1], k[0]]) / np.linalg.norm(k)
    assert plane_wave_gap(sim, k, p, sim.lame + 2 * sim.shear) < 5e-3
    assert plane_wave_gap(sim, k, s, sim.shear) < 5e-3


def test_a_plane_wave_in_3d_runs_at_the_lame_speeds():
    sim = aniso_sim((30, 32, 34))
    k = np.array([2.0, 3.0, 6.0]) / 7.0 * 0.05 / 0.01
    p = k / np.linalg.norm(k)
    s = np.cross(p, [1.0, 0.0, 0.0])
    s /= np.linalg.norm(s)
    assert plane_wave_gap(sim, k, p, sim.lame + 2 * sim.shear) < 5e-3
    assert plane_wave_gap(sim, k, s, sim.shear) < 5e-3


def test_an_orthotropic_c_sets_the_christoffel_speeds():
    C = np.array([[30.0, 6.0, 0.0], [6.0, 12.0, 0.0], [0.0, 0.0, 4.0]])
    sim = aniso_sim((60, 64), C=C)
    k = np.array([0.05 / sim.dx[0], 0.0])
    assert plane_wave_gap(sim, k, [1.0, 0.0], C[0, 0]) < 5e-3
    assert plane_wave_gap(sim, k, [0.0, 1.0], C[2, 2]) < 5e-3


def test_a_planted_corner_mirror_fails_the_plane_wave(monkeypatch):
    real = AN.AnisotropicElasticWave.stencil_table

    def mirrored_corners(self):
        # the corners read with axis 0 reversed (reversing every axis is a point
        # reflection, which elasticity does not see)
        K = real(self)
        n, nd = 2**self.ndim, self.ndim
        order = np.concatenate([np.arange(nd) + (c ^ 1) * nd for c in range(n)])
        return K[np.ix_(order, order)]

    monkeypatch.setattr(AN.AnisotropicElasticWave, "stencil_table", mirrored_corners)
    sim = aniso_sim((60, 64))
    k = np.array([3.0, 4.0]) * 0.05 / 0.01 / 5.0
    assert plane_wave_gap(sim, k, k / np.linalg.norm(k), sim.lame + 2 * sim.shear) > 0.1


@pytest.mark.parametrize("Nx", [(14, 16), (9, 10, 11)])
@pytest.mark.parametrize("boundary", ["traction", "clamped", "mixed"])
def test_the_operator_is_symmetric(Nx, boundary):
    faces = {"traction": None, "clamped": Clamped,
             "mixed": ((Traction, Clamped),) + ((Clamped, Traction),) * (len(Nx) - 1)}[boundary]
    sim = aniso_sim(Nx, boundary=faces)
    rng = np.random.default_rng(len(Nx))
    minv, cell, K = materials(sim, 0.5 + rng.random(sim.Nx_padded))
    free = np.broadcast_to(O.interior(Nx) & (minv > 0), (len(Nx), *Nx))
    mass = np.where(minv > 0, 1.0 / np.maximum(minv, 1e-300), 0.0)

    def apply(x):
        out = O.step(np.zeros_like(x), x, minv, cell, K, 1.0)
        return np.where(free, (2.0 * x - out) * mass, 0.0)

    x = np.where(free, rng.standard_normal(free.shape), 0.0)
    y = np.where(free, rng.standard_normal(free.shape), 0.0)
    ax, ay = apply(x), apply(y)
    assert abs(np.sum(y * ax) - np.sum(x * ay)) / (np.linalg.norm(ax) * np.linalg.norm(y)) < 1e-13


# ------------------------------------------------------------------ host side
@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_the_materials_are_the_oracles(precision):
    sim = aniso_sim((9, 11, 13), precision=precision,
                    boundary=((Clamped, Traction), (Traction, Traction), (Traction, Clamped)))
    gamma = (0.5 + np.random.default_rng(1).random(sim.Nx_padded)).astype(sim.np_dtype)
    mat = sim.build_materials(gamma)
    g = logical(sim, gamma)
    clamped = (1 << 0) | (1 << 5)
    assert np.array_equal(logical(sim, mat["minv"]),
                          O.inverse_inertia(g, sim.density, float(np.prod(sim.dx)), clamped))
    assert np.array_equal(logical(sim, mat["cell"])[:-1, :-1, :-1],
                          O.cell_average(g)[:-1, :-1, :-1])


def test_the_family_refuses_what_it_cannot_step():
    with pytest.raises(ValueError, match="order"):
        AN.AnisotropicElasticWave((10, 12), (0.1, 0.1), 5, 1e-3, space_order=4, density=1.0,
                                  wavespeed_p=2.0, wavespeed_s=1.0)
    with pytest.raises(ValueError, match="plane stress"):
        aniso_sim((8, 9, 10), plane="stress")
    with pytest.raises(ValueError, match="symmetric"):
        aniso_sim((10, 12), C=np.array([[1.0, 2.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]))
    with pytest.raises(ValueError, match="below"):
        AN.AnisotropicElasticWave((10, 12), (0.1, 0.1), 5, 1e-3, density=1.0, wavespeed_p=1.0,
                                  wavespeed_s=1.0)


def test_the_damped_update_is_the_damped_particle():
    # a rigid translation carries no force, so with damping each node follows
    # x2 = (2 x1 - (1 - beta) x0) / (1 + beta), beta = minv d dt / 2
    sim = aniso_sim((12, 13, 14))
    rng = np.random.default_rng(12)
    minv, cell, K = materials(sim, 0.5 + rng.random(sim.Nx_padded))
    damping = rng.random(sim.Nx)
    dt = 1e-3
    a, b = rng.standard_normal(3), rng.standard_normal(3)
    u0 = np.broadcast_to(a[:, None, None, None], (3, *sim.Nx)).copy()
    u1 = np.broadcast_to(b[:, None, None, None], (3, *sim.Nx)).copy()
    u2 = O.step(u0, u1, minv, cell, K, 1e-6, damping, dt)
    inner = (slice(1, -1),) * 3
    beta = 0.5 * minv[inner] * damping[inner] * dt
    for c in range(3):
        want = (2 * b[c] - (1 - beta) * a[c]) / (1 + beta)
        assert np.allclose(u2[c][inner], want, rtol=1e-12, atol=1e-12)


def test_the_reverse_march_refuses_a_damped_field():
    import dataclasses

    from cuwave_amd import boundary
    from cuwave_amd import sensitivity as S
    from cuwave_amd.wave import Source

    sim = aniso_sim((12, 13))
    sim = dataclasses.replace(sim, damping=boundary.sponge(sim, np.ones(sim.Nx_padded), 3, 0.1))
    source = Source(np.array([[0], [5], [6]]), np.zeros((sim.N, 1)))
    with pytest.raises(NotImplementedError, match="damped"):
        S.reconstruction_sensitivity(sim, source, np.ones(sim.Nx_padded),
                                     np.array([[1], [5], [7]]), lambda u: (0.0, u))
